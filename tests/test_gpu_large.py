"""Large forward + VJP through the batched entry (speedy_fwd_batch: a level-set
pass over every column, then the fused forward + VJP, then the per-member
parameter-gradient reduction), one request on the continuous-batching server:

  (a) T425, 1280 x 640 = 819200 columns, 8 levels, one field, land recipe
      (seed 1), parameters drawn from the adjoint ranges, cotangents N(0, 1);
  (b) 64 members at T31 with 16 levels (294912 columns), each member's T
      perturbed by N(0, 0.5 K), each with its own level set.

Each case is served twice and the two results must be bit-identical.
Checks, per member:
  * the level set equals the float32 oracle's over the whole grid (at most two
    of the 64 members may differ, through a float32 tie in one column); every
    member is checked, a member whose level set differs with the GPU's level
    set injected into the oracle, and the tie columns behind the difference
    are named from the GPU's own per-column cloud tops;
  * tendencies meet the forward bar of helpers.parity (float32 oracle on the
    whole grid, float64 oracle on a 4096-column subsample with the level set
    injected, helpers.subgrid);
  * state and parameter gradients against the float64 host adjoint on the whole
    grid (pinned to central differences of the float64 oracle in
    test_adjoint_host): the VJP bar of test_gpu_vjp per column, on all but
    0.1 % of the columns (_check_gradients);
  * <g, d> for a random direction d of the subsample's state equals the central
    difference of the float64 oracle on that subsample: relative error <= 1e-4,
    or <= 3x the float32 host adjoint's own error on the same direction.
"""
import numpy as np
import pytest

import helpers as H
from jcm_amd import synthetic
from jcm_amd.physics_interface import PhysicsTendency
from oracle import speedy_oracle as so
from test_gpu_vjp import GRAD_TOL

pytestmark = pytest.mark.gpu
FIELDS = ("u_wind", "v_wind", "temperature", "specific_humidity")


def _serve(phys, case, cot):
    """One forward + VJP request through the server's batched launch, served twice:
    the two results must be bit-identical (each column's arithmetic is fixed, so
    any difference would be a race)."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    st, fo, ge, da = case
    dev = phys.device
    st_d = st.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), device=dev) for k in H.STATE_FIELDS})
    cot_d = PhysicsTendency(*[torch.as_tensor(np.ascontiguousarray(c), device=dev) for c in cot])
    server = PhysicsServer(phys, n_streams=1, record_level_sets=True)
    reqs = [Request(state=st_d, forcing=fo, geometry=ge, date=da, cotangent=cot_d, rid=i) for i in range(2)]
    for r in reqs:
        assert server.run([r], realtime=False)[-1] is r
    torch.cuda.synchronize()
    assert server.n_batches == 2
    req, again = reqs
    assert req.level_set == again.level_set
    for f in FIELDS:
        assert torch.equal(getattr(req.tendency, f), getattr(again.tendency, f)), f
        assert torch.equal(getattr(req.state_grads, f), getattr(again.state_grads, f)), f
    assert torch.equal(req.state_grads.normalized_surface_pressure, again.state_grads.normalized_surface_pressure)
    for k in req.param_grads:
        assert torch.equal(req.param_grads[k], again.param_grads[k]), k
    out = lambda x: x.detach().cpu().numpy()
    tend = {f: out(getattr(req.tendency, f)) for f in FIELDS}
    sg = req.state_grads
    g = dict(u=out(sg.u_wind), v=out(sg.v_wind), t=out(sg.temperature), q=out(sg.specific_humidity),
             ps=out(sg.normalized_surface_pressure))
    from jcm_amd.params import VJP_PARAMS
    prm = np.stack([out(req.param_grads[f"{grp}.{nm}"]) for grp, nm in VJP_PARAMS], axis=-1)
    print(f"served {req.columns} columns in {req.latency * 1e3:.2f} ms (host to result-ready)", flush=True)
    return tend, g, prm, req.level_set


def _subsample(ix, il, n, seed):
    """A 4096-column subsample as a grid: random longitudes x random latitudes."""
    rng = np.random.default_rng(seed)
    ni = min(ix, 64)
    nj = min(il, n // ni)
    return np.sort(rng.choice(ix, ni, replace=False)), np.sort(rng.choice(il, nj, replace=False))


def _column_score(g, g64):
    """Per column: the largest error of any gradient field, relative to that field's
    largest float64 value."""
    sc = 0.0
    for k in ("t", "q", "u", "v", "ps"):
        e = np.abs(np.asarray(g[k], np.float64) - g64[k]) / max(np.abs(g64[k]).max(), 1e-30)
        sc = np.maximum(sc, e.max(axis=0) if e.ndim == 3 else e)
    return sc


def _check_gradients(label, g, prm, g64, g32, tight_params=None):
    """Gradients over a whole large grid.  A float32 evaluation is as far from
    float64 as its most ill-conditioned column lets it be (cancellation in the
    convective flux differences), and over 10^5-10^6 columns a few such columns
    always occur, where two float32 evaluations (with and without fused
    multiply-adds) land anywhere within that spread.  So the bar is per column:
      * at most 0.1 % of the columns (or 2) have an error above 1e-4 of the
        field's largest float64 value (the VJP bar of test_gpu_vjp);
      * columns above 1e-2 sit on a discrete-decision tie (a convection or
        cloud-top index that float32 and float64 round to different sides, where
        the derivative jumps); they are as rare as the float32 host adjoint's
        own (<= 2 + 1e-5 of the columns);
      * each parameter gradient within 1e-4 of the float64 sum, or within 3x the
        float32 host adjoint's own error (which carries the tie columns).  With
        tight_params (a list), the rigorous float32 bound of a sum over columns
        (the sum of the float32 host's per-column errors) is asserted instead and
        whether the tight bar held is appended, for a bar over many members."""
    s_gpu, s32 = _column_score(g, g64), _column_score(g32, g64)
    ncol = s_gpu.size
    ties_gpu, ties32 = int((s_gpu > 1e-2).sum()), int((s32 > 1e-2).sum())
    n_gpu, n32 = int((s_gpu > GRAD_TOL).sum()), int((s32 > GRAD_TOL).sum())
    q = lambda x: float(np.quantile(x, 0.999))
    print(f"{label}: columns above 1e-4: gpu {n_gpu}, float32 host {n32}; above 1e-2 (ties): gpu {ties_gpu}, "
          f"float32 host {ties32}; of {ncol}; 99.9th percentile gpu {q(s_gpu):.1e} float32 host {q(s32):.1e}",
          flush=True)
    for k in ("t", "q", "u", "v", "ps"):
        print(f"{label} grad {k}: rel err vs float64 adjoint {H.rel_inf(g[k], g64[k]):.2e} "
              f"(float32 host {H.rel_inf(g32[k], g64[k]):.2e})", flush=True)
    assert n_gpu <= max(2, 1e-3 * ncol), f"{label}: {n_gpu} columns above {GRAD_TOL}"
    assert ties_gpu <= max(2, 2 * ties32) and ties_gpu <= 2 + 1e-5 * ncol, f"{label}: {ties_gpu} tie columns"
    from jcm_amd.params import VJP_PARAMS
    for i, (grp, nm) in enumerate(VJP_PARAMS):
        ref = float(g64["prm"][..., i].sum())
        e = abs(float(prm[i]) - ref) / max(abs(ref), 1e-30)
        e32 = abs(float(g32["prm"][..., i].astype(np.float64).sum()) - ref) / max(abs(ref), 1e-30)
        l1 = float(np.abs(g32["prm"][..., i].astype(np.float64) - g64["prm"][..., i]).sum()) / max(abs(ref), 1e-30)
        print(f"{label} d/d {nm}: gpu {float(prm[i]):.6e} float64 {ref:.6e} rel {e:.1e} (float32 host {e32:.1e}, "
              f"rigorous float32 bound {l1:.1e})", flush=True)
        tight = e <= max(GRAD_TOL, 3 * e32) or ref == 0.0 and prm[i] == 0.0
        if tight_params is None:
            assert tight, f"{label} {nm}: {e:.3e}"
        else:
            assert tight or e <= l1, f"{label} {nm}: {e:.3e} above the float32 bound {l1:.3e}"
            tight_params.append(tight)


def _check_member(label, case, params, cot, tend, g, prm, S, member=None, inject=False, directional=True, ties=(),
                  tight_params=None):
    """inject: the float32 oracle runs with the GPU's level set S (a member whose
    free-running level set differs from the oracle's through a float32 tie, see
    _name_tie_columns); otherwise the oracle's own level set must equal S.
    ties: the named tie columns (lon, lat), whose discrete decisions differ by
    construction; they are left out of the tendency comparison (the gradient
    check counts them as tie columns)."""
    kx, ix, il = case[2].nodal_shape
    pick = (lambda a: a[member]) if member is not None else (lambda a: a)
    keep = np.ones((ix, il), bool)
    for i, j in ties:
        keep[i, j] = False
    drop = lambda a: np.where(keep, a, 0.0)
    # level set and tendencies against the float32 oracle on the whole grid
    ref, _, S_ref = H.run_oracle(*case, params=params, member=member, level_set=S if inject else None)
    assert S == S_ref, f"{label}: level set {bin(S)} != {bin(S_ref)}"
    I, J = _subsample(ix, il, 4096, 7 if member is None else 7 + member)
    sub = H.subgrid(case, I, J, member)
    ref64, _, _ = H.run_oracle(*sub, params=params, dtype=np.float64, level_set=S)
    cut = lambda a: np.asarray(a)[..., I, :][..., J]
    for f in FIELDS:
        gf = drop(pick(tend[f]))
        e_full = H.rel_inf(gf, drop(getattr(ref, f)))
        ok, e32, e64, eref = H.parity(cut(gf), cut(drop(getattr(ref, f))), np.where(cut(keep), getattr(ref64, f), 0.0))
        print(f"{label} {f}: whole grid vs fp32 oracle {e_full:.2e}; subsample vs fp32 {e32:.2e}, "
              f"vs fp64 {e64:.2e} (fp32 oracle vs fp64 {eref:.2e})", flush=True)
        assert e_full <= H.REL_TOL or ok, f"{label} {f}: {e_full:.3e} / {e64:.3e}"
    # gradients against the float64 (and float32) host adjoint on the whole grid
    one = H.subgrid(case, np.arange(ix), np.arange(il), member) if member is not None else case
    cm = [pick(c) for c in cot]
    _, g64 = H.run_host_vjp(one, params, cm, level_set=S)
    _, g32 = H.run_host_vjp(one, params, cm, level_set=S, dtype=np.float32)
    _check_gradients(label, {k: pick(v) for k, v in g.items()}, pick(prm), g64, g32, tight_params)
    if not directional:
        return
    # directional derivative on the subsample against the float64 oracle
    st, fo, ge, da = sub
    rng = np.random.default_rng(11)
    shp = st.temperature.shape
    d = dict(t=rng.standard_normal(shp), q=rng.standard_normal(shp) * 1e-2, u=rng.standard_normal(shp),
             v=rng.standard_normal(shp), ps=rng.standard_normal(shp[1:]) * 1e-3)
    lin = sum(float(np.sum(cut(pick(g[k])).astype(np.float64) * d[k])) for k in d)
    cts = [cut(c) for c in cm]
    J_of = lambda t: float(sum(np.sum(c.astype(np.float64) * x) for c, x in
                               zip(cts, (t.u_wind, t.v_wind, t.temperature, t.specific_humidity))))
    h = 1e-6

    def shifted(sgn):
        s = st.copy(temperature=st.temperature.astype(np.float64) + sgn * h * d["t"],
                    specific_humidity=st.specific_humidity.astype(np.float64) + sgn * h * d["q"],
                    u_wind=st.u_wind.astype(np.float64) + sgn * h * d["u"],
                    v_wind=st.v_wind.astype(np.float64) + sgn * h * d["v"],
                    normalized_surface_pressure=st.normalized_surface_pressure.astype(np.float64)
                    + sgn * h * d["ps"])
        t, _, _ = so.compute_tendencies(*H.oracle64_same_constants(s, fo, ge, da, params), level_set=S,
                                        return_level_set=True)
        return t

    fd = (J_of(shifted(1)) - J_of(shifted(-1))) / (2 * h)
    lin32 = sum(float(np.sum(cut(g32[k]).astype(np.float64) * d[k])) for k in d)
    e, e32 = abs(lin - fd) / abs(fd), abs(lin32 - fd) / abs(fd)
    print(f"{label} directional derivative on the subsample: gpu {lin:.8e} float64 oracle FD {fd:.8e} "
          f"(rel {e:.1e}; float32 host adjoint {e32:.1e})", flush=True)
    assert e <= max(GRAD_TOL, 3 * e32), f"{label}: {lin} vs {fd}"


def test_t425_forward_vjp_single_field():
    from jcm_amd.speedy_physics import SpeedyPhysics
    params = synthetic.draw_vjp_parameters(2)
    phys = SpeedyPhysics(parameters=params)
    case = synthetic.make_case(1280, 640, 8, seed=1, land=True)
    rng = np.random.default_rng(2)
    cot = [rng.standard_normal((8, 1280, 640)).astype(np.float32) for _ in range(4)]
    tend, g, prm, S = _serve(phys, case, cot)
    assert len(S) == 1
    _check_member("T425", case, params, cot, tend, g, prm, S[0])


def test_t425_forward_only():
    """T425 forward through the batched entry (the forward-only batch kernel), served
    twice (bit-identical): level set and tendencies against the float32 oracle on the
    whole grid and the float64 oracle on a 4096-column subsample."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    case = synthetic.make_case(1280, 640, 8, seed=1, land=True)
    st, fo, ge, da = case
    st_d = st.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), device=phys.device)
                      for k in H.STATE_FIELDS})
    server = PhysicsServer(phys, n_streams=1, record_level_sets=True)
    reqs = [Request(state=st_d, forcing=fo, geometry=ge, date=da, rid=i) for i in range(2)]
    for r in reqs:
        server.run([r], realtime=False)
    torch.cuda.synchronize()
    for f in FIELDS:
        assert torch.equal(getattr(reqs[0].tendency, f), getattr(reqs[1].tendency, f)), f
    ref, rdata, S_ref = H.run_oracle(*case)
    assert reqs[0].level_set == [S_ref], (reqs[0].level_set, bin(S_ref))
    I, J = _subsample(1280, 640, 4096, 5)
    ref64, _, _ = H.run_oracle(*H.subgrid(case, I, J), dtype=np.float64, level_set=S_ref)
    cut = lambda a: np.asarray(a)[..., I, :][..., J]
    off = {}
    for f in FIELDS:
        g = getattr(reqs[0].tendency, f).cpu().numpy()
        e_full = H.rel_inf(g, getattr(ref, f))
        ok, e32, e64, eref = H.parity(cut(g), cut(getattr(ref, f)), getattr(ref64, f))
        col_err = np.abs(g - getattr(ref, f)).max(axis=0) / max(np.abs(getattr(ref, f)).max(), 1e-30)
        n_off = int((col_err > H.REL_TOL).sum())
        print(f"T425 forward {f}: whole grid {e_full:.2e} ({n_off} columns above 1e-5); subsample vs fp32 {e32:.2e}, "
              f"vs fp64 {e64:.2e} (fp32 oracle vs fp64 {eref:.2e})", flush=True)
        assert ok and n_off <= 2 + 1e-5 * g[0].size, f
        for i, j in zip(*np.nonzero(col_err > H.REL_TOL)):
            off.setdefault((int(i), int(j)), []).append(f)
    # every column above the bar is explained, by name
    _explain_columns(case, reqs[0].tendency, ref, rdata, S_ref, off)


def _convection_margin(data64, sub, rhbl):
    """Smallest relative distance, in float64, of the comparisons that decide a
    column's convection (diagnose_convection, convection.py:15-97: the saturation
    static energies against the boundary-layer energies, and the boundary-layer
    humidities against rhbl * qsat).  Below a few float32 ulps the decision is a
    float32 tie: any float32 evaluation order may take either side."""
    st, _, ge, _ = sub
    se = np.asarray(data64.convection.se, np.float64)[:, 0, 0]
    qsat = np.asarray(data64.humidity.qsat, np.float64)[:, 0, 0]
    qa = np.asarray(st.specific_humidity, np.float64)[:, 0, 0]
    wvi = np.asarray(ge.wvi, np.float64)
    kx = se.size
    mss = se + so.alhc * qsat
    mse0 = se[-1] + so.alhc * qa[-1]
    mse1 = min(mse0, se[-2] + so.alhc * qa[-2])
    mss0 = max(mse0, mss[-1])
    mss2 = mss[:-1] + wvi[:-1, 1] * np.diff(mss)
    lv = range(2, kx - 3)
    return min([abs(mss0 - mss2[k]) / abs(mss0) for k in lv] + [abs(mse1 - mss2[k]) / abs(mse1) for k in lv]
               + [abs(qa[-1] - rhbl * qsat[-1]) / (rhbl * qsat[-1]), abs(qa[-2] - rhbl * qsat[-2]) / (rhbl * qsat[-2])])


def _explain_columns(case, tend, ref, rdata, S, off):
    """Each column whose tendencies miss the whole-grid bar must be one of:
    a column whose cloud top or convective top (after condensation) differs
    between the GPU and the float32 oracle; a column whose convection decision
    lies within 4e-7 (a few float32 ulps) of its threshold in float64; or a
    column where the GPU is as close to the float64 oracle as the float32 oracle
    is (x3, per column).  Each is named with the evidence."""
    import torch
    from jcm_amd import Parameters
    from jcm_amd.speedy_physics import SpeedyPhysics
    if not off:
        return
    st, fo, ge, da = case
    dphys = SpeedyPhysics(with_diagnostics=True)
    _, gdata = dphys.compute_tendencies(*case)
    torch.cuda.synchronize()
    ip_g = gdata.convection.iptop.cpu().numpy()
    icl_g = gdata.shortwave_rad.icltop.cpu().numpy()
    ip_r, icl_r = np.asarray(rdata.convection.iptop), np.asarray(rdata.shortwave_rad.icltop)
    rhbl = float(Parameters.default().convection.rhbl)
    for (i, j), fields in sorted(off.items()):
        if ip_g[i, j] != ip_r[i, j] or icl_g[i, j] != icl_r[i, j]:
            print(f"T425 tie column (lon {i}, lat {j}) {fields}: cloud top GPU {int(icl_g[i, j])} / oracle "
                  f"{int(icl_r[i, j])}, top after condensation GPU {int(ip_g[i, j])} / oracle {int(ip_r[i, j])}",
                  flush=True)
            continue
        sub = H.subgrid(case, np.array([i]), np.array([j]))
        r64, d64, _ = H.run_oracle(*sub, level_set=S, dtype=np.float64)
        margin = _convection_margin(d64, sub, rhbl)
        if margin < 4e-7:
            print(f"T425 tie column (lon {i}, lat {j}) {fields}: convection decided {margin:.1e} from its threshold "
                  "in float64", flush=True)
            continue
        for f in fields:
            scale = max(np.abs(getattr(ref, f)).max(), 1e-30)
            g = getattr(tend, f)[:, i, j].cpu().numpy().astype(np.float64)
            r32, r6 = np.asarray(getattr(ref, f))[:, i, j].astype(np.float64), np.ravel(getattr(r64, f))
            e_g, e_32 = np.abs(g - r6).max() / scale, np.abs(r32 - r6).max() / scale
            print(f"T425 column (lon {i}, lat {j}) {f}: GPU vs float64 {e_g:.1e}, float32 oracle vs float64 {e_32:.1e}",
                  flush=True)
            assert e_g <= max(H.REL_TOL, 3 * e_32), f"column ({i}, {j}) {f} is not explained"


def test_64_members_t31_16_levels():
    from jcm_amd.speedy_physics import SpeedyPhysics
    params = synthetic.draw_vjp_parameters(2)
    phys = SpeedyPhysics(parameters=params)
    st, fo, ge, da = synthetic.make_case(96, 48, 16, seed=1, land=True, nmember=64, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[1::2] *= np.float32(1.12)   # odd members condense widely: member-specific level sets
    case = (st.copy(specific_humidity=q), fo, ge, da)
    rng = np.random.default_rng(3)
    cot = [rng.standard_normal((64, 16, 96, 48)).astype(np.float32) for _ in range(4)]
    tend, g, prm, S = _serve(phys, case, cot)
    assert len(S) == 64 and len(set(S)) >= 2
    # every member's level set against the float32 oracle's; a member may differ
    # only through a float32 tie in one column's convection or condensation top
    # (the reference's float32 arithmetic and the kernel's round differently
    # there; about one column in 3e5), at most two of 64.  Such a member is
    # compared with the GPU's level set injected into the oracle, and its tie
    # columns are named.
    S_ref = [H.run_oracle(*case, params=params, member=m)[2] for m in range(64)]
    differ = [m for m in range(64) if S[m] != S_ref[m]]
    print(f"level sets differing from the float32 oracle: {[(m, bin(S[m]), bin(S_ref[m])) for m in differ]}",
          flush=True)
    assert len(differ) <= 2
    ties = {m: _name_tie_columns(case, params, m, S[m], S_ref[m]) for m in differ}
    # parameter gradients: each within the rigorous float32 bound of its sum over
    # columns; at least 90 % of the (member, parameter) pairs within the tight bar
    tight: list = []
    for m in range(64):
        _check_member(f"member {m}", case, params, cot, tend, g, prm, S[m], member=m, inject=m in differ,
                      directional=m in (0, 1, 2, 3, 31, 62, 63), ties=ties.get(m, ()), tight_params=tight)
    print(f"parameter gradients within the tight bar: {sum(tight)} of {len(tight)}", flush=True)
    assert sum(tight) >= 0.9 * len(tight)


def _name_tie_columns(case, params, m, S_gpu, S_ref):
    """The columns behind a level-set difference of member m: columns whose clipped
    cloud top (shortwave_radiation.py:63-70) on the GPU (member m alone, with
    diagnostics, which must reproduce the batch's level set) differs from the
    float32 oracle's, at a level in S_gpu ^ S_ref.  The float64 oracle's and the
    float32 host build's cloud tops at each named column show which side of the
    tie each evaluation takes."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    kx, ix, il = case[2].nodal_shape
    one = H.subgrid(case, np.arange(ix), np.arange(il), m)
    phys = SpeedyPhysics(parameters=params, with_diagnostics=True)
    _, data = phys.compute_tendencies(*one)
    torch.cuda.synchronize()
    assert phys.level_sets(ix, il, 1)[0] == S_gpu, f"member {m}: alone {bin(phys.level_sets(ix, il, 1)[0])}"
    clip = lambda a: np.clip(np.asarray(a).astype(np.int64), 0, kx - 1)
    c_gpu = clip(data.shortwave_rad.icltop.cpu().numpy())
    ip_gpu = data.convection.iptop.cpu().numpy().astype(np.int64)
    _, diag, S_host = H.run_host_column(one, params)
    _, rdata, _ = H.run_oracle(*one, params=params)
    _, rdata64, _ = H.run_oracle(*one, params=params, dtype=np.float64)
    c_host, c_ref, c64 = clip(diag["icltop"]), clip(rdata.shortwave_rad.icltop), clip(rdata64.shortwave_rad.icltop)
    bits = S_gpu ^ S_ref
    cols = [(int(i), int(j)) for i, j in zip(*np.nonzero(c_gpu != c_ref))
            if (bits >> int(c_gpu[i, j])) & 1 or (bits >> int(c_ref[i, j])) & 1]
    for i, j in cols:
        print(f"member {m}: tie column (lon {i}, lat {j}): cloud top GPU {c_gpu[i, j]}, float32 oracle "
              f"{c_ref[i, j]}, float64 oracle {c64[i, j]}, float32 host build {c_host[i, j]}; top after "
              f"condensation GPU {ip_gpu[i, j]}, float32 oracle {int(rdata.convection.iptop[i, j])}", flush=True)
    assert 1 <= len(cols) <= 2, f"member {m}: {len(cols)} columns behind the level-set difference"
    return cols
