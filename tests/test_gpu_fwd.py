"""GPU parity of the fused forward kernel (through the C ABI) against the oracle.

Bar: per tendency field ||F_gpu - F_ref||_inf / ||F_ref||_inf <= 1e-5 against the
float32 oracle on identical inputs (2e-5 at 16 levels, see test_host_column);
discrete decisions (iptop, icltop, level set S) identical.  The float64 oracle
error is printed per field.
"""
import os
import sys

import numpy as np
import pytest

import helpers as H
from jcm_amd import Geometry, Parameters, synthetic
from jcm_amd.forcing import ForcingData
from jcm_amd.physics_interface import PhysicsState
from oracle import speedy_oracle as so

pytestmark = pytest.mark.gpu
FIELDS = ("u_wind", "v_wind", "temperature", "specific_humidity")
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def phys():
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    assert torch.cuda.is_available()
    return SpeedyPhysics(with_diagnostics=True)


def _np(x):
    return x.detach().cpu().numpy()


def check_fwd(phys, case, params=None, tol=H.REL_TOL, label="", strict=True):
    """strict: every tendency field within `tol` of the float32 oracle; otherwise
    the rule of helpers.parity (float32 oracle or float64 oracle)."""
    import torch
    phys.parameters = params or Parameters.default()
    tend, data = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    ref, rdata, S_ref = H.run_oracle(*case, params=phys.parameters)
    ref64, rdata64, _ = H.run_oracle(*case, params=phys.parameters, dtype=np.float64)
    errs = {}
    for f in FIELDS:
        ok, e32, e64, eref = H.parity(_np(getattr(tend, f)), getattr(ref, f), getattr(ref64, f), tol)
        errs[f] = (ok, e32)
        print(f"{label} {f}: rel err vs fp32 oracle {e32:.2e}, vs fp64 oracle {e64:.2e} "
              f"(fp32 oracle vs fp64 {eref:.2e})", flush=True)
    kx, ix, il = case[2].nodal_shape
    S = phys.level_sets(ix, il, 1)[0]
    assert S == S_ref, f"level set {bin(S)} != {bin(S_ref)}"
    np.testing.assert_array_equal(_np(data.convection.iptop).astype(np.int32), rdata.convection.iptop)
    np.testing.assert_array_equal(_np(data.shortwave_rad.icltop).astype(np.int32), rdata.shortwave_rad.icltop)
    for f, (ok, e32) in errs.items():
        good = (e32 <= tol) if strict else ok
        assert good, f"{label} {f}: {e32:.3e}"
    assert phys.barrier_status(ix, il, 1) == 0, "level-set wait gave up"
    check_fwd.rdata64 = rdata64
    return tend, data, rdata


def test_fwd_t31_aquaplanet(phys):
    """Plain forward, T31 (96 x 48), 8 levels, aquaplanet recipe, seed 0."""
    check_fwd(phys, synthetic.make_case(96, 48, 8, seed=0, land=False), label="T31")


def test_fwd_t85_land(phys):
    """Plain forward, T85 (256 x 128 = 32768 columns), 8 levels, land recipe, seed 1."""
    check_fwd(phys, synthetic.make_case(256, 128, 8, seed=1, land=True), label="T85")


def wet(case, seed, hi=1.15):
    state, forcing, geometry, date = case
    rng = np.random.default_rng(seed)
    rh = rng.uniform(0.5, hi, state.temperature.shape)
    q = rh * synthetic.qsat_np(state.temperature, state.normalized_surface_pressure[None],
                               geometry.fsg[:, None, None].astype(np.float64))
    return (state.copy(specific_humidity=q.astype(np.float32)), forcing, geometry, date)


@pytest.mark.parametrize("kx,step,lfl", [(8, 0, True), (7, 0, True), (16, 0, True), (8, 1, True), (8, 0, False)])
def test_fwd_wet_heterogeneous_cloud_tops(phys, kx, step, lfl):
    """Condensing columns give cloud tops at many levels, so S (Q1) has many bits."""
    st, fo, ge, da = wet(synthetic.make_case(64, 32, kx, seed=3, land=True), 5)
    case = (st, fo.copy(lfluxland=lfl), ge, da.copy(model_step=step))
    _, _, rdata = check_fwd(phys, case, synthetic.draw_vjp_parameters(2), label=f"wet kx{kx}", strict=False)
    assert len(np.unique(rdata.shortwave_rad.icltop)) >= 4


@pytest.mark.parametrize("term", [3, 4, 6, 8, 9, 10])
def test_fwd_single_terms(phys, term):
    import torch
    case = wet(synthetic.make_case(64, 32, 8, seed=3, land=True), 5)
    phys.parameters = Parameters.default()
    tend, _ = phys.compute_tendencies(*case, term_mask=1 << term)
    torch.cuda.synchronize()
    ref, _, _ = H.run_oracle(*case, term_mask=1 << term)
    ref64, _, _ = H.run_oracle(*case, term_mask=1 << term, dtype=np.float64)
    for f in FIELDS:
        g, r, r64 = _np(getattr(tend, f)), getattr(ref, f), getattr(ref64, f)
        if np.max(np.abs(r)) == 0:
            assert np.max(np.abs(g)) == 0
            continue
        ok, e32, e64, eref = H.parity(g, r, r64)
        print(f"term {so.TERMS[term]} {f}: {e32:.2e} (fp64 {e64:.2e}, oracle {eref:.2e})", flush=True)
        assert ok and e32 <= 2e-5, f"{f}: {e32:.3e}"


def test_terms_api(phys):
    """phys.terms[i](state, physics_data, parameters, forcing, geometry)
    (speedy_physics.py:123-125): each term's tendency is the term_mask = 1 << i
    value bit for bit, the terms chained in order sum to the full call, and the
    returned PhysicsData carries the date and the chain's diagnostics."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    st, fo, ge, da = case = wet(synthetic.make_case(64, 32, 8, seed=3, land=True), 5)
    params = synthetic.draw_vjp_parameters(4)
    dphys = SpeedyPhysics(parameters=params, with_diagnostics=True)
    full, full_data = dphys.compute_tendencies(*case)
    terms = phys.terms
    assert [t.__name__ for t in terms] == list(so.TERMS)
    data = phys.get_empty_data(ge, da)
    total = None
    for i, term in enumerate(terms):
        tend, data_i = term(st, data, params, fo, ge)
        masked, _ = dphys.compute_tendencies(*case, term_mask=1 << i)
        for f in FIELDS:
            assert torch.equal(getattr(tend, f), getattr(masked, f)), (term.__name__, f)
        assert data_i.date is da
        data = data_i
        total = tend if total is None else total + tend
    torch.cuda.synchronize()
    for f in FIELDS:
        g, r = _np(getattr(total, f)), _np(getattr(full, f))
        assert np.max(np.abs(g - r)) <= 1e-6 * max(np.max(np.abs(r)), 1e-30), f
    assert torch.equal(data._diag, full_data._diag)
    assert phys.parameters is not params
    with pytest.raises(ValueError):
        terms[3](st, None, params, fo, ge)


def test_fwd_level_set_override(phys):
    import torch
    case = wet(synthetic.make_case(64, 32, 8, seed=3, land=True), 5)
    phys.parameters = Parameters.default()
    tend, _ = phys.compute_tendencies(*case, level_set=0b0101010)
    torch.cuda.synchronize()
    ref, _, _ = H.run_oracle(*case, level_set=0b0101010)
    e = H.rel_inf(_np(tend.temperature), ref.temperature)
    assert e <= H.REL_TOL, f"{e:.3e}"


def _data_fields(rdata):
    checks = {
        "convection.cbmf": rdata.convection.cbmf, "convection.precnv": rdata.convection.precnv,
        "condensation.precls": rdata.condensation.precls, "shortwave_rad.cloudc": rdata.shortwave_rad.cloudc,
        "shortwave_rad.cloudstr": rdata.shortwave_rad.cloudstr, "shortwave_rad.qcloud": rdata.shortwave_rad.qcloud,
        "shortwave_rad.gse": rdata.shortwave_rad.gse, "shortwave_rad.rsds": rdata.shortwave_rad.rsds,
        "shortwave_rad.rsns": rdata.shortwave_rad.rsns, "shortwave_rad.ftop": rdata.shortwave_rad.ftop,
        "shortwave_rad.fsol": rdata.shortwave_rad.fsol, "shortwave_rad.ozone": rdata.shortwave_rad.ozone,
        "shortwave_rad.ozupp": rdata.shortwave_rad.ozupp, "shortwave_rad.zenit": rdata.shortwave_rad.zenit,
        "shortwave_rad.stratz": rdata.shortwave_rad.stratz, "longwave_rad.ftop": rdata.longwave_rad.ftop,
        "mod_radcon.albsfc": rdata.mod_radcon.albsfc, "mod_radcon.alb_l": rdata.mod_radcon.alb_l,
        "mod_radcon.alb_s": rdata.mod_radcon.alb_s, "mod_radcon.snowc": rdata.mod_radcon.snowc,
        "surface_flux.rlds": rdata.surface_flux.rlds,
        "surface_flux.rlns": rdata.surface_flux.rlns, "surface_flux.tsfc": rdata.surface_flux.tsfc,
        "surface_flux.tskin": rdata.surface_flux.tskin, "surface_flux.t0": rdata.surface_flux.t0,
        "humidity.qsat": rdata.humidity.qsat, "humidity.rh": rdata.humidity.rh, "convection.se": rdata.convection.se,
        "shortwave_rad.dfabs": rdata.shortwave_rad.dfabs, "longwave_rad.dfabs": rdata.longwave_rad.dfabs,
    }
    for i in range(3):
        for fld in ("ustr", "vstr", "shf", "evap", "rlus"):
            checks[f"surface_flux.{fld}.{i}"] = getattr(rdata.surface_flux, fld)[..., i]
    for i in range(4):
        checks[f"mod_radcon.flux.{i}"] = rdata.mod_radcon.flux[..., i]
        checks[f"mod_radcon.tau2.{i}"] = rdata.mod_radcon.tau2[..., i]
    for i in range(2):
        checks[f"mod_radcon.st4a.{i}"] = rdata.mod_radcon.st4a[..., i]
        checks[f"mod_radcon.stratc.{i}"] = rdata.mod_radcon.stratc[..., i]
        checks[f"surface_flux.hfluxn.{i}"] = rdata.surface_flux.hfluxn[..., i]
    return checks


def test_fwd_diagnostics(phys):
    """PhysicsData subset written by the kernel against the oracle's PhysicsData.
    Bound per field: within 2e-5 of the float32 oracle, or as close to the float64
    oracle as the float32 oracle is (x2).  (cloudc contains sqrt(86.4*(precnv+precls)),
    which magnifies the float32 cancellation in precnv near zero.)"""
    case = wet(synthetic.make_case(64, 32, 8, seed=4, land=True), 6)
    _, data, rdata = check_fwd(phys, case, label="diag", strict=False)
    d = data.to_dict()
    checks = _data_fields(rdata)
    checks64 = _data_fields(check_fwd.rdata64)
    worst = 0.0
    for k, ref in checks.items():
        g = _np(d[k])
        e32, e64, eref = H.rel_inf(g, ref), H.rel_inf(g, checks64[k]), H.rel_inf(ref, checks64[k])
        worst = max(worst, min(e32, e64))
        assert e32 <= 2e-5 or e64 <= max(2e-5, 2 * eref), f"{k}: {e32:.3e} (fp64 {e64:.3e}, oracle {eref:.3e})"
    print(f"diagnostics: {len(checks)} fields, worst rel err {worst:.2e}", flush=True)


@pytest.mark.parametrize("seed", [2, 9])
def test_fwd_divisions_are_correctly_rounded(phys, seed):
    """The column code divides with the IEEE sequence minus its range-scaling
    steps (mdiv, speedy_column.hpp), which must equal a correctly rounded float32
    division for every operand the physics produces.  Checked bit for bit on a
    whole T85 grid against NumPy's float32 division on the host: the relative
    humidity rh = Q / qsat at every level and the cloud scheme's
    gse = (se[-2] - se[-1]) / (phi[-2] - phi[-1]), from the kernel's own qsat and
    se diagnostics."""
    import torch
    case = wet(synthetic.make_case(256, 128, 8, seed=seed, land=True), seed)
    phys.parameters = Parameters.default()
    _, data = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    d = data.to_dict()
    state = case[0]
    qsat, rh = _np(d["humidity.qsat"]), _np(d["humidity.rh"])
    assert qsat.dtype == np.float32 and qsat.shape == state.specific_humidity.shape
    np.testing.assert_array_equal(rh, state.specific_humidity / qsat)
    se, phi = _np(d["convection.se"]), state.geopotential
    gse = (se[-2] - se[-1]) / (phi[-2] - phi[-1])
    assert gse.dtype == np.float32
    np.testing.assert_array_equal(_np(d["shortwave_rad.gse"]), gse)


def test_fwd_increase_co2_snow_and_sea_ice(phys):
    """set_forcing (speedy/forcing.py:11-39) off its defaults: the CO2 trend on
    (increase_co2, 80 years after co2_year_ref), snow cover above 1 in part of the
    land (clamped to 1), sea ice between 0 and 1.  Tendencies meet the strict bar;
    snowc, alb_l, alb_s, albsfc and the longwave transmissivities (tau2, band 1
    carries ablco2) equal the float32 oracle's; the longwave tendencies move
    against the default forcing, so the trend branch really ran."""
    import torch
    st, fo, ge, da = synthetic.make_case(96, 48, 8, seed=9, land=True)
    rng = np.random.default_rng(11)
    snow = np.asarray(fo.snowc_am, np.float32).copy()
    snow[rng.random(snow.shape) < 0.3] = rng.uniform(1.0, 3.0)
    sice = np.where(rng.random(snow.shape) < 0.4, rng.uniform(0.0, 1.0, snow.shape), 0.0).astype(np.float32)
    case = (st, fo.copy(snowc_am=snow, sice_am=sice), ge, da.copy(model_year=2030, tyear=0.625))
    params = Parameters.default().replace(forcing=dict(increase_co2=True))
    tend, data, rdata = check_fwd(phys, case, params, label="co2 trend")
    d = data.to_dict()
    for k, ref in (("mod_radcon.snowc", rdata.mod_radcon.snowc), ("mod_radcon.alb_l", rdata.mod_radcon.alb_l),
                   ("mod_radcon.alb_s", rdata.mod_radcon.alb_s), ("mod_radcon.albsfc", rdata.mod_radcon.albsfc)):
        e = H.rel_inf(_np(d[k]), ref)
        assert e <= 1e-6, f"{k}: {e:.3e}"
    assert np.max(rdata.mod_radcon.snowc) == 1.0 and np.any(snow > 1.0)
    # transmissivities: the diagnostics bar (band 0 carries the cloud cover, whose
    # sqrt of the precipitation magnifies float32 cancellation in precnv)
    r64 = check_fwd.rdata64.mod_radcon.tau2
    for b in range(4):
        g = _np(d[f"mod_radcon.tau2.{b}"])
        e32, e64 = H.rel_inf(g, rdata.mod_radcon.tau2[..., b]), H.rel_inf(g, r64[..., b])
        eref = H.rel_inf(rdata.mod_radcon.tau2[..., b], r64[..., b])
        print(f"co2 trend tau2.{b}: {e32:.2e} (fp64 {e64:.2e}, oracle {eref:.2e})", flush=True)
        assert e32 <= 2e-5 or e64 <= max(2e-5, 2 * eref), f"tau2.{b}: {e32:.3e} (fp64 {e64:.3e}, oracle {eref:.3e})"
    phys.parameters = Parameters.default()
    tend0, _ = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    assert H.rel_inf(_np(tend.temperature), _np(tend0.temperature)) > 1e-4


def test_fwd_members_stacked(phys):
    """Three stacked members in one call: each equals its own single-member result,
    with its own level set."""
    import torch
    phys.parameters = Parameters.default()
    st, fo, ge, da = synthetic.make_case(64, 32, 8, seed=7, land=True, nmember=3, member_t_sigma=0.5)
    q = st.specific_humidity.copy()
    q[1] *= 1.15   # member 1 condenses widely: a different level set from member 0
    st = st.copy(specific_humidity=q)
    tend, _ = phys.compute_tendencies(st, fo, ge, da)
    torch.cuda.synchronize()
    S = phys.level_sets(64, 32, 3)
    for m in range(3):
        ref, _, S_ref = H.run_oracle(st, fo, ge, da, member=m)
        ref64, _, _ = H.run_oracle(st, fo, ge, da, member=m, dtype=np.float64)
        assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            ok, e32, e64, eref = H.parity(_np(getattr(tend, f)[m]), getattr(ref, f), getattr(ref64, f))
            print(f"member {m} {f}: {e32:.2e} (fp64 {e64:.2e}, oracle {eref:.2e})", flush=True)
            assert ok, f"member {m} {f}: {e32:.3e}"
    assert S[0] != S[1]


@pytest.mark.parametrize("nlon,nlat,nmember,kx", [(100, 50, 3, 8), (90, 45, 2, 8), (100, 50, 2, 16)])
def test_fwd_ragged_members(phys, nlon, nlat, nmember, kx):
    """Members whose column count is not a multiple of the 256-column workgroup (nor,
    for 90 x 45, of the 64-lane wave): workgroups straddle two members, whose level
    sets and float64 budgets must stay apart.  Every member matches its own oracle."""
    import torch
    phys.parameters = Parameters.default()
    st, fo, ge, da = synthetic.make_case(nlon, nlat, kx, seed=11, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[1::2] *= np.float32(1.15)
    case = (st.copy(specific_humidity=q), fo, ge, da)
    tend, _ = phys.compute_tendencies(*case, budgets=True)
    torch.cuda.synchronize()
    S = phys.level_sets(nlon, nlat, nmember)
    bud = phys.last_budgets.cpu().numpy()
    assert phys.barrier_status(nlon, nlat, nmember) == 0
    for m in range(nmember):
        ref, _, S_ref = H.run_oracle(*case, member=m)
        ref64, _, _ = H.run_oracle(*case, member=m, dtype=np.float64)
        assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            ok, e32, e64, eref = H.parity(_np(getattr(tend, f)[m]), getattr(ref, f), getattr(ref64, f))
            assert ok, f"member {m} {f}: {e32:.3e} (fp64 {e64:.2e}, oracle {eref:.2e})"
        one = H.subgrid(case, np.arange(nlon), np.arange(nlat), m)
        rb = H.oracle_budgets(one)
        for i in (0, 1, 2, 3):
            scale = max(abs(rb[2]), abs(rb[3]), 1e-12) if i else abs(rb[0])
            assert abs(bud[m, i] - rb[i]) <= 1e-5 * scale + (1e-3 if i == 0 else 0.0), (m, i, bud[m, i], rb[i])
    assert S[0] != S[1]


@pytest.mark.parametrize("nlon,nlat,nmember,level_set", [(256, 128, 3, -1), (16, 8, 5, -1), (64, 32, 1, 0b0101010)])
def test_fwd_one_segment_batch_route(nlon, nlat, nmember, level_set):
    """Calls the fused launch cannot take run as a one-segment batch through the batch
    kernels: beyond one co-resident launch (T85 x 3 members = 384 workgroups), members
    narrower than a workgroup (16 x 8 columns), and a given level set.  Every member
    matches its own oracle, with its own level set and float64 budgets."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()   # no diagnostics
    st, fo, ge, da = synthetic.make_case(nlon, nlat, 8, seed=13, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    if nmember > 1:
        q[1::2] *= np.float32(1.15)
    case = (st.copy(specific_humidity=q), fo, ge, da)
    tend, _ = phys.compute_tendencies(*case, level_set=level_set, budgets=True)
    torch.cuda.synchronize()
    S = phys.level_sets(nlon, nlat, nmember)
    bud = phys.last_budgets.cpu().numpy()
    for m in range(nmember):
        mm = m if nmember > 1 else None
        ref, _, S_ref = H.run_oracle(*case, member=mm, level_set=None if level_set < 0 else level_set)
        ref64, _, _ = H.run_oracle(*case, member=mm, dtype=np.float64, level_set=None if level_set < 0 else level_set)
        if level_set < 0:
            assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            g = _np(getattr(tend, f)[m]) if nmember > 1 else _np(getattr(tend, f))
            ok, e32, e64, eref = H.parity(g, getattr(ref, f), getattr(ref64, f))
            assert ok, f"member {m} {f}: {e32:.3e} (fp64 {e64:.2e}, oracle {eref:.2e})"
        if level_set < 0:
            one = H.subgrid(case, np.arange(nlon), np.arange(nlat), mm)
            rb = H.oracle_budgets(one)
            wscale = max(abs(rb[2]), abs(rb[3]), 1e-12)
            assert abs(bud[m, 0] - rb[0]) <= 1e-5 * abs(rb[0]) + 1e-3, (m, bud[m, 0], rb[0])
            for i in (1, 2, 3):
                assert abs(bud[m, i] - rb[i]) <= 1e-5 * wscale, (m, i, bud[m, i], rb[i])


@pytest.mark.parametrize("nlon,nlat", [(256, 128), (64, 32)])
def test_fwd_diagnostics_batch_route_equals_fused(nlon, nlat):
    """A call with diagnostics and a given level set runs as a one-segment batch
    through fwd_diag_batch_kernel.  Given the level set the fused launch computed
    for the same state, its tendencies and PhysicsData subset equal the fused
    call's (within 1e-6 relative; the same column code in both kernels)."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(with_diagnostics=True)
    case = wet(synthetic.make_case(nlon, nlat, 8, seed=19, land=True), 21)
    t0, d0 = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    S0 = phys.level_sets(nlon, nlat, 1)[0]
    t1, d1 = phys.compute_tendencies(*case, level_set=S0)
    torch.cuda.synchronize()
    identical = True
    for f in FIELDS:
        a, b = _np(getattr(t1, f)), _np(getattr(t0, f))
        err = float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-30)
        assert err <= 1e-6, f"{f}: {err:.2e}"
        identical &= err == 0.0
    a, b = _np(d1._diag), _np(d0._diag)
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-30)
    err = float((np.abs(a - b) / scale).max())
    assert err <= 1e-6, f"diagnostics: {err:.2e}"
    identical &= err == 0.0
    print(f"{nlon}x{nlat}: batch route == fused bit for bit: {identical}", flush=True)


@pytest.mark.parametrize("kx", [8, 7, 16])
def test_fwd_diagnostics_narrow_members(kx):
    """Members narrower than a workgroup (16 x 8 columns, 5 members) with
    diagnostics: level-set launch, then fwd_diag_batch_kernel, at 8, 7 and 16
    levels.  Per member: level set, tendencies and the PhysicsData subset against
    the member's oracle run (diagnostic fields bound as in test_fwd_diagnostics)."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(with_diagnostics=True)
    nlon, nlat, nmember = 16, 8, 5
    st, fo, ge, da = synthetic.make_case(nlon, nlat, kx, seed=23, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[1::2] *= np.float32(1.15)
    case = (st.copy(specific_humidity=q), fo, ge, da)
    tend, data = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    S = phys.level_sets(nlon, nlat, nmember)
    d = data.to_dict()
    for m in range(nmember):
        ref, rdata, S_ref = H.run_oracle(*case, member=m)
        ref64, rdata64, _ = H.run_oracle(*case, member=m, dtype=np.float64)
        assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            ok, e32, e64, eref = H.parity(_np(getattr(tend, f)[m]), getattr(ref, f), getattr(ref64, f))
            assert ok, f"member {m} {f}: {e32:.3e} (fp64 {e64:.2e}, oracle {eref:.2e})"
        checks, checks64 = _data_fields(rdata), _data_fields(rdata64)
        for k, r in checks.items():
            g = _np(d[k])[m]
            e32, e64, eref = H.rel_inf(g, r), H.rel_inf(g, checks64[k]), H.rel_inf(r, checks64[k])
            assert e32 <= 2e-5 or e64 <= max(2e-5, 2 * eref), f"member {m} {k}: {e32:.3e} (fp64 {e64:.3e})"


@pytest.mark.parametrize("nlon,nlat,nmember", [(256, 128, 3), (16, 8, 5)])
def test_fwd_batch_route_plain_kernel_equals_extra(nlon, nlat, nmember):
    """One-segment batch forwards without the verify steps or budgets run
    fwd_plain_batch_kernel (the forward without that optional code); with budgets
    they run fwd_only_batch_kernel, which test_fwd_one_segment_batch_route holds to
    the oracle.  The tendencies agree bit for bit."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    st, fo, ge, da = synthetic.make_case(nlon, nlat, 8, seed=37, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[1::2] *= np.float32(1.15)
    case = (st.copy(specific_humidity=q), fo, ge, da)
    t_plain, _ = phys.compute_tendencies(*case)
    t_extra, _ = phys.compute_tendencies(*case, budgets=True)
    torch.cuda.synchronize()
    for f in FIELDS:
        np.testing.assert_array_equal(_np(getattr(t_plain, f)), _np(getattr(t_extra, f)), err_msg=f)


@pytest.mark.parametrize("kx", [7, 16])
def test_fwd_batch_route_other_level_counts(kx):
    """Members narrower than a workgroup (16 x 8 columns, 5 members) at 7 and 16
    levels through the batch kernels without the optional code
    (fwd_plain_batch_kernel<7>, <16>): every member's level set and tendencies
    against its oracle run, and bit for bit against the same call with budgets
    (fwd_only_batch_kernel)."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    nlon, nlat, nmember = 16, 8, 5
    st, fo, ge, da = synthetic.make_case(nlon, nlat, kx, seed=41, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[1::2] *= np.float32(1.15)
    case = (st.copy(specific_humidity=q), fo, ge, da)
    t_plain, _ = phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    S = phys.level_sets(nlon, nlat, nmember)
    t_extra, _ = phys.compute_tendencies(*case, budgets=True)
    torch.cuda.synchronize()
    for m in range(nmember):
        ref, _, S_ref = H.run_oracle(*case, member=m)
        ref64, _, _ = H.run_oracle(*case, member=m, dtype=np.float64)
        assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            ok, e32, e64, eref = H.parity(_np(getattr(t_plain, f)[m]), getattr(ref, f), getattr(ref64, f))
            assert ok, f"kx {kx} member {m} {f}: {e32:.3e} (fp64 {e64:.2e}, oracle {eref:.2e})"
    for f in FIELDS:
        np.testing.assert_array_equal(_np(getattr(t_plain, f)), _np(getattr(t_extra, f)), err_msg=f)


def test_fwd_repeated_calls_reset_level_set(phys):
    """The level set of one call does not leak into the next (self-resetting workspace)."""
    import torch
    phys.parameters = Parameters.default()
    wet_case = wet(synthetic.make_case(64, 32, 8, seed=3, land=True), 5)
    dry_case = synthetic.make_case(64, 32, 8, seed=3, land=True)
    for case in (wet_case, dry_case, wet_case):
        phys.compute_tendencies(*case)
        torch.cuda.synchronize()
        _, _, S_ref = H.run_oracle(*case)
        S = phys.level_sets(64, 32, 1)[0]
        assert S == S_ref, (bin(S), bin(S_ref))


def test_fwd_forcing_climatology_day_rows():
    """Forcing climatologies (ix, il, 365), as NumPy arrays and as device tensors, give
    exactly the tendencies of their day slice (the library lays each out once as
    [day][ncol] and passes the day's row); the day is clamped like the reference's
    gather (day 365 -> 364); the oracle, slicing the same climatologies, agrees; and
    an in-place update of a device climatology is seen by the next call."""
    import torch
    from jcm_amd.forcing import FORCING_FIELDS
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    st, fo, ge, da = synthetic.make_case(64, 32, 8, seed=2, land=True)
    days = np.arange(365, dtype=np.float32)
    wave = lambda amp, ph: (1.0 + amp * np.sin(2 * np.pi * (days + ph) / 365.0)).astype(np.float32)
    clim = {f: (np.asarray(getattr(fo, f), np.float32)[..., None] * wave(0.02, 30 * i)).astype(np.float32)
            for i, f in enumerate(FORCING_FIELDS)}
    fo3 = fo.copy(**clim)
    dev = phys.device
    fo3_dev = fo.copy(**{f: torch.as_tensor(v, device=dev) for f, v in clim.items()})
    for tyear, day in ((0.0, 0), (100.4 / 365.2425, 100), (1.0, 364)):
        date = da.copy(tyear=tyear)
        assert min(date.model_day(), 364) == day
        fo2 = fo.copy(**{f: np.ascontiguousarray(v[:, :, day]) for f, v in clim.items()})
        t2, _ = phys.compute_tendencies(st, fo2, ge, date)
        t3, _ = phys.compute_tendencies(st, fo3, ge, date)
        t3d, _ = phys.compute_tendencies(st, fo3_dev, ge, date)
        torch.cuda.synchronize()
        for f in FIELDS:
            ref2 = _np(getattr(t2, f))
            np.testing.assert_array_equal(_np(getattr(t3, f)), ref2, err_msg=f"day {day} {f} numpy")
            np.testing.assert_array_equal(_np(getattr(t3d, f)), ref2, err_msg=f"day {day} {f} device")
        if day == 100:
            ref, _, _ = H.run_oracle(st, fo3, ge, date)
            ref64, _, _ = H.run_oracle(st, fo3, ge, date, dtype=np.float64)
            for f in FIELDS:
                ok, e32, e64, eref = H.parity(_np(getattr(t3, f)), getattr(ref, f), getattr(ref64, f))
                print(f"climatology day 100 {f}: vs fp32 oracle {e32:.2e}, vs fp64 {e64:.2e}", flush=True)
                assert ok, f"{f}: {e32:.3e}"
    # an in-place update of a device climatology reaches the next call
    date = da.copy(tyear=100.4 / 365.2425)
    fo3_dev.sea_surface_temperature[:, :, 100] += 1.0
    t3d, _ = phys.compute_tendencies(st, fo3_dev, ge, date)
    fo2 = fo.copy(**{f: np.ascontiguousarray(v[:, :, 100]) for f, v in clim.items()})
    fo2 = fo2.copy(sea_surface_temperature=fo2.sea_surface_temperature + np.float32(1.0))
    t2, _ = phys.compute_tendencies(st, fo2, ge, date)
    torch.cuda.synchronize()
    for f in FIELDS:
        np.testing.assert_array_equal(_np(getattr(t3d, f)), _np(getattr(t2, f)), err_msg=f"updated {f}")


def test_fwd_convection_fortran_fixture(phys):
    """The reference's Fortran convection fixture (convection_test.py:109-140) through the
    full kernel with only the convection tendency summed."""
    import torch
    f90 = np.load(os.path.join(HERE, "golden", "convection_f90.npz"), allow_pickle=False)
    ix, il, kx = 96, 48, 8
    radang, sia, coa = so.speedy_latitudes(il)
    ge = Geometry.from_latitudes(radang, ix, kx, sia=sia, coa=coa)
    fsg = ge.fsg
    ps = np.ones((ix, il), np.float32)
    expo = np.float32(0.05) * np.cos(np.float32(3) * np.arange(il, dtype=np.float32) / np.float32(il)) ** 3
    ta = (np.float32(300) * np.ones((kx, ix, il), np.float32) * (fsg[:, None, None] ** expo)).astype(np.float32)
    qsat = so.get_qsat(ta, ps, fsg[:, None, None])
    qa = (np.sin(np.float32(2) * np.arange(ix, dtype=np.float32)[:, None] / np.float32(ix)) ** 2 * qsat
          * np.float32(3.5)).astype(np.float32)
    phi = (np.float32(so.rgas) * ta * np.log(fsg[:, None, None])).astype(np.float32)
    z = np.zeros((kx, ix, il), np.float32)
    st = PhysicsState(z, z, ta, qa, phi, ps)
    fo = ForcingData.zeros((ix, il), lfluxland=False)
    from jcm_amd.date import DateData
    phys.parameters = Parameters.default()
    tend, data = phys.compute_tendencies(st, fo, ge, DateData(), term_mask=1 << 3)
    torch.cuda.synchronize()
    np.testing.assert_allclose(_np(data.convection.cbmf), f90["cbmf"], atol=1e-4, rtol=0)
    np.testing.assert_allclose(_np(data.convection.precnv), f90["precnv"], atol=1e-4, rtol=0)
    dfse, dfqa = f90["dfse"], f90["dfqa"]
    tt = dfse.copy(); tt[1:] = dfse[1:] * ge.grdscp[1:, None, None]
    qt = dfqa.copy(); qt[1:] = dfqa[1:] * ge.grdsig[1:, None, None]
    np.testing.assert_allclose(_np(tend.temperature), tt, atol=1e-4, rtol=0)
    np.testing.assert_allclose(_np(tend.specific_humidity), qt, atol=1e-4, rtol=0)


@pytest.mark.parametrize("nlon,nlat,nmember", [(128, 64, 1), (128, 64, 3), (256, 128, 3)])
def test_fwd_graph_replay_matches_eager(nlon, nlat, nmember):
    """Calls captured in a HIP graph (a plain launch of the fused kernel, or at T85 x 3
    members the one-segment batch with its descriptor kernel) give, on every replay,
    exactly the eager call's tendencies; eager work runs between replays."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    st, fo, ge, da = synthetic.make_case(nlon, nlat, 8, seed=4, land=True, nmember=nmember, member_t_sigma=1.0)
    st = st.copy(specific_humidity=(st.specific_humidity * np.float32(1.1)).astype(np.float32))
    dev = phys.device
    st = st.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), device=dev)
                    for k in ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential",
                              "normalized_surface_pressure")})
    eager, _ = phys.compute_tendencies(st, fo, ge, da)
    eager = [getattr(eager, f).clone() for f in FIELDS]
    S_eager = phys.level_sets(nlon, nlat, nmember)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        phys.compute_tendencies(st, fo, ge, da)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        outs = [phys.compute_tendencies(st, fo, ge, da)[0] for _ in range(3)]
    for rep in range(3):
        for o in outs:
            for f in FIELDS:
                getattr(o, f).fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        phys.compute_tendencies(st, fo, ge, da)   # eager work between replays
        torch.cuda.synchronize()
        for o in outs:
            for f, e in zip(FIELDS, eager):
                assert torch.equal(getattr(o, f), e), (rep, f)
        assert phys.level_sets(nlon, nlat, nmember) == S_eager
        assert phys.barrier_status(nlon, nlat, nmember) == 0


TWO_LAUNCH_SKIP = ("GPU memory faults were reported at the first synchronize of this test twice (T85 x 8 members, "
                   "then 8 x 4 x 6 members), each time right after the graph-replay test, with a plain-forward "
                   "two-launch kernel that has since been taken out (the library no longer has the two-launch "
                   "kernels; these grids now run as a one-segment batch); a CPU simulation of every index of the "
                   "path is in bounds.  No cause has been shown, so the test stays off the GPU.")


@pytest.mark.skip(reason=TWO_LAUNCH_SKIP)
@pytest.mark.parametrize("nlon,nlat,nmember", [(256, 128, 8), (8, 4, 6)])
def test_fwd_two_kernel_level_set_path(nlon, nlat, nmember):
    """Grids beyond one co-resident launch (T85 x 8 members = 262144 columns) and
    members narrower than a workgroup (8 x 4 columns) take the two-launch path
    (level-set kernel, then forward); every member matches its own oracle run."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    st, fo, ge, da = synthetic.make_case(nlon, nlat, 8, seed=9, land=True, nmember=nmember, member_t_sigma=1.0)
    q = st.specific_humidity.copy()
    q[1::2] *= np.float32(1.12)   # odd members condense widely: member-specific level sets
    st = st.copy(specific_humidity=q)
    phys._lib.speedy_timer_enable(8)
    tend, _ = phys.compute_tendencies(st, fo, ge, da)
    torch.cuda.synchronize()
    from jcm_amd import _native
    launches = _native.timer_read()
    phys._lib.speedy_timer_disable()
    assert launches["levelset"][1] == 1 and launches["fwd"][1] == 1, launches
    S = phys.level_sets(nlon, nlat, nmember)
    for m in range(nmember):
        ref, _, S_ref = H.run_oracle(st, fo, ge, da, member=m)
        ref64, _, _ = H.run_oracle(st, fo, ge, da, member=m, dtype=np.float64)
        assert S[m] == S_ref, (m, bin(S[m]), bin(S_ref))
        for f in FIELDS:
            ok, e32, e64, eref = H.parity(_np(getattr(tend, f)[m]), getattr(ref, f), getattr(ref64, f))
            assert ok, f"member {m} {f}: {e32:.3e} (fp64 {e64:.2e}, oracle {eref:.2e})"


def test_fwd_fused_verify_steps(phys):
    """verify_state -> compute_tendencies -> verify_tendencies (physics_interface.py:371-375)
    fused into the kernel, against the oracle's sequence; negative humidities and a
    one-hour step so that both steps act."""
    import torch
    from test_host_column import _negative_q_case
    case = _negative_q_case()
    phys.parameters = Parameters.default()
    tend, _ = phys.compute_tendencies(*case, verify=True)
    torch.cuda.synchronize()
    ref, _, S_ref = H.run_oracle_verified(*case)
    ref64, _, _ = H.run_oracle_verified(*case, dtype=np.float64)
    assert phys.level_sets(64, 32, 1)[0] == S_ref
    for f in FIELDS:
        ok, e32, e64, eref = H.parity(_np(getattr(tend, f)), getattr(ref, f), getattr(ref64, f))
        print(f"verify {f}: {e32:.2e} (fp64 {e64:.2e})", flush=True)
        assert ok, f"{f}: {e32:.3e}"


@pytest.mark.parametrize("nlon,nlat,verify", [(256, 128, False), (64, 32, True)])
def test_fwd_float64_budgets(phys, nlon, nlat, verify):
    """Float64 column energy / water budgets reduced on the GPU against the same sums of
    the float64 oracle's tendencies: each global mean within 1e-5 of its scale.  Without
    the verify steps the column water budget closes (residual <= 1e-5 of the water-flux
    scale; the float64 oracle's is ~2e-7); with them the limiter's non-conservation
    matches the oracle's."""
    import torch
    from test_host_column import _negative_q_case
    case = _negative_q_case() if verify else synthetic.make_case(nlon, nlat, 8, seed=1, land=True)
    phys.parameters = Parameters.default()
    phys.compute_tendencies(*case, verify=verify, budgets=True)
    torch.cuda.synchronize()
    b = phys.last_budgets.cpu().numpy()[0]
    ref = H.oracle_budgets(case, verify=verify)
    print("budgets gpu", b, "oracle64", ref, flush=True)
    wscale = max(abs(ref[2]), abs(ref[3]))
    assert abs(b[0] - ref[0]) <= 1e-5 * abs(ref[0]) + 1e-3
    for i in (1, 2, 3):
        assert abs(b[i] - ref[i]) <= 1e-5 * wscale, i
    assert abs(b[4] - ref[4]) <= 1e-5 * wscale
    if not verify:   # the column water budget closes (verify_tendencies' limiter breaks it, as in the reference)
        assert abs(b[4]) <= 1e-5 * wscale and abs(ref[4]) <= 1e-6 * wscale
    assert b[5] >= 0.0


def test_diagnostics_collector_running_mean():
    """DiagnosticsCollector (model.py:46-64) on the device: over 2 outputs of 3 physics
    steps each, with a second accumulate per step that must be a no-op, the averages
    equal the reference's float32 sequence data[i] += new / steps bit for bit."""
    import torch
    from jcm_amd import DiagnosticsCollector
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(with_diagnostics=True)
    st, fo, ge, da = synthetic.make_case(64, 32, 8, seed=7, land=True)
    col = DiagnosticsCollector(steps_to_average=3)
    col.init_data(phys.get_empty_data(ge), outer_steps=2)
    ref = np.zeros(tuple(col.data._diag.shape), np.float32)
    for o in range(2):
        for k in range(3):
            case = wet((st, fo, ge, da.copy(model_step=k)), 10 * o + k)
            col.physical_step = True
            _, data = phys.compute_tendencies(*case)
            col.accumulate_if_physical_step(data)
            col.accumulate_if_physical_step(data)   # physical_step is now False: no-op
            torch.cuda.synchronize()
            ref[o] = ref[o] + _np(data._diag) / np.float32(3)
        col.i += 1
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(col.data._diag), ref)
    rsds = col.data.shortwave_rad.rsds            # the struct views work on the stacked averages
    assert tuple(rsds.shape) == (2, 64, 32)


@pytest.mark.parametrize("n,offset", [(1, 0), (4099, 0), (4099, 1), (1 << 20, 3)])
def test_diag_accumulate_tails_and_alignment(n, offset):
    """speedy_diag_accumulate on lengths that are not a multiple of 4 and on buffers
    that are not 16-byte aligned (scalar path) against the float32 sequence."""
    import ctypes
    import torch
    from jcm_amd import _native
    lib = _native.load()
    rng = np.random.default_rng(n + offset)
    d = rng.standard_normal(n + 8).astype(np.float32)
    a = rng.standard_normal(n + 8).astype(np.float32)
    dd, ad = torch.as_tensor(d, device="cuda"), torch.as_tensor(a, device="cuda")
    for _ in range(2):
        rc = lib.speedy_diag_accumulate(dd[offset:].data_ptr(), ad[offset:].data_ptr(), n, 7.0,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0, _native.load().speedy_last_error()
        a[offset:offset + n] = a[offset:offset + n] + d[offset:offset + n] / np.float32(7.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(ad), a)


@pytest.mark.parametrize("vjp", [False, True])
def test_failed_level_set_wait_is_reported_and_reset(vjp):
    """A fused call whose level-set wait gave up must not pass unnoticed.  The
    condition is forced from the host, as a timed-out wait leaves it: the
    workspace's error word is set and its level-set words are left stale (no
    kernel waits or stalls).  barrier_status() reports the flag and resets the
    whole workspace, after which a call equals a clean call bit for bit; with
    check_barrier=True the call itself raises."""
    import torch
    from jcm_amd import _native
    from jcm_amd.physics_interface import PhysicsTendency
    from jcm_amd.speedy_physics import SpeedyPhysics
    case = synthetic.make_case(64, 32, 8, seed=3, land=True)
    kx, ix, il = case[2].nodal_shape
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    cot = PhysicsTendency(*[torch.randn((kx, ix, il), device=phys.device) for _ in range(4)])
    run = (lambda p: p.compute_tendencies_vjp(*case, cot)) if vjp else (lambda p: p.compute_tendencies(*case))
    clean = run(phys)
    torch.cuda.synchronize()
    ws = phys.workspace(ix, il, 1)
    ws.view(torch.int32)[1] = 1            # the error word, as a timed-out wait sets it
    assert phys.barrier_status(ix, il, 1) == 1
    assert int(ws.view(torch.int32).abs().max()) == 0, "the workspace was not reset"
    again = run(phys)
    torch.cuda.synchronize()
    for f in FIELDS:
        assert torch.equal(getattr(again[0], f), getattr(clean[0], f)), f
    assert phys.barrier_status(ix, il, 1) == 0
    checked = SpeedyPhysics(parameters=phys.parameters, check_barrier=True)
    run(checked)
    checked.workspace(ix, il, 1).view(torch.int32)[1] = 1
    with pytest.raises(_native.SpeedyError, match="gave up"):
        run(checked)
    again2 = run(checked)   # the workspace was reset: clean again
    torch.cuda.synchronize()
    assert torch.equal(again2[0].temperature, clean[0].temperature)


@pytest.mark.parametrize("kx", [7, 8, 16])
@pytest.mark.parametrize("variant", ["plain", "diag", "verify_budgets", "diag_budgets"])
def test_fwd_fused_equals_batch_route(kx, variant):
    """Every single-launch forward kernel (plain, with diagnostics, with the verify
    steps and float64 budgets) gives, bit for bit, what the batch route gives with
    the same level set: the same column code, whichever kernel runs it.  Three
    calls in a row on one workspace (no state carried between calls)."""
    import torch
    from jcm_amd.speedy_physics import SpeedyPhysics
    case = synthetic.make_case(64, 32, kx, seed=1, land=True)
    diag = variant.startswith("diag")
    budgets = variant.endswith("budgets")
    verify = variant == "verify_budgets"
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2), with_diagnostics=diag)
    for _ in range(3):
        t_f, d_f = phys.compute_tendencies(*case, verify=verify, budgets=budgets)
        b_f = phys.last_budgets.clone() if budgets else None
        S = phys.level_sets(64, 32, 1)[0]
        t_b, d_b = phys.compute_tendencies(*case, verify=verify, budgets=budgets, level_set=S)
        b_b = phys.last_budgets.clone() if budgets else None
        torch.cuda.synchronize()
        for f in FIELDS:
            bad = int((getattr(t_f, f) != getattr(t_b, f)).sum())
            assert bad == 0, f"kx{kx} {variant} {f}: {bad} entries differ from the batch route"
        if diag:
            assert torch.equal(d_f._diag, d_b._diag), f"kx{kx} {variant}: diagnostics differ"
        if budgets:
            assert torch.equal(b_f, b_b), f"kx{kx} {variant}: budgets {b_f} vs {b_b}"
