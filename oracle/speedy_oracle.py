"""CPU restatement of JCM's SPEEDY column physics -- TEST INFRASTRUCTURE ONLY.

This module is the oracle for the HIP kernels.  It restates, expression by
expression, the eleven SPEEDY physics terms of the reference
(``jcm/physics/speedy/*.py``) in NumPy.  It is generic in the floating-point
type: with ``float32`` inputs every intermediate stays float32 (operation
order follows the reference expressions), with ``float64`` inputs it is the
FP64 oracle the kernels' error is reported against.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker.  The product path
(``jcm_amd``) never imports it and has no CPU fallback.

Arrays are level-first: a 3-D field has shape ``(kx,) + H`` where ``H`` is any
horizontal shape, e.g. ``(ix, il)`` like the reference or ``(ncol,)``.
Per-latitude geometry (``sia``, ``coa``) must broadcast against ``H``.

Quirks of the reference that are reproduced on purpose are marked ``Q<n>``
(see the reference lines cited next to them):
  Q1 grid-wide cloud albedo scatter  (shortwave_radiation.py:63-70)
  Q2 LW-up consumes the SW-written flux  (shortwave_radiation.py:208,
     longwave_radiation.py:150,160)
  Q4 condensation resets iptop to 1  (large_scale_condensation.py:74)
  Q5 cloud top by float equality  (shortwave_radiation.py:367-373)
  Q6 out-of-range convective top scatter is dropped  (convection.py:321-323)

Parity status of each function (how it is pinned) is listed in
``PINNED_BY`` at the bottom of this file and checked by tests/test_oracle_*.
No reference vector exists for the composed call: ``compute_tendencies`` (the
eleven terms in sequence, Q1 with heterogeneous cloud tops and the Q2 hand-off
in composition) is parity unpinned by the reference; it is checked by the
pinned terms it composes, the column water budget identity and a brute-force
emulation of the Q1 scatter, and only an import of the reference could pin it.
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# constants  (jcm/constants.py:3-5, jcm/physics/speedy/physical_constants.py)
# --------------------------------------------------------------------------
p0 = 1e5           # constants.py:3
grav = 9.81        # constants.py:4
cp = 1004.0        # constants.py:5
akap = 2.0 / 7.0   # physical_constants.py:15
rgas = akap * cp   # physical_constants.py:16
alhc = 2501.0      # physical_constants.py:17
sbc = 5.67e-8      # physical_constants.py:19
solc = 342.0       # physical_constants.py:20
epssw = 0.020      # physical_constants.py:21
sd2sc = 60.0       # physical_constants.py:37
epsilon = 1e-9     # physical_constants.py:42
nstrad = 3         # physical_constants.py:44
ablco2_ref = 6.0   # physics_data.py:6
del_co2 = 0.005    # speedy/forcing.py:9
DAYS_YEAR = 365.2425  # date.py:7

# physical_constants.py:46-50; the 16-level set is this build's own choice
# (each 8-level layer split in two); the reference rejects kx outside {7, 8}.
SIGMA_LAYER_BOUNDARIES = {
    7: [0.0, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
    8: [0.0, 0.05, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
    16: [0.0, 0.025, 0.05, 0.095, 0.14, 0.2, 0.26, 0.34, 0.42, 0.51, 0.6,
         0.685, 0.77, 0.835, 0.9, 0.95, 1.0],
}


class Struct:
    """Minimal record with ``copy(**changes)`` (the reference uses tree_math structs)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def copy(self, **kw):
        d = dict(self.__dict__)
        d.update(kw)
        return Struct(**d)

    def __repr__(self):
        return "Struct(" + ", ".join(sorted(self.__dict__)) + ")"


def _lev(v, nh):
    """Reshape a per-level vector to broadcast against (kx,)+H."""
    v = np.asarray(v)
    return v.reshape(v.shape[:1] + (1,) * nh)


# --------------------------------------------------------------------------
# parameters  (params.py:9-254)
# --------------------------------------------------------------------------
def default_parameters():
    """Parameters.default() (params.py:244-254) plus this build's two scale
    factors (co2_scale on ablco2, drag_scale on cdl and cds; both 1.0 leaves
    every expression identical to the reference)."""
    return Struct(
        convection=Struct(psmin=0.8, trcnv=6.0, rhil=0.7, rhbl=0.9, entmax=0.5, smf=0.8),  # :19-27
        condensation=Struct(trlsc=4.0, rhlsc=0.9, drhlsc=0.1, rhblsc=0.95),  # :57-63
        shortwave_radiation=Struct(  # :102-126
            albcl=0.43, albcls=0.50, absdry=0.033, absaer=0.033, abswv1=0.022, abswv2=15.0,
            abscl1=0.015, abscl2=0.15, ablwin=0.3, ablwv1=0.7, ablwv2=50.0, ablcl1=12.0,
            ablcl2=0.6, rhcl1=0.30, rhcl2=1.00, qacl=0.20, wpcl=0.2, pmaxcl=10.0,
            clsmax=0.60, clsminl=0.15, gse_s0=0.25, gse_s1=0.40),
        mod_radcon=Struct(albsea=0.07, albice=0.60, albsn=0.60, epslw=0.05, emisfc=0.98),  # :143-150
        surface_flux=Struct(  # :186-204
            fwind0=0.95, ftemp0=1.0, fhum0=0.0, cdl=2.4e-3, cds=1.0e-3, chl=1.2e-3,
            chs=0.9e-3, vgust=5.0, ctday=1.0e-2, dtheta=3.0, fstab=0.67, clambda=7.0,
            clambsn=7.0, lscasym=True, lskineb=True, hdrag=2000.0, drag_scale=1.0),
        vertical_diffusion=Struct(trshc=6.0, trvdi=24.0, trvds=6.0, redshc=0.5, rhgrad=0.5,
                                  segrad=0.1),  # :221-229
        forcing=Struct(increase_co2=False, co2_year_ref=1950, co2_scale=1.0),  # :38-42
    )


def _cast_params(params, dt):
    """Float parameters become 0-d arrays of ``dt`` (jnp.array(x) is float32 in the
    reference); bools and ints stay Python values."""
    out = {}
    for grp, s in params.__dict__.items():
        d = {}
        for k, v in s.__dict__.items():
            d[k] = v if isinstance(v, (bool, int, np.bool_)) and not isinstance(v, float) else dt(v)
        out[grp] = Struct(**d)
    return Struct(**out)


# --------------------------------------------------------------------------
# geometry  (geometry.py:62-85 vertical; 130-132 horizontal)
# --------------------------------------------------------------------------
def vertical_geometry(kx, dtype=np.float32):
    """_initialize_vertical (geometry.py:62-85)."""
    dt = np.dtype(dtype).type
    hsg = np.asarray(SIGMA_LAYER_BOUNDARIES[kx], dtype=dt)
    fsg = (hsg[1:] + hsg[:-1]) / dt(2.0)
    dhs = np.diff(hsg)
    sigl = np.log(fsg)
    grdsig = dt(grav) / (dhs * dt(p0))
    grdscp = grdsig / dt(cp)
    wvi = np.zeros((kx, 2), dtype=dt)
    wvi[:-1, 0] = dt(1.0) / np.diff(sigl)
    wvi[:-1, 1] = (np.log(hsg[1:-1]) - sigl[:-1]) * wvi[:-1, 0]
    wvi[-1, 1] = (np.log(dt(0.99)) - sigl[-1]) * wvi[-2, 0]
    return Struct(kx=kx, hsg=hsg, fsg=fsg, dhs=dhs, sigl=sigl, grdsig=grdsig,
                  grdscp=grdscp, wvi=wvi)


def make_geometry(kx, sia, coa, fmask, orog, phis0=None, dtype=np.float32):
    """Geometry with horizontal fields; ``sia``/``coa`` broadcast against H."""
    dt = np.dtype(dtype).type
    g = vertical_geometry(kx, dtype)
    orog = np.asarray(orog, dt)
    if phis0 is None:
        phis0 = dt(grav) * orog
    return g.copy(sia=np.asarray(sia, dt), coa=np.asarray(coa, dt),
                  fmask=np.asarray(fmask, dt), orog=orog, phis0=np.asarray(phis0, dt))


def speedy_latitudes(il, dtype=np.float32):
    """convert_to_speedy_latitudes (test_utils.py:4-16): returns (radang, sia, coa)."""
    dt = np.dtype(dtype).type
    iy = (il + 1) // 2
    j = np.arange(1, iy + 1, dtype=dt)
    sia_half = np.cos(dt(np.pi) * (j - dt(0.25)) / (dt(il) + dt(0.5)))
    radang = np.concatenate((-np.arcsin(sia_half), np.arcsin(sia_half)[::-1]))
    sia = np.concatenate((-sia_half, sia_half[::-1]))
    coa = np.cos(radang)
    return radang.astype(dt), sia.astype(dt), coa.astype(dt)


# --------------------------------------------------------------------------
# PhysicsData.zeros  (physics_data.py:462-472 and the per-struct zeros)
# --------------------------------------------------------------------------
def physics_data_zeros(H, kx, dtype=np.float32, date=None):
    dt = np.dtype(dtype).type
    z2 = lambda *extra: np.zeros(tuple(H) + extra, dt)
    z3 = lambda *extra: np.zeros((kx,) + tuple(H) + extra, dt)
    return Struct(
        shortwave_rad=Struct(qcloud=z2(), fsol=z2(), rsds=z2(), rsns=z2(), ozone=z2(),
                             ozupp=z2(), zenit=z2(), stratz=z2(), gse=z2(),
                             icltop=np.zeros(tuple(H), np.int32), cloudc=z2(), cloudstr=z2(),
                             ftop=z2(), dfabs=z3(), compute_shortwave=False),
        longwave_rad=Struct(dfabs=z3(), ftop=z2()),
        convection=Struct(se=z3(), iptop=np.zeros(tuple(H), np.int32), cbmf=z2(), precnv=z2()),
        mod_radcon=Struct(ablco2=dt(ablco2_ref), alb_l=z2(), alb_s=z2(), albsfc=z2(), snowc=z2(),
                          tau2=z3(4), st4a=z3(2), stratc=z2(2), flux=z2(4)),
        humidity=Struct(rh=z3(), qsat=z3()),
        condensation=Struct(precls=z2(), dtlsc=z3(), dqlsc=z3()),
        surface_flux=Struct(ustr=z2(3), vstr=z2(3), shf=z2(3), evap=z2(3), rlus=z2(3), rlds=z2(),
                            rlns=z2(), hfluxn=z2(2), tsfc=z2(), tskin=z2(), u0=z2(), v0=z2(), t0=z2()),
        date=date if date is not None else Struct(tyear=dt(0.0), model_year=1950, model_step=0,
                                                  dt_seconds=dt(1800.0)),
        land_model=Struct(stl_lm=np.full(tuple(H), 288.0, dt)),
    )


def tendency_zeros(shape, dtype):
    z = lambda: np.zeros(shape, dtype)
    return Struct(u_wind=z(), v_wind=z(), temperature=z(), specific_humidity=z())


def _hshape(state):
    return state.temperature.shape[1:]


def _dt(state):
    return state.temperature.dtype.type


# --------------------------------------------------------------------------
# humidity.py
# --------------------------------------------------------------------------
def get_qsat(ta, ps, sig):
    """get_qsat (humidity.py:73-104); ``sig`` is a scalar or broadcasts against ta."""
    dt = np.asarray(ta).dtype.type
    e0, c1, c2, t0, t1, t2 = (dt(x) for x in (6.108e-3, 17.269, 21.875, 273.16, 35.86, 7.66))
    with np.errstate(all="ignore"):
        qsat = np.where(ta >= t0, e0 * np.exp(c1 * (ta - t0) / (ta - t1)),
                        e0 * np.exp(c2 * (ta - t0) / (ta - t2)))
        sig = np.asarray(sig, dt)
        ps = np.asarray(ps, dt)
        ps00 = ps.reshape(-1)[0] if ps.size else dt(0)
        qsat = np.where(sig <= 0.0, dt(622.0) * qsat / (ps00 - dt(0.378) * qsat),
                        dt(622.0) * qsat / (sig * ps - dt(0.378) * qsat))
    return qsat


def qsat_levels(ta, ps, fsg):
    """jax.vmap(get_qsat, in_axes=(0, None, 0)) over levels (humidity.py:43-44)."""
    return np.stack([get_qsat(ta[k], ps, fsg[k]) for k in range(ta.shape[0])])


def spec_hum_to_rel_hum(state, data, params, forcing, geometry):
    """humidity.py:16-51."""
    qsat = qsat_levels(state.temperature, state.normalized_surface_pressure, geometry.fsg)
    with np.errstate(all="ignore"):
        rh = state.specific_humidity / qsat
    data = data.copy(humidity=data.humidity.copy(rh=rh, qsat=qsat))
    return tendency_zeros(state.temperature.shape, _dt(state)), data


def rel_hum_to_spec_hum(ta, ps, sig, rh):
    """humidity.py:53-71."""
    qsat = get_qsat(ta, ps, sig)
    return rh * qsat, qsat


# --------------------------------------------------------------------------
# convection.py
# --------------------------------------------------------------------------
def diagnose_convection(psa, se, qa, qsat, params, geometry):
    """diagnose_convection (convection.py:15-97).  Returns (iptop, qdif)."""
    dt = se.dtype.type
    kx = se.shape[0]
    H = se.shape[1:]
    nh = len(H)
    pc = params.convection
    mss = se + dt(alhc) * qsat                                      # :50
    rlhc = dt(1.0 / alhc)                                           # :52
    mask_psa = psa > pc.psmin                                       # :56
    mse0 = se[kx - 1] + dt(alhc) * qa[kx - 1]                       # :58
    mse1 = se[kx - 2] + dt(alhc) * qa[kx - 2]
    mse1 = np.minimum(mse0, mse1)
    mss0 = np.maximum(mse0, mss[kx - 1])                            # :63
    mss2 = np.zeros_like(mss)                                       # :65-68
    mss2[:-1] = mss[:-1] + _lev(geometry.wvi[:-1, 1], nh) * np.diff(mss, axis=0)
    levels = np.arange(2, kx - 3)                                   # :73

    def cloud_top(instab):                                          # :74-78
        return np.where(instab.any(axis=0), (levels + 1)[np.argmax(instab, axis=0)], kx)

    ktop1 = cloud_top(mss0 > mss2[2:kx - 3])                        # :81
    ktop2 = cloud_top(mse1 > mss2[2:kx - 3])                        # :84
    msthr = np.take_along_axis(mss2, (ktop2 - 1)[None], axis=0)[0]  # :85
    qthr0 = pc.rhbl * qsat[kx - 1]                                  # :88
    qthr1 = pc.rhbl * qsat[kx - 2]
    lqthr = (qa[kx - 1] > qthr0) & (qa[kx - 2] > qthr1)
    case_1 = mask_psa & (ktop1 < kx) & (ktop2 < kx)                 # :92
    case_2 = mask_psa & (ktop1 < kx) & ~(ktop2 < kx) & lqthr
    iptop = np.where(case_1 | case_2, ktop1, kx + 1).astype(np.int32)
    qdif = np.where(case_1, np.maximum(qa[kx - 1] - qthr0, (mse0 - msthr) * rlhc),
                    np.zeros(H, dt))
    qdif = np.where(case_2, qa[kx - 1] - qthr0, qdif)
    return iptop, qdif.astype(dt)


def get_convection_tendencies(state, data, params, forcing, geometry):
    """get_convection_tendencies (convection.py:99-237)."""
    dt = _dt(state)
    pc = params.convection
    se = dt(cp) * state.temperature + state.geopotential            # :124
    qa = state.specific_humidity
    qsat = data.humidity.qsat
    kx = se.shape[0]
    H = se.shape[1:]
    nh = len(H)
    psa = state.normalized_surface_pressure
    zeros3 = lambda: np.zeros((kx,) + H, dt)
    dfse, dfqa = zeros3(), zeros3()
    with np.errstate(all="ignore"):
        entr = np.maximum(dt(0.0), geometry.fsg[1:kx - 1] - dt(0.5)) ** 2    # :136
        sentr = np.sum(entr)
        entr = entr * (pc.entmax / sentr)
        fqmax = dt(5.0)
        fm0 = dt(p0) * geometry.dhs[-1] / (dt(grav) * pc.trcnv * dt(3600.0))  # :141
        rdps = dt(2.0) / (dt(1.0) - pc.psmin)
        iptop, qdif = diagnose_convection(psa, se, qa, qsat, params, geometry)  # :145
        mask = iptop < kx
        k = kx - 1
        qmax = np.maximum(dt(1.01) * qa[-1], qsat[-1])              # :153
        w = _lev(geometry.wvi[:-1, 1], nh)
        interpolate = lambda tr: tr[:-1] + w * np.diff(tr, axis=0)  # :155
        sb3 = zeros3(); sb3[1:] = interpolate(se)
        qb3 = zeros3(); qb3[1:] = interpolate(qa)
        sb, qb = sb3[k], np.minimum(qb3, qa)[k]                     # :159
        fpsa = psa * np.minimum(dt(1.0), (psa - pc.psmin) * rdps)   # :162
        fmass = fm0 * fpsa * np.minimum(fqmax, qdif / (qmax - qb))
        cbmf = mask * fmass
        fus, fuq = fmass * se[k], fmass * qmax                      # :167
        fds, fdq = fmass * sb, fmass * qb
        dfse[k] = fds - fus
        dfqa[k] = fdq - fuq
        ar = _lev(np.arange(kx), nh)
        loop_mask = (kx - 2 >= ar) & (ar >= iptop)                  # :178
        en = zeros3()
        en[1:-1] = _lev(entr, nh) * psa * cbmf                      # :182
        en = loop_mask * en
        rcum = lambda x: np.cumsum(x[::-1], axis=0)[::-1]
        FM = fmass + rcum(en * dt(1))                               # :185-188
        FUS = fus + rcum(en * se)
        FUQ = fuq + rcum(en * qa)
        FDS = FM * sb3; FDS[-1] = fds                               # :191
        FDQ = FM * qb3; FDQ[-1] = fdq
        dfse[:-1] = loop_mask[:-1] * np.diff(FUS - FDS, axis=0)     # :194
        dfqa[:-1] = loop_mask[:-1] * np.diff(FUQ - FDQ, axis=0)
        delq = loop_mask * (pc.rhil * qsat - qa)                    # :198
        fsq = (delq > 0.0) * pc.smf * cbmf * delq
        dfqa = dfqa + fsq
        dfqa[-1] = dfqa[-1] + (-np.sum(fsq, axis=0))               # :202
        pad2 = lambda a: np.concatenate([a, np.zeros((2,) + a.shape[1:], a.dtype)], axis=0)
        gather = lambda a, idx: np.take_along_axis(a, idx[None], axis=0)[0]
        fmass_t, fus_t, fuq_t, fds_t, fdq_t = (gather(pad2(a), iptop) for a in (FM, FUS, FUQ, FDS, FDQ))
        kt = iptop - 1                                              # :211
        qsatb = gather(pad2(interpolate(qsat)), kt)
        precnv = np.maximum(fuq_t - fmass_t * qsatb, dt(0.0))       # :215
        valid = kt < kx                                             # Q6: JAX drops out-of-range scatter
        kk = np.where(valid, kt, 0)
        for arr, val in ((dfse, fus_t - fds_t + dt(alhc) * precnv), (dfqa, fuq_t - fdq_t - precnv)):
            cur = gather(arr, kk)
            np.put_along_axis(arr, kk[None], np.where(valid, val, cur)[None], axis=0)
        rps = dt(1.0) / psa                                         # :226
        ttend = dfse.copy(); ttend[1:] = dfse[1:] * rps * _lev(geometry.grdscp[1:], nh)
        qtend = dfqa.copy(); qtend[1:] = dfqa[1:] * rps * _lev(geometry.grdsig[1:], nh)
    conv = data.convection.copy(se=se, iptop=iptop.astype(np.int32), cbmf=cbmf.astype(dt),
                                precnv=precnv.astype(dt))
    tend = tendency_zeros(state.temperature.shape, dt)
    tend.temperature, tend.specific_humidity = ttend, qtend
    return tend, data.copy(convection=conv)


# --------------------------------------------------------------------------
# large_scale_condensation.py
# --------------------------------------------------------------------------
def get_large_scale_condensation_tendencies(state, data, params, forcing, geometry):
    """large_scale_condensation.py:14-86."""
    dt = _dt(state)
    pl = params.condensation
    qa = state.specific_humidity
    qsat = data.humidity.qsat
    kx = qa.shape[0]
    nh = qa.ndim - 1
    psa = state.normalized_surface_pressure
    dtlsc = np.zeros_like(qa)
    dqlsc = np.zeros_like(qa)
    with np.errstate(all="ignore"):
        qsmax = dt(10.0)
        rtlsc = dt(1.0) / (pl.trlsc * dt(3600.0))                   # :48
        tfact = dt(alhc / cp)
        prg = dt(p0 / grav)
        psa2 = psa ** 2
        sig2 = geometry.fsg ** 2                                    # :59
        rhref = pl.rhlsc + pl.drhlsc * (sig2 - dt(1.0))
        rhref = np.maximum(rhref, pl.rhblsc)                        # Q10
        dqmax = qsmax * sig2 * rtlsc
        dqa = _lev(rhref, nh) * qsat - qa                           # :66
        neg = dqa < 0
        dqlsc[1:] = np.where(neg[1:], dqa[1:] * rtlsc, dt(0.0))     # :70
        dtlsc[1:] = np.where(neg[1:], tfact * np.minimum(-dqlsc[1:], _lev(dqmax[1:], nh) * psa2), dt(0.0))
        iptop = np.minimum(np.argmin(dqa[1:] >= 0, axis=0) + 1, data.convection.iptop)  # :74 (Q4)
        pfact = geometry.dhs * prg                                  # :77
        precls = dt(0.0) - np.sum(_lev(pfact[1:], nh) * dqlsc[1:], axis=0)
        precls = precls * psa
    data = data.copy(condensation=data.condensation.copy(precls=precls.astype(dt)),
                     convection=data.convection.copy(iptop=iptop.astype(np.int32)))
    tend = tendency_zeros(qa.shape, dt)
    tend.temperature, tend.specific_humidity = dtlsc, dqlsc
    return tend, data


# --------------------------------------------------------------------------
# shortwave_radiation.py
# --------------------------------------------------------------------------
def solar(tyear, csol=4.0 * solc, sia=None, coa=None, dtype=np.float32):
    """solar (shortwave_radiation.py:406-451): daily-mean TOA insolation per latitude."""
    dt = np.dtype(dtype).type
    with np.errstate(all="ignore"):
        asin1 = np.arcsin(dt(1.0))
        pigr = dt(2.0) * asin1                                      # :421
        alpha = dt(2.0) * pigr * dt(tyear)
        ca1 = np.cos(alpha); sa1 = np.sin(alpha)
        ca2 = ca1 ** 2 - sa1 ** 2
        sa2 = dt(2.0) * sa1 * ca1
        ca3 = ca1 * ca2 - sa1 * sa2
        sa3 = sa1 * ca2 + sa2 * ca1
        decl = (dt(0.006918) - dt(0.399912) * ca1 + dt(0.070257) * sa1 - dt(0.006758) * ca2
                + dt(0.000907) * sa2 - dt(0.002697) * ca3 + dt(0.001480) * sa3)   # :432
        fdis = dt(1.000110) + dt(0.034221) * ca1 + dt(0.001280) * sa1 + dt(0.000719) * ca2 + dt(0.000077) * sa2
        cdecl = np.cos(decl); sdecl = np.sin(decl)
        tdecl = sdecl / cdecl
        csolp = dt(csol) / pigr                                     # :443
        # :446 clips to +-(1-1e-9); in the reference's float32 those bounds round to
        # exactly +-1, so both precisions here clip to the float32-rounded bounds.
        lo, hi = dt(np.float32(-1 + epsilon)), dt(np.float32(1 - epsilon))
        ch0 = np.clip(-tdecl * sia / coa, lo, hi)
        h0 = np.arccos(ch0)
        sh0 = np.sin(h0)
        topsr = csolp * fdis * (h0 * sia * sdecl + sh0 * coa * cdecl)  # :450
    return topsr


def zonal_fields(tyear, sia, coa, dtype=np.float32):
    """Per-latitude part of get_zonal_average_fields (shortwave_radiation.py:220-296):
    returns fsol, ozupp, ozone, zenit, stratz on the latitude axis of sia/coa."""
    dt = np.dtype(dtype).type
    with np.errstate(all="ignore"):
        asin1 = np.arcsin(dt(1.0))
        alpha = dt(4.0) * asin1 * (dt(tyear) + dt(10.0 / 365.0))  # :250
        coz1 = np.maximum(dt(0.0), np.cos(alpha - dt(0.0)))
        coz2 = dt(1.8)
        rzen = -np.cos(alpha) * dt(23.45) * asin1 / dt(90.0)       # :259
        topsr = solar(tyear, 4.0 * solc, sia, coa, dtype)
        flat2 = dt(1.5) * sia ** 2 - dt(0.5)                       # :268
        fsol = topsr
        ozupp = dt(0.5 * epssw)
        ozone = dt(0.4 * epssw) * (dt(1.0) + coz1 * sia + coz2 * flat2)
        zenit = dt(1.0) + dt(1.0) * (dt(1.0) - (coa * np.cos(rzen) + sia * np.sin(rzen))) ** 2
        ozupp = fsol * ozupp * zenit                                # :281
        ozone = fsol * ozone * zenit
        stratz = np.maximum(dt(6.0) - fsol, dt(0.0))               # :285
    return fsol, ozupp, ozone, zenit, stratz


def get_zonal_average_fields(state, data, forcing, geometry):
    """get_zonal_average_fields (shortwave_radiation.py:220-296)."""
    dt = _dt(state)
    H = _hshape(state)
    f = zonal_fields(data.date.tyear, geometry.sia, geometry.coa, dt)
    fsol, ozupp, ozone, zenit, stratz = (np.broadcast_to(x, H).astype(dt) for x in f)
    sw = data.shortwave_rad.copy(fsol=fsol, ozupp=ozupp, ozone=ozone, zenit=zenit, stratz=stratz)
    return data.copy(shortwave_rad=sw)


def clouds(state, data, params, forcing, geometry):
    """clouds (shortwave_radiation.py:313-404)."""
    dt = _dt(state)
    ps_ = params.shortwave_radiation
    se = data.convection.se
    phig = state.geopotential
    qa = state.specific_humidity
    rh = data.humidity.rh
    kx = qa.shape[0]
    nh = qa.ndim - 1
    with np.errstate(all="ignore"):
        gse = (se[-2] - se[-1]) / (phig[-2] - phig[-1])            # :338
        nl1 = kx - 2
        nlp = kx
        rrcl = dt(1.0) / (ps_.rhcl2 - ps_.rhcl1)
        mask = rh[nl1] > ps_.rhcl1                                  # :360
        cloudc = np.where(mask, rh[nl1] - ps_.rhcl1, dt(0.0))
        icltop = np.where(mask, nl1, nlp)
        drh = rh[2:kx - 2] - ps_.rhcl1                              # :365
        m = (drh > cloudc) & (qa[2:kx - 2] > ps_.qacl)
        cloudc_update = np.where(m, drh, cloudc)
        cloudc = np.max(cloudc_update, axis=0)
        k_idx = _lev(np.arange(2, kx - 2), nh)
        icltop_update = np.where(m, k_idx, icltop)                  # :372
        icltop = np.where(cloudc == cloudc_update, icltop_update, icltop).max(axis=0)  # Q5
        pr1 = np.minimum(ps_.pmaxcl, dt(86.4) * (data.convection.precnv + data.condensation.precls))
        cloudc = np.minimum(dt(1.0), ps_.wpcl * np.sqrt(np.maximum(dt(epsilon), pr1))
                            + np.minimum(dt(1.0), cloudc * rrcl) ** 2)   # :378
        cloudc = np.where(np.isnan(cloudc), dt(1.0), cloudc)
        icltop = np.minimum(data.convection.iptop, icltop)         # :380
        qcloud = qa[nl1]
        clfact = dt(1.2)
        rgse = dt(1.0) / (ps_.gse_s1 - ps_.gse_s0)
        fstab = np.clip(rgse * (gse - ps_.gse_s0), dt(0.0), dt(1.0))   # :391
        clstr = fstab * np.maximum(ps_.clsmax - clfact * cloudc, dt(0.0))
        clstrl = np.maximum(clstr, ps_.clsminl) * rh[kx - 1]
        clstr = clstr + geometry.fmask * (clstrl - clstr)
    sw = data.shortwave_rad.copy(gse=gse.astype(dt), icltop=icltop.astype(np.int32),
                                 cloudc=cloudc.astype(dt), cloudstr=clstr.astype(dt),
                                 qcloud=qcloud.astype(dt))
    return tendency_zeros(qa.shape, dt), data.copy(shortwave_rad=sw)


def get_clouds(state, data, params, forcing, geometry):
    """get_clouds (shortwave_radiation.py:298-311): tendencies are zero either way."""
    return clouds(state, data, params, forcing, geometry)


def cloud_level_set(icltop, kx):
    """Q1: the set S = {clip(icltop[c], 0, kx-1)} over all columns of a call, as a
    bitmask (bit k set <=> level k in S)."""
    lv = np.unique(np.clip(np.asarray(icltop).reshape(-1), 0, kx - 1))
    return int(sum(1 << int(k) for k in lv))


def shortwave_rad_fluxes(state, data, params, forcing, geometry, level_set=None):
    """shortwave_rad_fluxes (shortwave_radiation.py:27-217).

    ``level_set`` overrides the Q1 cloud-top level set S (bitmask); by default S
    is computed from ``icltop`` over every column of this call, which is what
    the reference's mixed advanced-index scatter produces.
    Returns (tendency, data, level_set).
    """
    dt = _dt(state)
    pr = params.shortwave_radiation
    kx = state.temperature.shape[0]
    H = _hshape(state)
    nh = len(H)
    dhs, fsg = geometry.dhs, geometry.fsg
    psa = state.normalized_surface_pressure
    qa = state.specific_humidity
    sw = data.shortwave_rad
    icltop, cloudc, clstr = sw.icltop, sw.cloudc, sw.cloudstr
    albsfc = data.mod_radcon.albsfc
    nl1 = kx - 1
    fband2 = dt(0.05)
    fband1 = dt(1.0 - 0.05)
    if level_set is None:
        level_set = cloud_level_set(icltop, kx)
    with np.errstate(all="ignore"):
        tau2 = np.zeros((kx,) + H + (4,), dt)
        mask = icltop < kx                                          # :63
        # Q1 (:64-68): every level in S gets albcl*cloudc in every column with icltop<kx
        val = np.where(mask, pr.albcl * cloudc, dt(0.0))
        for k in range(kx):
            if (level_set >> k) & 1:
                tau2[k, ..., 2] = val
        tau2[kx - 1, ..., 2] = pr.albcls * clstr                    # :70
        psaz = psa * sw.zenit                                       # :75
        acloud = cloudc * np.minimum(pr.abscl1 * sw.qcloud, pr.abscl2)
        tau2[0, ..., 0] = np.exp(-psaz * dhs[0] * pr.absdry)        # :80
        abs1 = pr.absdry + pr.absaer * fsg[1:nl1] ** 2
        cloudy = _lev(np.arange(1, nl1), nh) >= icltop
        tau2[1:nl1, ..., 0] = np.exp(-psaz * _lev(dhs[1:nl1], nh) * (
            _lev(abs1, nh) + pr.abswv1 * qa[1:nl1] + cloudy * acloud))   # :85-91
        abs1 = pr.absdry + pr.absaer * fsg[kx - 1] ** 2
        tau2[kx - 1, ..., 0] = np.exp(-psaz * dhs[kx - 1] * (abs1 + pr.abswv1 * qa[kx - 1]))
        tau2[1:kx, ..., 1] = np.exp(-psaz * _lev(dhs[1:kx], nh) * pr.abswv2 * qa[1:kx])  # :95
        # 3. downward flux
        dfabs = np.zeros((kx,) + H, dt)
        ftop = sw.fsol
        flux_1 = np.zeros((kx,) + H, dt)
        flux_2 = np.zeros((kx,) + H, dt)
        flux_1[0] = sw.fsol * fband1                                # :107
        flux_2[0] = sw.fsol * fband2
        dfabs[0] = flux_1[0]                                        # :112
        flux_1[0] = tau2[0, ..., 0] * (flux_1[0] - sw.ozupp * psa)
        dfabs[0] = dfabs[0] + (-flux_1[0])
        flux_1[1] = flux_1[0]                                       # :117
        dfabs[1] = flux_1[1]
        flux_1[1] = tau2[1, ..., 0] * (flux_1[1] - sw.ozone * psa)
        dfabs[1] = dfabs[1] + (-flux_1[1])
        for k in range(2, kx):                                      # :127-134 scan
            flux_1[k] = flux_1[k - 1] * tau2[k, ..., 0] * (dt(1) - tau2[k, ..., 2])
        dfabs[2:kx] = flux_1[1:kx - 1] * (dt(1) - tau2[2:kx, ..., 2]) * (dt(1) - tau2[2:kx, ..., 0])
        tau2[2:kx, ..., 2] = tau2[2:kx, ..., 2] * flux_1[1:kx - 1]  # :138
        flux_2[1] = flux_2[0]                                       # :140
        carry = flux_2[1]
        for k in range(1, kx):                                      # :142-146 scan
            carry = carry * tau2[k, ..., 1]
            flux_2[k] = carry
        dfabs[1:kx] = dfabs[1:kx] + flux_2[:kx - 1] * (dt(1) - tau2[1:kx, ..., 1])
        # 4. upward flux
        rsds = flux_1[kx - 1] + flux_2[kx - 1]                      # :152
        flux_1[kx - 1] = flux_1[kx - 1] * albsfc
        rsns = rsds - flux_1[kx - 1]
        carry = flux_1[-1]
        up = []
        for k in range(kx - 1, 0, -1):                              # :158-164 scan
            carry = carry * tau2[k, ..., 0] + tau2[k, ..., 2]
            up.append(carry)
        flux_1[:-1] = np.stack(up[::-1])
        dfabs = dfabs + flux_1 * (dt(1) - tau2[..., 0])             # :166
        flux_1[1:] = flux_1[:-1].copy()
        flux_1[0] = tau2[0, ..., 0] * flux_1[0] + tau2[0, ..., 2]   # :169
        ftop = ftop - flux_1[0]                                     # :172
        # 5. longwave transmissivity
        ablco2 = data.mod_radcon.ablco2
        ones = np.ones_like(qa)
        absorptivity = np.stack([pr.ablwin * ones, ablco2 * ones, pr.ablwv1 * qa, pr.ablwv2 * qa], axis=-1)
        absorptivity[0, ..., 2:] = 0                                # :188
        acloud1 = cloudc[..., None] * pr.ablcl1
        acloud2 = cloudc[..., None] * pr.ablcl2
        lvl = _lev(np.arange(2, kx - 1), nh)
        absorptivity[2:kx - 1, ..., 0] = absorptivity[2:kx - 1, ..., 0] + np.where(
            lvl > icltop, acloud1[..., 0], acloud2[..., 0])         # :196
        absorptivity[2:kx - 1, ..., 2:] = np.maximum(absorptivity[2:kx - 1, ..., 2:], acloud2[None])
        tau_lw = np.exp(-absorptivity * psa[..., None] * _lev(dhs, nh)[..., None])   # :200
        eps1 = params.mod_radcon.epslw / (dhs[0] + dhs[1])          # :203
        stratc = np.zeros(H + (2,), dt)
        stratc[..., 0] = sw.stratz * psa
        stratc[..., 1] = eps1 * psa
        flux = data.mod_radcon.flux.copy()                          # :208 (Q2)
        flux[..., 0] = flux_1[0]
        flux[..., 1] = flux_2[kx - 1]
        ttend = dfabs * _lev(geometry.grdscp, nh) / psa             # :214
    mod = data.mod_radcon.copy(tau2=tau_lw, stratc=stratc, flux=flux)
    swo = sw.copy(rsns=rsns, ftop=ftop, dfabs=dfabs, rsds=rsds)
    tend = tendency_zeros(state.temperature.shape, dt)
    tend.temperature = ttend
    return tend, data.copy(shortwave_rad=swo, mod_radcon=mod), level_set


def get_shortwave_rad_fluxes(state, data, params, forcing, geometry, level_set=None):
    """get_shortwave_rad_fluxes (shortwave_radiation.py:12-25): the tendency is gated
    by compute_shortwave; the data are always updated (Q7)."""
    tend, data, S = shortwave_rad_fluxes(state, data, params, forcing, geometry, level_set)
    if not bool(data.shortwave_rad.compute_shortwave):
        tend = tendency_zeros(state.temperature.shape, _dt(state))
    return tend, data


# --------------------------------------------------------------------------
# longwave_radiation.py
# --------------------------------------------------------------------------
def radset(temp, epslw):
    """radset (longwave_radiation.py:204-225): 4-band energy fractions, last axis."""
    dt = np.asarray(temp).dtype.type
    jt = np.clip(temp, dt(200), dt(320))
    f1 = dt(0.148) - dt(3.0e-6) * (jt - dt(247)) ** 2
    f2 = dt(0.356) - dt(5.2e-6) * (jt - dt(282)) ** 2
    f3 = dt(0.314) + dt(1.0e-5) * (jt - dt(315)) ** 2
    f0 = dt(1.0) - (f1 + f2 + f3)
    return (dt(1.0) - epslw) * np.stack([f0, f1, f2, f3], axis=-1)


def get_downward_longwave_rad_fluxes(state, data, params, forcing, geometry):
    """longwave_radiation.py:13-117."""
    dt = _dt(state)
    kx = state.temperature.shape[0]
    nh = state.temperature.ndim - 1
    ta = state.temperature
    st4a = data.mod_radcon.st4a.copy()
    tau2 = data.mod_radcon.tau2
    nl1 = kx - 1
    pm = params.mod_radcon
    st4a[:nl1, ..., 0] = ta[:nl1] + _lev(geometry.wvi[:nl1, 1], nh) * (ta[1:nl1 + 1] - ta[:nl1])  # :49
    st4a[0, ..., 1] = dt(0.75) * ta[0] + dt(0.25) * st4a[0, ..., 0]
    st4a[1, ..., 1] = dt(0.50) * ta[1] + dt(0.25) * (st4a[0, ..., 0] + st4a[1, ..., 0])
    anis = dt(1)
    st4a[2:nl1, ..., 1] = dt(0.5) * anis * np.maximum(st4a[2:nl1, ..., 0] - st4a[1:nl1 - 1, ..., 0], dt(0.0))
    st4a[kx - 1, ..., 1] = anis * np.maximum(ta[kx - 1] - st4a[nl1 - 1, ..., 0], dt(0.0))
    st4a[:2, ..., 0] = dt(sbc) * st4a[:2, ..., 1] ** 4                   # :62
    st4a[:2, ..., 1] = 0.0
    st3a = dt(sbc) * ta[2:kx] ** 3
    st4a[2:kx, ..., 0] = st3a * ta[2:kx]
    st4a[2:kx, ..., 1] = dt(4.0) * st3a * st4a[2:kx, ..., 1]
    emis = dt(1) - tau2                                              # :81
    brad = radset(ta, pm.epslw) * (st4a[..., 0, None] + emis * st4a[..., 1, None])
    emis_brad = emis * brad
    flux = emis_brad[0].copy()
    flux[..., 2:] = 0.0
    dfabs = np.zeros(ta.shape, dt)
    dfabs[0] = dfabs[0] + (-np.sum(flux, axis=-1))
    f3d = np.zeros(ta.shape + (4,), dt)
    f3d[0] = flux
    for k in range(1, kx):                                           # :89-93 scan
        flux = tau2[k] * flux + emis_brad[k]
        f3d[k] = flux
    dfabs[1:] = dfabs[1:] + (-np.diff(np.sum(f3d, axis=-1), axis=0))
    rlds = np.sum(pm.emisfc * flux, axis=-1)                         # :98
    corlw = pm.epslw * pm.emisfc * st4a[kx - 1, ..., 0]
    dfabs[-1] = dfabs[-1] + (-corlw)
    rlds = rlds + corlw
    data = data.copy(surface_flux=data.surface_flux.copy(rlds=rlds),
                     longwave_rad=data.longwave_rad.copy(dfabs=dfabs),
                     mod_radcon=data.mod_radcon.copy(st4a=st4a))
    return tendency_zeros(ta.shape, dt), data


def get_upward_longwave_rad_fluxes(state, data, params, forcing, geometry):
    """longwave_radiation.py:119-202."""
    dt = _dt(state)
    kx = state.temperature.shape[0]
    nh = state.temperature.ndim - 1
    ta = state.temperature
    pm = params.mod_radcon
    dfabs = data.longwave_rad.dfabs.copy()
    rlds = data.surface_flux.rlds
    st4a = data.mod_radcon.st4a
    flux = data.mod_radcon.flux
    tau2 = data.mod_radcon.tau2
    stratc = data.mod_radcon.stratc
    rlus = data.surface_flux.rlus[..., 2]
    ts = data.surface_flux.tsfc
    refsfc = dt(1.0) - pm.emisfc
    epslw = pm.epslw
    fsfc = rlus - rlds                                               # :158
    flux = radset(ts, epslw) * rlus[..., None] + refsfc * flux       # :160 (Q2)
    dfabs[-1] = dfabs[-1] + epslw * rlus
    emis = dt(1.0) - tau2
    brad = radset(ta, epslw) * (st4a[..., 0, None] - emis * st4a[..., 1, None])
    emis_brad = emis * brad
    f3d = np.zeros(ta.shape + (4,), dt)
    f3d[-1] = flux
    for k in range(kx - 1, 0, -1):                                   # :171-175 scan
        flux = tau2[k] * flux + emis_brad[k]
        f3d[k - 1] = flux
    flux = f3d[0].copy()
    dfabs[1:] = dfabs[1:] + np.diff(np.sum(f3d, axis=-1), axis=0)   # :178
    flux[..., :2] = (tau2[0] * flux + emis_brad[0])[..., :2]
    dfabs[0] = dfabs[0] + np.sum((f3d[0] - flux)[..., :2], axis=-1)
    corlw1 = geometry.dhs[0] * stratc[..., 1] * st4a[0, ..., 0] + stratc[..., 0]   # :183
    corlw2 = geometry.dhs[1] * stratc[..., 1] * st4a[1, ..., 0]
    dfabs[0] = dfabs[0] + (-corlw1)
    dfabs[1] = dfabs[1] + (-corlw2)
    ftop = corlw1 + corlw2
    ftop = ftop + np.sum(flux, axis=-1)
    with np.errstate(all="ignore"):
        ttend = dfabs * _lev(geometry.grdscp, nh) / state.normalized_surface_pressure  # :199
    data = data.copy(surface_flux=data.surface_flux.copy(rlns=fsfc),
                     longwave_rad=data.longwave_rad.copy(ftop=ftop, dfabs=dfabs),
                     mod_radcon=data.mod_radcon.copy(flux=flux))
    tend = tendency_zeros(ta.shape, dt)
    tend.temperature = ttend
    return tend, data


# --------------------------------------------------------------------------
# surface_flux.py
# --------------------------------------------------------------------------
def get_orog_land_sfc_drag(phis0, hdrag):
    """surface_flux.py:302-315."""
    dt = np.asarray(phis0).dtype.type
    rhdrag = dt(1.0) / (dt(grav) * dt(hdrag))
    return dt(1.0) + rhdrag * (dt(1.0) - np.exp(-np.maximum(phis0, dt(0.0)) * rhdrag))


def get_surface_fluxes(state, data, params, forcing, geometry):
    """get_surface_fluxes (surface_flux.py:15-300)."""
    dt = _dt(state)
    sf = params.surface_flux
    kx = state.temperature.shape[0]
    H = _hshape(state)
    psa = state.normalized_surface_pressure
    ua, va, ta = state.u_wind, state.v_wind, state.temperature
    qa, phi = state.specific_humidity, state.geopotential
    fmask = geometry.fmask
    rsds = data.shortwave_rad.rsds
    rlds = data.surface_flux.rlds
    rh = data.humidity.rh
    phi0 = geometry.orog * dt(grav)                                  # :68
    snowc = data.mod_radcon.snowc
    alb_l = data.mod_radcon.alb_l
    alb_s = data.mod_radcon.alb_s
    stl_am = forcing.stl_am
    sst = forcing.sea_surface_temperature
    esbc = params.mod_radcon.emisfc * dt(sbc)                        # :75
    ghum0 = dt(1.0) - sf.fhum0
    z = lambda n: np.zeros(H + (n,), dt)
    ustr, vstr, shf, evap, rlus = z(3), z(3), z(3), z(3), z(3)
    hfluxn, t1, q1, t2, qsat0 = z(2), z(2), z(2), z(2), z(2)
    denvvs = z(3)
    cdl = sf.cdl * sf.drag_scale if sf.drag_scale != 1 else sf.cdl
    cds = sf.cds * sf.drag_scale if sf.drag_scale != 1 else sf.cds
    u0 = sf.fwind0 * ua[kx - 1]                                      # :90
    v0 = sf.fwind0 * va[kx - 1]
    tskin = np.zeros(H, dt)
    with np.errstate(all="ignore"):
        astab = dt(0.5) if sf.lscasym else dt(1.0)                   # :137
        rdth = sf.fstab / sf.dtheta
        if forcing.lfluxland:                                        # :243 land_fluxes
            rcp = dt(1.0 / cp)
            nl1 = kx - 1
            gtemp0 = dt(1.0) - sf.ftemp0
            dt1 = geometry.wvi[kx - 1, 1] * (ta[kx - 1] - ta[nl1 - 1])   # :106
            t1[..., 0] = t1[..., 0] + (ta[kx - 1] + dt1)
            t1[..., 1] = t1[..., 0] - phi0 * dt1 / (dt(rgas * 288.0) * geometry.sigl[kx - 1])
            t2[..., 1] = ta[kx - 1] + rcp * phi[kx - 1]              # :115
            t2[..., 0] = t2[..., 1] - rcp * phi0
            t1 = np.where((ta[kx - 1] > ta[nl1 - 1])[..., None], sf.ftemp0 * t1 + gtemp0 * t2,
                          ta[kx - 1][..., None])                     # :119
            t0 = t1[..., 1] + fmask * (t1[..., 0] - t1[..., 1])
            denvvs[..., 0] = (dt(p0) * psa / (dt(rgas) * t0)) * np.sqrt(u0 ** 2 + v0 ** 2 + sf.vgust ** 2)
            tskin = stl_am + sf.ctday * np.sqrt(geometry.coa) * rsds * (dt(1.0) - alb_l) * psa   # :132
            dthl = np.where(tskin > t2[..., 0], np.minimum(sf.dtheta, tskin - t2[..., 0]),
                            np.maximum(-sf.dtheta, astab * (tskin - t2[..., 0])))   # :139
            denvvs[..., 1] = denvvs[..., 0] * (dt(1.0) + dthl * rdth)
            forog = get_orog_land_sfc_drag(geometry.phis0, sf.hdrag)  # :148
            cdldv = cdl * denvvs[..., 0] * forog
            ustr[..., 0] = -cdldv * ua[kx - 1]
            vstr[..., 0] = -cdldv * va[kx - 1]
            chlcp = sf.chl * dt(cp)
            shf[..., 0] = chlcp * denvvs[..., 1] * (tskin - t1[..., 0])   # :155

            def evap_q1(idx):                                        # :158-170
                if sf.fhum0 > 0.0:
                    q1_val, qsat0_val = rel_hum_to_spec_hum(t1[..., idx], psa, dt(1.0), rh[kx - 1])
                    q1[..., idx] = sf.fhum0 * q1_val + ghum0 * qa[kx - 1]
                    qsat0[..., idx] = qsat0_val
                else:
                    q1[..., idx] = qa[kx - 1]

            evap_q1(0)
            qsat0[..., 0] = get_qsat(tskin, psa, dt(1.0))            # :172
            evap[..., 0] = sf.chl * denvvs[..., 1] * np.maximum(dt(0.0), forcing.soilw_am * qsat0[..., 0] - q1[..., 0])
            tsk3 = tskin ** 3                                        # :179
            drls = dt(4.0) * esbc * tsk3
            rlus[..., 0] = esbc * tsk3 * tskin
            hfluxn[..., 0] = rsds * (dt(1.0) - alb_l) + rlds - (rlus[..., 0] + shf[..., 0] + (dt(alhc) * evap[..., 0]))
            if sf.lskineb:                                           # :189-221
                clamb = sf.clambda + (snowc * (sf.clambsn - sf.clambda))
                hfluxn[..., 0] = hfluxn[..., 0] - (clamb * (tskin - stl_am))
                dtskin = tskin + dt(1.0)
                qsat0[..., 1] = get_qsat(dtskin, psa, dt(1.0))
                qsat0[..., 1] = np.where(evap[..., 0] > 0.0, forcing.soilw_am * (qsat0[..., 1] - qsat0[..., 0]), dt(0.0))
                dtskin = hfluxn[..., 0] / (clamb + drls + (sf.chl * denvvs[..., 1] * (dt(cp) + (dt(alhc) * qsat0[..., 1]))))
                tskin = tskin + dtskin
                shf[..., 0] = shf[..., 0] + chlcp * denvvs[..., 1] * dtskin
                evap[..., 0] = evap[..., 0] + sf.chl * denvvs[..., 1] * qsat0[..., 1] * dtskin
                rlus[..., 0] = rlus[..., 0] + drls * dtskin
                hfluxn[..., 0] = clamb * (tskin - stl_am)
            dths = np.where(sst > t2[..., 1], np.minimum(sf.dtheta, sst - t2[..., 1]),
                            np.maximum(-sf.dtheta, astab * (sst - t2[..., 1])))   # :223
            denvvs[..., 2] = denvvs[..., 0] * (dt(1.0) + dths * rdth)
            evap_q1(1)
            cdsdv = cds * denvvs[..., 2]                             # :236
            ustr[..., 1] = -cdsdv * ua[kx - 1]
            vstr[..., 1] = -cdsdv * va[kx - 1]
        # sea surface (:250-261)
        shf[..., 1] = sf.chs * dt(cp) * denvvs[..., 2] * (sst - t1[..., 1])
        qsat0[..., 1] = get_qsat(sst, psa, dt(1.0))
        evap[..., 1] = sf.chs * denvvs[..., 2] * (qsat0[..., 1] - q1[..., 1])
        rlus[..., 1] = esbc * (sst ** 4)
        hfluxn[..., 1] = rsds * (dt(1.0) - alb_s) + rlds - rlus[..., 1] + shf[..., 1] + dt(alhc) * evap[..., 1]
        wavg = lambda var: var[..., 1] + fmask * (var[..., 0] - var[..., 1])
        t0 = np.zeros(H, dt)
        tsfc = np.zeros(H, dt)
        if forcing.lfluxland:                                        # :266-286
            for arr in (ustr, vstr, shf, evap, rlus):
                arr[..., 2] = wavg(arr)
            t0 = wavg(t1)
            tsfc = sst + fmask * (stl_am - sst)
            tskin = sst + fmask * (tskin - sst)
        rps = dt(1.0) / psa                                          # :293
        shape3 = state.temperature.shape
        utend = np.zeros(shape3, dt); utend[-1] = utend[-1] + ustr[..., 2] * rps * geometry.grdsig[-1]
        vtend = np.zeros(shape3, dt); vtend[-1] = vtend[-1] + vstr[..., 2] * rps * geometry.grdsig[-1]
        ttend = np.zeros(shape3, dt); ttend[-1] = ttend[-1] + shf[..., 2] * rps * geometry.grdscp[-1]
        qtend = np.zeros(shape3, dt); qtend[-1] = qtend[-1] + evap[..., 2] * rps * geometry.grdsig[-1]
    sfo = data.surface_flux.copy(ustr=ustr, vstr=vstr, shf=shf, evap=evap, rlus=rlus, hfluxn=hfluxn,
                                 tsfc=tsfc, tskin=tskin, u0=u0, v0=v0, t0=t0)
    return Struct(u_wind=utend, v_wind=vtend, temperature=ttend, specific_humidity=qtend), \
        data.copy(surface_flux=sfo)


# --------------------------------------------------------------------------
# vertical_diffusion.py
# --------------------------------------------------------------------------
def get_vertical_diffusion_tend(state, data, params, forcing, geometry):
    """vertical_diffusion.py:10-135."""
    dt = _dt(state)
    pv = params.vertical_diffusion
    se = data.convection.se
    rh = data.humidity.rh
    qsat = data.humidity.qsat
    qa = state.specific_humidity
    phi = state.geopotential
    kx = qa.shape[0]
    nh = qa.ndim - 1
    icnv = kx - data.convection.iptop - 1                            # :39
    ttenvd = np.zeros(qa.shape, dt)
    qtenvd = np.zeros(qa.shape, dt)
    hsg, fsg, dhs = geometry.hsg, geometry.fsg, geometry.dhs
    nl1 = kx - 1
    with np.errstate(all="ignore"):
        cshc = dhs[kx - 1] / dt(3600.0)                              # :45
        cvdi = (hsg[nl1] - hsg[1]) / dt((nl1 - 1) * 3600.0)
        fshcq = cshc / pv.trshc
        fshcse = cshc / (pv.trshc * dt(cp))
        fvdiq = cvdi / pv.trvdi
        fvdise = cvdi / (pv.trvds * dt(cp))
        rsig = dt(1.0) / dhs
        rsig1 = np.zeros(kx, dt)
        rsig1[:-1] = dt(1.0) / (dt(1.0) - hsg[1:-1])
        drh0 = pv.rhgrad * (fsg[kx - 1] - fsg[nl1 - 1])              # :59
        fvdiq2 = fvdiq * hsg[nl1]
        dmse = se[kx - 1] - se[nl1 - 1] + dt(alhc) * (qa[kx - 1] - qsat[nl1 - 1])
        drh = rh[kx - 1] - rh[nl1 - 1]
        fcnv = np.where(np.logical_and(icnv > 0, dmse >= 0), pv.redshc, dt(1.0))   # :70
        fluxse = np.where(dmse >= 0.0, fcnv * fshcse * dmse, dt(0.0))
        ttenvd[nl1 - 1] = np.where(dmse >= 0.0, fluxse * rsig[nl1 - 1], ttenvd[nl1 - 1])
        ttenvd[kx - 1] = np.where(dmse >= 0.0, -fluxse * rsig[kx - 1], ttenvd[kx - 1])
        c1 = (dmse >= 0.0) & (drh >= 0.0)
        fluxq = np.where(c1, fcnv * fshcq * qsat[kx - 1] * drh, dt(0.0))   # :80
        qtenvd[nl1 - 1] = np.where(c1, fluxq * rsig[nl1 - 1], qtenvd[nl1 - 1])
        qtenvd[kx - 1] = np.where(c1, -fluxq * rsig[kx - 1], qtenvd[kx - 1])
        c2 = (dmse < 0.0) & (drh > drh0)
        fluxq = np.where(c2, fvdiq2 * qsat[nl1 - 1] * drh, dt(0.0))  # :88
        qtenvd[nl1 - 1] = np.where(c2, fluxq * rsig[nl1 - 1], qtenvd[nl1 - 1])
        qtenvd[kx - 1] = np.where(c2, -fluxq * rsig[kx - 1], qtenvd[kx - 1])
        kr = np.arange(2, kx - 2)                                    # :99
        cond = hsg[kr + 1] > 0.5
        drh0v = pv.rhgrad * (fsg[kr + 1] - fsg[kr])
        fvdiq2v = fvdiq * hsg[kr + 1]
        drhv = rh[kr + 1] - rh[kr]
        fluxqv = np.where((drhv >= _lev(drh0v, nh)) & _lev(cond, nh), _lev(fvdiq2v, nh) * qsat[kr] * drhv, dt(0.0))
        qtenvd[kr] = qtenvd[kr] + fluxqv * _lev(rsig[kr], nh)        # :117
        qtenvd[kr + 1] = qtenvd[kr + 1] + (-fluxqv * _lev(rsig[kr + 1], nh))
        se0 = se[1:] - pv.segrad * np.diff(phi, axis=0)              # :121
        cnd = se[:nl1] < se0
        fse = np.where(cnd, fvdise * (se0 - se[:nl1]), dt(0.0))
        ttenvd[:nl1] = ttenvd[:nl1] + fse * _lev(rsig[:nl1], nh)
        cum = np.cumsum(fse * _lev(rsig1[:nl1], nh), axis=0)
        ttenvd[1:] = ttenvd[1:] + (-cum)                             # :131
    tend = tendency_zeros(qa.shape, dt)
    tend.temperature, tend.specific_humidity = ttenvd, qtenvd
    return tend, data


# --------------------------------------------------------------------------
# speedy_physics.py / speedy/forcing.py
# --------------------------------------------------------------------------
def set_physics_flags(state, data, params, forcing, geometry):
    """speedy_physics.py:13-32."""
    cs = (int(data.date.model_step) % nstrad) == 0
    data = data.copy(shortwave_rad=data.shortwave_rad.copy(compute_shortwave=cs))
    return tendency_zeros(state.temperature.shape, _dt(state)), data


def set_forcing(state, data, params, forcing, geometry):
    """speedy/forcing.py:11-39 (co2_scale is this build's multiplier; 1.0 = reference)."""
    dt = _dt(state)
    data = get_zonal_average_fields(state, data, forcing, geometry)
    tyear = data.date.tyear
    model_year = data.date.model_year
    pm = params.mod_radcon
    fmask = geometry.fmask
    snowc = np.minimum(dt(1.0), forcing.snowc_am)                    # :27
    alb_l = forcing.alb0 + snowc * (pm.albsn - forcing.alb0)
    alb_s = pm.albsea + forcing.sice_am * (pm.albice - pm.albsea)
    albsfc = alb_s + fmask * (alb_l - alb_s)
    pf = params.forcing
    ablco2 = dt(ablco2_ref) * np.exp(dt(float(pf.increase_co2)) * dt(del_co2)
                                     * (dt(model_year) + dt(tyear) - dt(pf.co2_year_ref)))   # :33
    if pf.co2_scale != 1:
        ablco2 = ablco2 * pf.co2_scale
    mod = data.mod_radcon.copy(snowc=snowc, alb_l=alb_l, alb_s=alb_s, albsfc=albsfc, ablco2=dt(ablco2))
    return tendency_zeros(state.temperature.shape, dt), data.copy(mod_radcon=mod)


TERMS = ("set_physics_flags", "set_forcing", "spec_hum_to_rel_hum", "get_convection_tendencies",
         "get_large_scale_condensation_tendencies", "get_clouds", "get_shortwave_rad_fluxes",
         "get_downward_longwave_rad_fluxes", "get_surface_fluxes", "get_upward_longwave_rad_fluxes",
         "get_vertical_diffusion_tend")   # speedy_physics.py:67-80

ALL_TERMS_MASK = (1 << len(TERMS)) - 1


def day_slice(forcing, tyear):
    """tree_index_3d(forcing, round(tyear*365.2425)) (speedy_physics.py:120-121,
    utils.py:104-108); an out-of-range day is clamped like a JAX gather."""
    day = int(np.round(np.float32(tyear) * np.float32(DAYS_YEAR)))
    out = {}
    for k, v in forcing.__dict__.items():
        if isinstance(v, np.ndarray) and v.ndim > 2:
            out[k] = v[..., min(max(day, 0), v.shape[-1] - 1)]
        else:
            out[k] = v
    return Struct(**out)


def compute_tendencies(state, forcing, geometry, date, params=None, level_set=None,
                       term_mask=ALL_TERMS_MASK, return_level_set=False):
    """SpeedyPhysics.compute_tendencies (speedy_physics.py:88-127).

    ``term_mask`` selects which terms' tendencies are summed (all terms always
    run, so the data they hand on are unchanged); ``level_set`` overrides Q1's S.
    """
    params = _cast_params(params if params is not None else default_parameters(), _dt(state))
    dt = _dt(state)
    H = _hshape(state)
    kx = state.temperature.shape[0]
    data = physics_data_zeros(H, kx, dt, date=date)                  # :109-113
    tend = tendency_zeros(state.temperature.shape, dt)               # :117
    f2d = day_slice(forcing, date.tyear)                             # :120-121
    S_used = None
    g = globals()
    for i, name in enumerate(TERMS):                                 # :123-125
        if name == "get_shortwave_rad_fluxes":
            t, data, S_used = shortwave_rad_fluxes(state, data, params, f2d, geometry, level_set)
            if not bool(data.shortwave_rad.compute_shortwave):
                t = tendency_zeros(state.temperature.shape, dt)
        else:
            t, data = g[name](state, data, params, f2d, geometry)
        if (term_mask >> i) & 1:
            for fld in ("u_wind", "v_wind", "temperature", "specific_humidity"):
                setattr(tend, fld, getattr(tend, fld) + getattr(t, fld))
    if return_level_set:
        return tend, data, S_used
    return tend, data


def verify_state(state):
    """physics_interface.py:305-319: negative specific humidity is clamped to 0."""
    q = state.specific_humidity
    return state.copy(specific_humidity=np.where(q < 0.0, q.dtype.type(0.0), q))


def verify_tendencies(state, tend, time_step):
    """physics_interface.py:321-342: q tendency limited so q stays non-negative."""
    q, dq = state.specific_humidity, tend.specific_humidity
    ts = q.dtype.type(time_step)
    return tend.copy(specific_humidity=np.where(q + ts * dq >= 0, dq, -q / ts))


# How each restated function is pinned (checked in tests/test_oracle_*.py).
PINNED_BY = {
    "get_qsat": "closed form + reference property tests (humidity_test.py:72-148)",
    "diagnose_convection": "Fortran fixtures iptop/qdif.npy (convection_test.py:30-46), isothermal + moist-adiabat columns",
    "get_convection_tendencies": "Fortran fixtures cbmf/precnv/dfse/dfqa.npy (convection_test.py:109-140), moist adiabat (:171-219)",
    "get_large_scale_condensation_tendencies": "realistic column (large_scale_condensation_test.py:45-74)",
    "clouds": "SW composite goldens (shortwave_radiation_test.py:145-225)",
    "solar": "6 x 48 golden insolation values (shortwave_radiation_test.py:23-97)",
    "zonal_fields": "identities (shortwave_radiation_test.py:245-305)",
    "shortwave_rad_fluxes": "SW composite goldens; Q1 heterogeneous cloud tops by brute-force scatter emulation",
    "radset": "closed form (band fractions sum to 1-epslw)",
    "get_downward_longwave_rad_fluxes": "Fortran rlds/dfabs/st4a (longwave_radiation_test.py:44-106)",
    "get_upward_longwave_rad_fluxes": "column values (longwave_radiation_test.py:108-142)",
    "get_surface_fluxes": "point values (surface_flux_test.py:67-106) and max/min/mean triples (:108-354)",
    "get_orog_land_sfc_drag": "surface_flux_test.py:355-364",
    "get_vertical_diffusion_tend": "column values (vertical_diffusion_test.py:25-50)",
    "set_forcing": "albedo limits, snow clamp and the float64 closed form of the CO2 trend (tests/test_oracle_pins.py)",
    "verify_state": "physics_interface_test.py:46-64 properties, exact zeros, idempotence (tests/test_oracle_pins.py)",
    "verify_tendencies": "element-by-element brute force and the non-negativity it guarantees (tests/test_oracle_pins.py)",
    "compute_tendencies": "composition of the pinned terms + conservation identities (column water budget)",
}
