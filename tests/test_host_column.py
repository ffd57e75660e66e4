"""The kernel's column routine (speedy_column.hpp), built for the host, against the
oracle on the same inputs: checks the device code's logic on the CPU.

Tolerance: ||F - F_ref||_inf / ||F_ref||_inf <= 1e-5 per tendency field against
the float32 oracle (2e-5 at 16 levels, where the float32 oracle itself is
1.2e-5 from float64 because of cancellation in the convective flux
differences); discrete decisions (iptop, icltop, level set S) must be equal.
"""
import numpy as np
import pytest

import helpers as H
from jcm_amd import Parameters, synthetic

FIELDS = ("u_wind", "v_wind", "temperature", "specific_humidity")


def wet(case, seed, hi=1.15):
    """RH ~ U(0.5, hi): condensation at many levels -> heterogeneous cloud tops (Q1)."""
    state, forcing, geometry, date = case
    rng = np.random.default_rng(seed)
    rh = rng.uniform(0.5, hi, state.temperature.shape)
    q = rh * synthetic.qsat_np(state.temperature, state.normalized_surface_pressure[None],
                               geometry.fsg[:, None, None].astype(np.float64))
    return (state.copy(specific_humidity=q.astype(np.float32)), forcing, geometry, date)


CASES = {
    "T31_aqua": dict(res=(96, 48), kx=8, land=False, wet=False),
    "T21_land": dict(res=(64, 32), kx=8, land=True, wet=False),
    "T21_wet": dict(res=(64, 32), kx=8, land=True, wet=True),
    "T21_wet_kx16": dict(res=(64, 32), kx=16, land=True, wet=True),
    "T21_wet_kx7": dict(res=(64, 32), kx=7, land=True, wet=True),
    "T21_wet_swoff": dict(res=(64, 32), kx=8, land=True, wet=True, step=1),
    "T21_wet_nofluxland": dict(res=(64, 32), kx=8, land=True, wet=True, lfluxland=False),
}


def make(name, seed=3):
    c = CASES[name]
    case = synthetic.make_case(*c["res"], c["kx"], seed=seed, land=c["land"])
    if c["wet"]:
        case = wet(case, seed + 2)
    st, fo, ge, da = case
    if "step" in c:
        da = da.copy(model_step=c["step"])
    if "lfluxland" in c:
        fo = fo.copy(lfluxland=c["lfluxland"])
    return st, fo, ge, da


@pytest.mark.parametrize("name", list(CASES))
def test_host_column_matches_oracle(name):
    case = make(name)
    p = synthetic.draw_vjp_parameters(2) if CASES[name]["wet"] else Parameters.default()
    outs, diag, S = H.run_host_column(case, p)
    tend, data, S_ref = H.run_oracle(*case, params=p)
    assert S == S_ref
    np.testing.assert_array_equal(diag["iptop"], data.convection.iptop)
    np.testing.assert_array_equal(diag["icltop"], data.shortwave_rad.icltop)
    tol = 2e-5 if CASES[name]["kx"] == 16 else H.REL_TOL
    for f, a in zip(FIELDS, outs):
        assert H.rel_inf(a, getattr(tend, f)) <= tol, f
    for k, ref in (("cbmf", data.convection.cbmf), ("precnv", data.convection.precnv),
                   ("precls", data.condensation.precls), ("cloudc", data.shortwave_rad.cloudc),
                   ("clstr", data.shortwave_rad.cloudstr), ("rsds", data.shortwave_rad.rsds),
                   ("rlds", data.surface_flux.rlds), ("tskin", data.surface_flux.tskin)):
        assert H.rel_inf(diag[k], ref) <= 1e-5, k


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_column_increase_co2(dtype):
    """The CO2 trend of set_forcing (speedy/forcing.py:31-33) in the kernel's
    uniforms: increase_co2 on, model_year 2030, co2_year_ref 1990.  float32 against
    the float32 oracle (1e-5), float64 against the float64 oracle fed the same
    float32 constants (1e-12); the result differs from the trend-free one."""
    st, fo, ge, da = make("T21_land")
    case = (st, fo, ge, da.copy(model_year=2030, tyear=0.625))
    p = Parameters.default().replace(forcing=dict(increase_co2=True, co2_year_ref=1990))
    outs, _, S = H.run_host_column(case, p, dtype=dtype)
    if dtype == np.float32:
        tend, _, S_ref = H.run_oracle(*case, params=p)
        tol = H.REL_TOL
    else:
        tend, _, S_ref = H.run_oracle64_same_constants(case, p)
        tol = F64_TOL
    assert S == S_ref
    for f, a in zip(FIELDS, outs):
        assert H.rel_inf(a, getattr(tend, f)) <= tol, f
    outs0, _, _ = H.run_host_column(case, Parameters.default(), dtype=dtype)
    assert H.rel_inf(outs[2], outs0[2]) > 1e-4


@pytest.mark.parametrize("term", [3, 4, 6, 8, 9, 10])
def test_host_column_single_terms(term):
    """Each term's tendency alone (term_mask) against the oracle's.  A single term's
    field can be small next to its float32 rounding (e.g. shortwave at level 0 is a
    difference of two ~1300 W/m^2 fluxes), so the bound is: within 2e-5 of the float32
    oracle and as close to the float64 oracle as the float32 oracle is (or 1e-5)."""
    case = make("T21_wet")
    p = Parameters.default()
    outs, _, _ = H.run_host_column(case, p, term_mask=1 << term)
    tend, _, _ = H.run_oracle(*case, params=p, term_mask=1 << term)
    t64, _, _ = H.run_oracle(*case, params=p, term_mask=1 << term, dtype=np.float64)
    for f, a in zip(FIELDS, outs):
        ref, ref64 = getattr(tend, f), getattr(t64, f)
        if np.max(np.abs(ref)) == 0:
            assert np.max(np.abs(a)) == 0, f
        else:
            assert H.rel_inf(a, ref) <= 2e-5, f
            assert H.rel_inf(a, ref64) <= max(1e-5, 2 * H.rel_inf(ref, ref64)), f


def test_host_column_level_set_override():
    """With S given, the tendencies equal the oracle's with the same S."""
    case = make("T21_wet")
    p = Parameters.default()
    S = 0b0101010
    outs, _, _ = H.run_host_column(case, p, level_set=S)
    tend, _, S_used = H.run_oracle(*case, params=p, level_set=S)
    assert S_used == S
    assert H.rel_inf(outs[2], tend.temperature) <= H.REL_TOL
    ref_free, _, _ = H.run_oracle(*case, params=p)
    assert H.rel_inf(tend.temperature, ref_free.temperature) > 1e-4   # S matters


F64_TOL = 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_host_column_float64_matches_float64_oracle(name):
    """The same column routine instantiated in double against the float64 oracle
    fed the library's float32 parameters and per-level constants: every expression
    of the device code is pinned to the oracle up to operation-order rounding
    (bound 1e-12 per field, measured ~1e-14), including the discrete decisions."""
    case = make(name)
    p = synthetic.draw_vjp_parameters(2) if CASES[name]["wet"] else Parameters.default()
    outs, diag, S = H.run_host_column(case, p, dtype=np.float64)
    tend, data, S_ref = H.run_oracle64_same_constants(case, p)
    assert S == S_ref
    np.testing.assert_array_equal(diag["iptop"], data.convection.iptop)
    np.testing.assert_array_equal(diag["icltop"], data.shortwave_rad.icltop)
    for f, a in zip(FIELDS, outs):
        e = H.rel_inf(a, getattr(tend, f))
        assert e <= F64_TOL, f"{f}: {e:.2e}"
    for k, ref in (("cbmf", data.convection.cbmf), ("precnv", data.convection.precnv),
                   ("precls", data.condensation.precls), ("cloudc", data.shortwave_rad.cloudc),
                   ("clstr", data.shortwave_rad.cloudstr), ("rsds", data.shortwave_rad.rsds),
                   ("rlds", data.surface_flux.rlds), ("tskin", data.surface_flux.tskin)):
        assert H.rel_inf(diag[k], ref) <= F64_TOL, k


@pytest.mark.parametrize("term", [3, 4, 6, 8, 9, 10])
def test_host_column_float64_single_terms(term):
    case = make("T21_wet")
    p = Parameters.default()
    outs, _, _ = H.run_host_column(case, p, term_mask=1 << term, dtype=np.float64)
    tend, _, _ = H.run_oracle64_same_constants(case, p, term_mask=1 << term)
    for f, a in zip(FIELDS, outs):
        ref = getattr(tend, f)
        if np.max(np.abs(ref)) == 0:
            assert np.max(np.abs(a)) == 0, f
        else:
            assert H.rel_inf(a, ref) <= F64_TOL, f


def _negative_q_case(seed=3, frac=0.05):
    """T21 wet land case with a fraction of negative humidities (as the dynamics can
    leave them) and a long time step, so both verify steps act."""
    st, fo, ge, da = make("T21_wet", seed)
    rng = np.random.default_rng(seed)
    q = st.specific_humidity.copy()
    neg = rng.random(q.shape) < frac
    q[neg] = -np.abs(q[neg]) * np.float32(0.05)
    return st.copy(specific_humidity=q), fo, ge, da.copy(dt_seconds=3600.0)


@pytest.mark.parametrize("dtype,tol", [(np.float32, H.REL_TOL), (np.float64, F64_TOL)])
def test_host_column_fused_verify_steps(dtype, tol):
    """verify_state and verify_tendencies fused into the column routine equal the
    oracle's verify_state -> compute_tendencies -> verify_tendencies sequence."""
    import jcm_amd._native as nat
    case = _negative_q_case()
    p = Parameters.default()
    flags = nat.FLAG_VERIFY_STATE | nat.FLAG_VERIFY_TENDENCIES
    outs, _, S = H.run_host_column(case, p, dtype=dtype, extra_flags=flags)
    tend, _, S_ref = H.run_oracle_verified(*case, params=p, dtype=dtype, same_constants=(dtype == np.float64))
    assert S == S_ref
    q = case[0].specific_humidity
    lim = q + 3600.0 * tend.specific_humidity.astype(np.float64) < 0
    assert (q < 0).any() and np.any(tend.specific_humidity == -q / np.float32(3600.0))
    for f, a in zip(FIELDS, outs):
        ref = getattr(tend, f)
        if dtype == np.float32:
            ref64 = H.run_oracle_verified(*case, params=p, dtype=np.float64)[0]
            ok, e32, e64, eref = H.parity(a, ref, getattr(ref64, f))
            assert ok, f"{f}: {e32:.2e}"
        else:
            assert H.rel_inf(a, ref) <= tol, f
