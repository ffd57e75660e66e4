"""Pins of the oracle functions that the reference's golden vectors leave open.

set_forcing (jcm/physics/speedy/forcing.py:11-39) has no golden vector in the
reference; verify_state / verify_tendencies (jcm/physics_interface.py:305-342)
are tested there only by the properties of physics_interface_test.py:46-64.
Each check below is derived from the mathematics of the expression (special
cases that reduce to one operand, the closed-form CO2 trend evaluated
independently in float64, brute force per element), chosen so that a dropped
term, a wrong sign, a swapped operand or a wrong reference year fails.
"""
import math

import numpy as np
import pytest

from oracle import speedy_oracle as so

F32, F64 = np.float32, np.float64


def _params(dtype, **forcing):
    p = so.default_parameters()
    p.forcing = p.forcing.copy(**forcing)
    return so._cast_params(p, np.dtype(dtype).type)


def _case(dtype, ix=6, il=4, kx=8, fmask=None, alb0=None, snowc_am=None, sice_am=None, tyear=0.3,
          model_year=1950):
    dt = np.dtype(dtype).type
    _, sia, coa = so.speedy_latitudes(il, dtype)
    H = (ix, il)
    fm = np.zeros(H, dt) if fmask is None else np.asarray(fmask, dt)
    geom = so.make_geometry(kx, sia, coa, fm, np.zeros(H, dt), dtype=dtype)
    f = dict(alb0=np.full(H, 0.2, dt), sice_am=np.zeros(H, dt), snowc_am=np.zeros(H, dt),
             soilw_am=np.zeros(H, dt), stl_am=np.full(H, 280.0, dt),
             sea_surface_temperature=np.full(H, 290.0, dt), lfluxland=True)
    for k, v in (("alb0", alb0), ("snowc_am", snowc_am), ("sice_am", sice_am)):
        if v is not None:
            f[k] = np.broadcast_to(np.asarray(v, dt), H).copy()
    forcing = so.Struct(**f)
    z3 = np.zeros((kx,) + H, dt)
    state = so.Struct(u_wind=z3, v_wind=z3, temperature=z3 + dt(280.0), specific_humidity=z3,
                      geopotential=z3, normalized_surface_pressure=np.ones(H, dt))
    date = so.Struct(tyear=dt(tyear), model_year=model_year, model_step=0, dt_seconds=dt(1800.0))
    data = so.physics_data_zeros(H, kx, dtype, date=date)
    return state, data, forcing, geom


@pytest.mark.parametrize("dtype", [F32, F64])
def test_set_forcing_albedo_limits(dtype):
    """forcing.py:26-30.  fmask = 0 gives the sea albedo, fmask = 1 the land albedo;
    snowc_am >= 1 clamps to full snow cover (land albedo = albsn), snowc_am = 0
    leaves alb0; sice_am = 1 gives albice and sice_am = 0 gives albsea."""
    pm = so.default_parameters().mod_radcon
    ix, il = 6, 4
    fmask = np.zeros((ix, il)); fmask[:3] = 1.0
    snow = np.zeros((ix, il)); snow[:, 0] = 0.0; snow[:, 1] = 1.0; snow[:, 2] = 3.5; snow[:, 3] = 0.25
    sice = np.zeros((ix, il)); sice[::2] = 1.0
    alb0 = np.linspace(0.1, 0.4, ix * il).reshape(ix, il)
    st, data, forcing, geom = _case(dtype, ix, il, fmask=fmask, alb0=alb0, snowc_am=snow, sice_am=sice)
    _, out = so.set_forcing(st, data, _params(dtype), forcing, geom)
    m = out.mod_radcon
    tol = dict(rtol=2 * np.finfo(dtype).eps, atol=0)
    # snow cover: min(1, snowc_am)
    np.testing.assert_array_equal(m.snowc, np.minimum(1.0, snow).astype(dtype))
    # land albedo: alb0 without snow, albsn with full snow, linear in between
    np.testing.assert_array_equal(m.alb_l[:, 0], alb0[:, 0].astype(dtype))
    np.testing.assert_allclose(m.alb_l[:, 1], pm.albsn, **tol)
    np.testing.assert_allclose(m.alb_l[:, 2], pm.albsn, **tol)
    np.testing.assert_allclose(m.alb_l[:, 3], 0.75 * alb0[:, 3] + 0.25 * pm.albsn, rtol=4 * np.finfo(dtype).eps)
    # sea albedo: albice under full sea ice, albsea without
    np.testing.assert_allclose(m.alb_s[::2], pm.albice, **tol)
    np.testing.assert_array_equal(m.alb_s[1::2], np.full((ix // 2, il), pm.albsea, dtype))
    # blend by the land-sea mask
    np.testing.assert_allclose(m.albsfc[:3], m.alb_l[:3], **tol)
    np.testing.assert_array_equal(m.albsfc[3:], m.alb_s[3:])


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("model_year,tyear", [(1950, 0.0), (1950, 0.5), (2000, 0.25), (1900, 0.75), (2100, 0.0)])
def test_set_forcing_co2_trend(dtype, model_year, tyear):
    """forcing.py:31-33: ablco2 = ablco2_ref * exp(increase_co2 * del_co2 *
    (model_year + tyear - co2_year_ref)) with ablco2_ref = 6 (physics_data.py:6),
    del_co2 = 0.005 (forcing.py:9), co2_year_ref = 1950 (params.py:41).  The
    expected value is the closed form in Python float64; without the trend the
    absorptivity is 6 in every year."""
    st, data, forcing, geom = _case(dtype, tyear=tyear, model_year=model_year)
    _, out = so.set_forcing(st, data, _params(dtype, increase_co2=True), forcing, geom)
    tyear_d = float(np.dtype(dtype).type(tyear))
    expect = 6.0 * math.exp(0.005 * (model_year + tyear_d - 1950))
    assert abs(float(out.mod_radcon.ablco2) - expect) <= 4 * np.finfo(dtype).eps * expect
    _, out0 = so.set_forcing(st, data, _params(dtype, increase_co2=False), forcing, geom)
    assert float(out0.mod_radcon.ablco2) == 6.0
    # another reference year moves the trend by exactly that many years
    _, out2 = so.set_forcing(st, data, _params(dtype, increase_co2=True, co2_year_ref=2000), forcing, geom)
    expect2 = 6.0 * math.exp(0.005 * (model_year + tyear_d - 2000))
    assert abs(float(out2.mod_radcon.ablco2) - expect2) <= 4 * np.finfo(dtype).eps * expect2


def test_set_forcing_co2_trend_values():
    """Spot values of the trend: 50 years after the reference year the absorptivity
    is 6 e^0.25 = 7.7042; 50 years before, 6 e^-0.25 = 4.6728."""
    for year, val in ((2000, 7.704152), (1900, 4.672804)):
        st, data, forcing, geom = _case(F64, tyear=0.0, model_year=year)
        _, out = so.set_forcing(st, data, _params(F64, increase_co2=True), forcing, geom)
        assert abs(float(out.mod_radcon.ablco2) - val) < 1e-6


def _brute_verify(q, dq, dt):
    out = np.empty_like(dq)
    for i, (qi, di) in enumerate(zip(q.ravel(), dq.ravel())):
        qn = qi + dt * di
        out.ravel()[i] = di if qn >= 0 else -qi / dt
    return out


def test_verify_state_reference_properties():
    """physics_interface_test.py:46-64: q = -1 and q = -1e-5 everywhere become
    non-negative (here exactly 0); non-negative q passes unchanged, bit for bit."""
    kx, ix, il = 8, 96, 48
    for v in (-1.0, -1e-5):
        st = so.Struct(specific_humidity=np.full((kx, il, ix), v, F32))
        out = so.verify_state(st).specific_humidity
        assert np.all(out >= 0) and np.all(out == 0)
    rng = np.random.default_rng(5)
    q = rng.normal(0, 3, (kx, 12, 7)).astype(F32)
    q[0, 0, 0] = -0.0
    out = so.verify_state(so.Struct(specific_humidity=q)).specific_humidity
    np.testing.assert_array_equal(out[q >= 0], q[q >= 0])
    assert np.all(out[q < 0] == 0)
    # idempotent
    np.testing.assert_array_equal(so.verify_state(so.Struct(specific_humidity=out)).specific_humidity, out)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_verify_tendencies_brute_force(dtype):
    """physics_interface.py:321-342 element by element: where q + dt*dq >= 0 the
    tendency passes unchanged, elsewhere it becomes -q/dt, so the next state
    q + dt*dq' is non-negative up to rounding."""
    rng = np.random.default_rng(7)
    kx, H = 8, (10, 6)
    q = np.abs(rng.normal(0, 2, (kx,) + H)).astype(dtype)
    dq = rng.normal(0, 2e-3, (kx,) + H).astype(dtype)
    q[0, 0, :3] = 0.0
    dq[0, 0, :3] = (-1e-6, 0.0, 1e-6)
    time_step = 1800.0
    st = so.Struct(specific_humidity=q)
    tend = so.Struct(specific_humidity=dq)
    out = so.verify_tendencies(st, tend, time_step).specific_humidity
    ref = _brute_verify(q, dq, dtype(time_step))
    np.testing.assert_array_equal(out, ref)
    keep = q + dtype(time_step) * dq >= 0
    assert 0 < keep.sum() < keep.size
    np.testing.assert_array_equal(out[keep], dq[keep])
    qn = q + dtype(time_step) * out
    assert np.all(qn >= -4 * np.finfo(dtype).eps * np.abs(q))
