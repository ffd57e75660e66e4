"""Pin the CPU oracle against every golden vector the reference's own tests hold.

Each test restates the inputs of one reference test (cited) and checks the
oracle's output against the reference's expected values, at the reference's
tolerance.  Fixture data: tests/golden/convection_f90.npz (from
jcm/data/test/*.npy, made by tests/golden/make_fixtures.py).
"""
import os

import numpy as np
import pytest

from oracle import speedy_oracle as so

HERE = os.path.dirname(os.path.abspath(__file__))
F32 = np.float32


def grid_geometry(ix=96, il=48, kx=8, dtype=F32, fmask=0.0, orog=0.0):
    """Geometry.from_grid_shape + convert_to_speedy_latitudes (test_utils.py:4-16)."""
    dt = np.dtype(dtype).type
    _, sia, coa = so.speedy_latitudes(il, dtype)
    fm = np.full((ix, il), fmask, dt)
    og = np.full((ix, il), orog, dt)
    return so.make_geometry(kx, sia, coa, fm, og, dtype=dtype)


def column_geometry(kx=8, dtype=F32):
    """Geometry.single_column_geometry (geometry.py:207-236)."""
    dt = np.dtype(dtype).type
    z = np.zeros((1, 1), dt)
    return so.make_geometry(kx, np.zeros((1, 1), dt), np.ones((1, 1), dt), z, z, dtype=dtype)


def params(dtype=F32):
    return so._cast_params(so.default_parameters(), np.dtype(dtype).type)


def state_zeros(shape, dtype=F32, **kw):
    dt = np.dtype(dtype).type
    d = dict(u_wind=np.zeros(shape, dt), v_wind=np.zeros(shape, dt), temperature=np.zeros(shape, dt),
             specific_humidity=np.zeros(shape, dt), geopotential=np.zeros(shape, dt),
             normalized_surface_pressure=np.zeros(shape[1:], dt))
    d.update({k: np.asarray(v, dt) for k, v in kw.items()})
    return so.Struct(**d)


def forcing_const(H, dtype=F32, value=0.0, lfluxland=False, **kw):
    dt = np.dtype(dtype).type
    d = {k: np.full(H, value, dt) for k in ("alb0", "sice_am", "snowc_am", "soilw_am", "stl_am",
                                            "sea_surface_temperature")}
    d.update({k: np.asarray(v, dt) for k, v in kw.items()})
    d["lfluxland"] = lfluxland
    return so.Struct(**d)


# ---------------------------------------------------------------- convection
def varying_convection_inputs(dtype=F32):
    """convection_test.py:30-37."""
    dt = np.dtype(dtype).type
    ix, il, kx = 96, 48, 8
    g = grid_geometry(ix, il, kx, dtype)
    fsg = g.fsg
    ps = np.ones((ix, il), dt)
    expo = dt(0.05) * np.cos(dt(3) * np.arange(il, dtype=dt) / dt(il)) ** 3
    ta = dt(300) * np.ones((kx, ix, il), dt) * (fsg[:, None, None] ** expo)
    qsat = so.get_qsat(ta, ps, fsg[:, None, None])
    qa = np.sin(dt(2) * np.arange(ix, dtype=dt)[:, None] / dt(ix)) ** 2 * qsat * dt(3.5)
    phi = dt(so.rgas) * ta * np.log(fsg[:, None, None])
    return g, ps, ta, qsat, qa, phi


@pytest.fixture(scope="module")
def f90():
    return np.load(os.path.join(HERE, "golden", "convection_f90.npz"), allow_pickle=False)


def test_diagnose_convection_varying(f90):
    """convection_test.py:30-46 (atol 1e-4 against F90 iptop/qdif)."""
    g, ps, ta, qsat, qa, phi = varying_convection_inputs()
    se = F32(so.cp) * ta + phi
    iptop, qdif = so.diagnose_convection(ps, se, qa, qsat, params(), g)
    np.testing.assert_array_equal(iptop, f90["iptop"])
    np.testing.assert_allclose(qdif, f90["qdif"], atol=1e-4, rtol=0)


def test_get_convection_tendencies_varying(f90):
    """convection_test.py:109-140."""
    g, ps, ta, qsat, qa, phi = varying_convection_inputs()
    kx, ix, il = ta.shape
    st = state_zeros((kx, ix, il), temperature=ta, geopotential=phi, specific_humidity=qa,
                     normalized_surface_pressure=ps)
    data = so.physics_data_zeros((ix, il), kx)
    data = data.copy(humidity=data.humidity.copy(qsat=qsat))
    tend, out = so.get_convection_tendencies(st, data, params(), forcing_const((ix, il)), g)
    np.testing.assert_array_equal(out.convection.iptop, f90["iptop"])
    np.testing.assert_allclose(out.convection.cbmf, f90["cbmf"], atol=1e-4, rtol=0)
    np.testing.assert_allclose(out.convection.precnv, f90["precnv"], atol=1e-4, rtol=0)
    dfse, dfqa = f90["dfse"].copy(), f90["dfqa"].copy()
    tt = dfse.copy(); tt[1:] = dfse[1:] * (1 / ps) * g.grdscp[1:, None, None]
    qt = dfqa.copy(); qt[1:] = dfqa[1:] * (1 / ps) * g.grdsig[1:, None, None]
    np.testing.assert_allclose(tend.temperature, tt, atol=1e-4, rtol=0)
    np.testing.assert_allclose(tend.specific_humidity, qt, atol=1e-4, rtol=0)


ISO_SE = np.array([594060., 483714.2, 422181.7, 378322.1, 344807.97, 320423.78, 304056.8, 293391.7])
MA_SE = np.array([482562.19904568, 404459.50322158, 364997.46113127, 343674.54474717,
                  328636.42287272, 316973.69544231, 301500., 301500.])
MA_QA = np.array([0., 0.00035438, 0.00347954, 0.00472337, 0.00700214, 0.01416442, 0.01782708, 0.0216505])
MA_QSAT = np.array([0., 0.00037303, 0.00366268, 0.00787228, 0.01167024, 0.01490992, 0.01876534, 0.02279])


def _tile(a, ix=96, il=48):
    return np.tile(np.asarray(a, F32)[:, None, None], (1, ix, il))


def test_diagnose_convection_isothermal():
    """convection_test.py:48-63."""
    g = grid_geometry()
    psa = np.ones((96, 48), F32)
    qsat = so.get_qsat(np.ones((1, 1, 1), F32) * F32(288.), np.ones((1, 1, 1), F32), g.fsg[:, None, None])
    itop, qdif = so.diagnose_convection(psa, _tile(ISO_SE), _tile(np.zeros(8)), np.tile(qsat, (1, 96, 48)),
                                        params(), g)
    assert np.all(itop == 9) and np.all(qdif == 0)


def test_diagnose_convection_moist_adiabat():
    """convection_test.py:89-107."""
    g = grid_geometry()
    psa = np.ones((96, 48), F32)
    itop, qdif = so.diagnose_convection(psa, _tile(MA_SE), _tile(MA_QA) * F32(1000.),
                                        _tile(MA_QSAT) * F32(1000.), params(), g)
    assert itop[0, 0] == 5
    assert abs(float(qdif[0, 0]) - 1.1395) < 5e-5


def test_get_convection_tendencies_isothermal():
    """convection_test.py:142-169."""
    g = grid_geometry()
    kx, ix, il = 8, 96, 48
    qsat = so.get_qsat(np.ones((1, 1, 1), F32) * F32(288.), np.ones((1, 1, 1), F32), g.fsg[:, None, None])
    st = state_zeros((kx, ix, il), temperature=_tile(ISO_SE / so.cp), specific_humidity=_tile(np.zeros(8)),
                     normalized_surface_pressure=np.ones((ix, il)))
    data = so.physics_data_zeros((ix, il), kx)
    data = data.copy(humidity=data.humidity.copy(qsat=np.tile(qsat, (1, ix, il))))
    tend, out = so.get_convection_tendencies(st, data, params(), forcing_const((ix, il)), g)
    assert np.all(out.convection.iptop == 9)
    assert np.allclose(out.convection.cbmf, 0) and np.allclose(out.convection.precnv, 0)
    assert np.allclose(tend.temperature, 0) and np.allclose(tend.specific_humidity, 0)


def test_get_convection_tendencies_moist_adiabat():
    """convection_test.py:171-219."""
    g = grid_geometry()
    kx, ix, il = 8, 96, 48
    st = state_zeros((kx, ix, il), temperature=_tile(MA_SE / so.cp), specific_humidity=_tile(MA_QA) * 1000.,
                     normalized_surface_pressure=np.ones((ix, il)))
    data = so.physics_data_zeros((ix, il), kx)
    data = data.copy(humidity=data.humidity.copy(qsat=_tile(MA_QSAT) * F32(1000.)))
    tend, out = so.get_convection_tendencies(st, data, params(), forcing_const((ix, il)), g)
    assert abs(float(out.convection.cbmf[0, 0]) - 0.019614903) < 5e-5
    assert abs(float(out.convection.precnv[0, 0]) - 0.21752352) < 5e-5
    dfse = np.array([0., 0., 0., 0., -29.774475, 402.0166, 171.78418, 0.])
    dfqa = np.array([0., 0., 0., 0.01235308, 0.07379276, -0.15330768, -0.08423203, -0.05377656])
    tt = dfse.copy(); tt[1:] = dfse[1:] * g.grdscp[1:]
    qt = dfqa.copy(); qt[1:] = dfqa[1:] * g.grdsig[1:]
    for k in (4, 5, 6):   # assertAlmostEqual(places=2)
        assert abs(float(tend.temperature[k, 0, 0]) - tt[k]) < 5e-3
        assert abs(float(tend.specific_humidity[k, 0, 0]) - qt[k]) < 5e-3


# -------------------------------------------------------------- condensation
def test_large_scale_condensation_realistic():
    """large_scale_condensation_test.py:45-74."""
    g = column_geometry()
    kx = 8
    qa = np.array([16.148024, 10.943978, 5.851813, 2.4522789, 0.02198645, 0.16069981, 0., 0.], F32)
    qsat = np.array([1.64229703e-01, 1.69719307e-02, 1.45193088e-01, 1.98833509e+00, 4.58917155e+00,
                     9.24226425e+00, 1.48490220e+01, 2.02474803e+01], F32)
    st = state_zeros((kx, 1, 1), specific_humidity=qa[:, None, None],
                     normalized_surface_pressure=np.full((1, 1), 1.0110737))
    data = so.physics_data_zeros((1, 1), kx)
    data = data.copy(humidity=data.humidity.copy(qsat=qsat[:, None, None].copy()),
                     convection=data.convection.copy(iptop=np.full((1, 1), 4, np.int32)))
    tend, out = so.get_large_scale_condensation_tendencies(st, data, params(), forcing_const((1, 1)), g)
    np.testing.assert_allclose(tend.temperature[:, 0, 0], [0.0, 1.59599063e-05, 7.07364228e-05, 1.45072684e-04,
                                                           0, 0, 0, 0], atol=1e-4, rtol=0)
    np.testing.assert_allclose(tend.specific_humidity[:, 0, 0], [0.0, -7.59054545e-04, -3.98269278e-04,
                                                                 -5.82378946e-05, 0, 0, 0, 0], atol=1e-4, rtol=0)
    assert abs(float(out.condensation.precls[0, 0]) - 1.293) < 0.05
    assert out.convection.iptop[0, 0] == 1   # Q4


# ---------------------------------------------------------- vertical diffusion
def test_vertical_diffusion_column():
    """vertical_diffusion_test.py:25-50 (atol 1e-9)."""
    g = column_geometry()
    kx = 8
    col = lambda a: np.asarray(a, F32)[:, None, None]
    se = col(np.linspace(400, 300, kx))
    rh = col(np.linspace(0.1, 0.9, kx))
    qa = col([1, 4, 7.3, 8.8, 12, 18, 24, 26])
    qsat = col([5, 8, 10, 13, 16, 21, 28, 31])
    phi = col(np.linspace(150000, 0, kx))
    st = state_zeros((kx, 1, 1), specific_humidity=qa, geopotential=phi)
    data = so.physics_data_zeros((1, 1), kx)
    data = data.copy(humidity=data.humidity.copy(rh=rh, qsat=qsat),
                     convection=data.convection.copy(iptop=np.ones((1, 1), np.int32), se=se))
    tend, _ = so.get_vertical_diffusion_tend(st, data, params(), forcing_const((1, 1)), g)
    np.testing.assert_allclose(tend.temperature[:, 0, 0], [2.78098357e-04, 1.39862334e-04, 8.50690617e-05,
                                                           3.73100450e-05, 3.67983799e-06, -2.65383318e-05,
                                                           -6.18272365e-05, -3.07837296e-04], atol=1e-9, rtol=0)
    np.testing.assert_allclose(tend.specific_humidity[:, 0, 0], [0, 0, 0, 0, 9.99411916e-06, 7.24206425e-06,
                                                                 1.30163815e-05, -4.72222083e-05], atol=1e-9, rtol=0)
    assert np.all(tend.u_wind == 0) and np.all(tend.v_wind == 0)


# ---------------------------------------------------------------- longwave
def test_downward_longwave_fortran():
    """longwave_radiation_test.py:8-25,44-106 (Fortran values, atol 1e-4)."""
    ix, il, kx = 96, 48, 8
    g = grid_geometry(ix, il, kx)
    n = ix * il * kx
    step = F32((250.0 - 130.0) / (n - 1))
    ta = F32(130.0) + step * np.arange(n, dtype=F32).reshape((kx, il, ix)).transpose((0, 2, 1))
    st = state_zeros((kx, ix, il), temperature=ta)
    data = so.physics_data_zeros((ix, il), kx)
    _, out = so.get_downward_longwave_rad_fluxes(st, data, params(), forcing_const((ix, il), value=1.0), g)
    f90_rlds = [[186.6984, 187.670515, 188.646319, 189.625957, 190.609469],
                [186.708473, 187.680627, 188.656572, 189.636231, 190.6197],
                [186.718628, 187.69074, 188.666658, 189.646441, 190.630014],
                [186.728719, 187.700953, 188.676876, 189.656632, 190.640263],
                [186.738793, 187.711066, 188.687129, 189.666908, 190.650495]]
    f90_dfabs = [-3.799531, -20.11071, -17.83563, -17.667264, -22.200773, -27.997842, -33.615657, -47.10823]
    f90_st4a = [[[76.56151, 9.97944], [77.0403, 10.02566], [77.5214, 10.07201], [78.0048, 10.11851], [78.49052, 10.16516]],
                [[76.56649, 9.97992], [77.04531, 10.02614], [77.52642, 10.0725], [78.00985, 10.119], [78.4956, 10.16564]],
                [[76.57147, 9.9804], [77.0503, 10.02662], [77.53144, 10.07297], [78.01489, 10.11948], [78.50067, 10.16613]],
                [[76.57644, 9.98088], [77.0553, 10.0271], [77.53647, 10.07346], [78.01994, 10.11996], [78.50574, 10.16662]],
                [[76.58142, 9.98136], [77.0603, 10.02758], [77.54149, 10.07395], [78.02499, 10.12045], [78.51081, 10.1671]]]
    np.testing.assert_allclose(out.surface_flux.rlds[:5, :5], f90_rlds, atol=1e-4, rtol=0)
    np.testing.assert_allclose(out.longwave_rad.dfabs[:, 0, 0], f90_dfabs, atol=1e-4, rtol=0)
    np.testing.assert_allclose(np.mean(out.mod_radcon.st4a[:, :5, :5, :], axis=0), f90_st4a, atol=1e-4, rtol=0)


def test_upward_longwave_column():
    """longwave_radiation_test.py:108-142 (atol 1e-5)."""
    ix, il, kx = 96, 48, 8
    g = grid_geometry(ix, il, kx)
    st = state_zeros((kx, ix, il), temperature=np.full((kx, ix, il), 300.0))
    st.normalized_surface_pressure = np.zeros((il,), F32)   # PhysicsState.zeros((ix, il), kx) quirk
    data = so.physics_data_zeros((ix, il), kx)
    tau2 = (np.ones((kx, ix, il, 4)) + np.arange(kx)[:, None, None, None] * .1).astype(F32)
    data = data.copy(
        longwave_rad=data.longwave_rad.copy(dfabs=np.ones((kx, ix, il), F32)),
        mod_radcon=data.mod_radcon.copy(st4a=np.ones((kx, ix, il, 2), F32), flux=np.ones((ix, il, 4), F32),
                                        tau2=tau2, stratc=np.ones((ix, il, 2), F32)),
        surface_flux=data.surface_flux.copy(rlus=np.zeros((ix, il, 3), F32) + np.array([0, 0, 1], F32),
                                            rlds=np.ones((ix, il), F32), tsfc=np.full((ix, il), 300.0, F32)))
    _, out = so.get_upward_longwave_rad_fluxes(st, data, params(), forcing_const((ix, il)), g)
    assert abs(float(out.surface_flux.rlns[0, 0]) - 0.0) < 1e-5
    assert abs(float(out.longwave_rad.ftop[0, 0]) - (-2.037812334328966)) < 1e-5
    np.testing.assert_allclose(out.longwave_rad.dfabs[:, 0, 0],
                               [-5e-2, 1.2938921, 1.6556535, 1.7784461, 1.7685201, 1.6721002, 1.5397001, 1.4595001],
                               atol=1e-5, rtol=0)
    np.testing.assert_allclose(out.mod_radcon.flux[0, 0, :], [-0.55618826, -0.35694631, -1.20774518, -1.05693259],
                               atol=1e-5, rtol=0)


def test_radset_closed_form():
    """radset (longwave_radiation.py:216-225): fractions sum to 1-epslw; clip at 200/320 K."""
    t = np.array([150., 200., 250., 300., 320., 400.], F32)
    fb = so.radset(t, F32(0.05))
    np.testing.assert_allclose(fb.sum(-1), 0.95, rtol=1e-6)
    np.testing.assert_array_equal(fb[0], fb[1])
    np.testing.assert_array_equal(fb[4], fb[5])
    np.testing.assert_allclose(fb[3, 1], 0.95 * (0.148 - 3e-6 * 53 ** 2), rtol=1e-6)


# -------------------------------------------------------------- shortwave
SOLAR_GOLD = {
    0.2: [59.64891891, 82.51370562, 109.0996075, 135.94454033, 162.48195582, 188.46471746, 213.72891835,
          238.14170523, 261.58627434, 283.95547202, 305.15011948, 325.07762082, 343.65189868, 360.79323687,
          376.42841812, 390.49090207, 402.92092072, 413.66583083, 422.68006932, 429.9254984, 435.37150003,
          438.9950085, 440.78070068, 440.7209988, 438.81611994, 435.07404132, 429.51050427, 422.14893274,
          413.02032164, 402.16320111, 389.62332055, 375.45360549, 359.71400001, 342.47101119, 323.7977572,
          303.77351671, 282.48360014, 260.01911561, 236.4767785, 211.95903738, 186.57407167, 160.43718712,
          133.67240691, 106.41888862, 78.84586166, 51.20481384, 24.06562443, 0.89269878],
    0.4: [0.0, 0.0, 0.0, 0.0, 0.0, 1.17528392e-01, 1.13271540e+01, 2.91320240e+01, 5.00775958e+01,
          7.28770444e+01, 9.68131455e+01, 1.21415906e+02, 1.46344316e+02, 1.71332241e+02, 1.96160737e+02,
          2.20642698e+02, 2.44613680e+02, 2.67926725e+02, 2.90448515e+02, 3.12057588e+02, 3.32642980e+02,
          3.52103122e+02, 3.70345744e+02, 3.87287495e+02, 4.02853935e+02, 4.16979874e+02, 4.29609344e+02,
          4.40696200e+02, 4.50204647e+02, 4.58109880e+02, 4.64399211e+02, 4.69073258e+02, 4.72147819e+02,
          4.73656558e+02, 4.73654825e+02, 4.72225630e+02, 4.69489091e+02, 4.65618250e+02, 4.60867185e+02,
          4.55625373e+02, 4.50536488e+02, 4.46820735e+02, 4.47873663e+02, 4.58140604e+02, 4.66603495e+02,
          4.73109251e+02, 4.77630650e+02, 4.80148724e+02],
    0.6: [0., 0., 0., 0., 2.42301138, 17.44981519, 37.44706963, 59.86771264, 83.6333103, 108.1344301,
          132.97031768, 157.84825598, 182.53801702, 206.84837586, 230.61437093, 253.6899679, 275.94351445,
          297.25534724, 317.5157371, 336.62422101, 354.48898098, 371.02626785, 386.16057506, 399.82446689,
          411.95866549, 422.51235541, 431.44315853, 438.71756928, 444.31126415, 448.20948277, 450.40765545,
          450.9120464, 449.74077685, 446.92519666, 442.51191674, 436.56582757, 429.17485652, 420.45766136,
          410.57670499, 399.7619425, 388.35679371, 376.91876172, 366.48029222, 359.54828853, 363.72218759,
          368.79349031, 372.31796687, 374.28083132],
    0.8: [2.40672590e+02, 2.39410416e+02, 2.37278513e+02, 2.48984331e+02, 2.66799442e+02, 2.86134104e+02,
          3.05646230e+02, 3.24707974e+02, 3.42958056e+02, 3.60158149e+02, 3.76136095e+02, 3.90759256e+02,
          4.03921448e+02, 4.15535691e+02, 4.25530154e+02, 4.33845751e+02, 4.40434599e+02, 4.45259173e+02,
          4.48291587e+02, 4.49513271e+02, 4.48914644e+02, 4.46494901e+02, 4.42261840e+02, 4.36231709e+02,
          4.28429095e+02, 4.18886672e+02, 4.07645224e+02, 3.94753408e+02, 3.80267620e+02, 3.64252011e+02,
          3.46778141e+02, 3.27925150e+02, 3.07779834e+02, 2.86436505e+02, 2.63997727e+02, 2.40574768e+02,
          2.16288991e+02, 1.91274040e+02, 1.65679673e+02, 1.39678886e+02, 1.13480705e+02, 8.73568473e+01,
          6.16981674e+01, 3.71583316e+01, 1.51012308e+01, 1.34429313e-01, 0.0, 0.0],
    1.0: [553.93421795, 551.02918596, 545.81297397, 538.30746507, 528.54406252, 516.56378888, 506.85181087,
          506.40750073, 508.57359122, 511.41450948, 514.02258691, 515.87725366, 516.65036719, 516.12420873,
          514.15095359, 510.63050328, 505.49750198, 498.71321538, 490.2604143, 480.13978746, 468.36747184,
          454.973485, 440.00027589, 423.50189151, 405.54332338, 386.19977815, 365.55671479, 343.709479,
          320.76336998, 296.8341639, 272.04837717, 246.54473496, 220.47604586, 194.01174781, 167.34305754,
          140.69016619, 114.31490876, 88.54240315, 63.80108395, 40.70440853, 20.24490036, 4.43498764, 0., 0.,
          0., 0., 0., 0.],
}
SOLAR_1300_06 = [0., 0., 0., 0., 2.30256929, 16.58242672, 35.58566559, 56.89183219, 79.47609897, 102.75932685,
                 126.36068201, 150.00199764, 173.46448986, 196.56643905, 219.15108349, 241.07964786, 262.22702397,
                 282.47949664, 301.7327911, 319.89143809, 336.86818368, 352.58344167, 366.96545876, 379.95015129,
                 391.48118796, 401.51027927, 409.99715357, 416.90997081, 422.22561652, 425.93006404, 428.01897082,
                 428.49828971, 427.38524116, 424.7096167, 420.51571036, 414.86518702, 407.84160341, 399.55771912,
                 390.16792141, 379.89073483, 369.05250864, 358.18303379, 348.26343559, 341.67600518, 345.64242972,
                 350.46165015, 353.81093342, 355.67622859]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_solar(dtype):
    """shortwave_radiation_test.py:23-97.  The reference checks atol 1e-4; the goldens
    are themselves a float32 computation (the float64 result differs from them by the
    float32 rounding of the values, up to ~4 ulp at 500 W/m^2 = 1.3e-4), so this uses
    atol 1.5e-4.  A dropped or mis-signed term moves values by >1e-2."""
    _, sia, coa = so.speedy_latitudes(48, dtype)
    for tyear, gold in SOLAR_GOLD.items():
        np.testing.assert_allclose(so.solar(tyear, 4 * so.solc, sia, coa, dtype), gold, atol=1.5e-4, rtol=0)
    np.testing.assert_allclose(so.solar(0.6, 1300, sia, coa, dtype), SOLAR_1300_06, atol=1.5e-4, rtol=0)


def test_zonal_identities():
    """shortwave_radiation_test.py:245-305: fsol = solar; polar night; ozone form; non-negativity."""
    _, sia, coa = so.speedy_latitudes(48)
    for tyear in (0.2466, 0.25, 0.6):
        fsol, ozupp, ozone, zenit, stratz = so.zonal_fields(tyear, sia, coa)
        np.testing.assert_array_equal(fsol, so.solar(tyear, 4 * so.solc, sia, coa))
        assert np.all(stratz >= 0) and np.all(np.maximum(6.0 - fsol, 0) == stratz)
        flat2 = 1.5 * sia ** 2 - 0.5
        exp_oz = 0.4 * so.epssw * (1.0 + max(0.0, np.cos(4.0 * np.arcsin(1.0) * (tyear + 10.0 / 365.0))) + 1.8 * flat2)
        np.testing.assert_allclose(ozone[0], fsol[0] * exp_oz[0], atol=1e-4)
        assert np.all(fsol >= 0) and np.all(ozupp >= 0) and np.all(ozone >= 0) and np.all(zenit >= 0)


def sw_composite(dtype=F32):
    """Inputs of shortwave_radiation_test.py:145-187."""
    dt = np.dtype(dtype).type
    ix, il, kx = 96, 48, 8
    qa = dt(0.5) * dt(1000.) * np.array(MA_QA, dt)
    qsat = dt(1000.) * np.array(MA_QSAT, dt)
    with np.errstate(all="ignore"):
        rh = qa / qsat
    geop = np.arange(7, -1, -1, dtype=dt)
    se = dt(.1) * geop
    bc = lambda a: np.tile(np.asarray(a, dt)[:, None, None], (1, ix, il))
    g = grid_geometry(ix, il, kx, dtype, fmask=0.7)
    st = state_zeros((kx, ix, il), dtype, specific_humidity=bc(qa), geopotential=bc(geop),
                     normalized_surface_pressure=np.ones((ix, il)))
    data = so.physics_data_zeros((ix, il), kx, dtype)
    data = data.copy(humidity=data.humidity.copy(rh=bc(rh), qsat=bc(qsat)),
                     convection=data.convection.copy(iptop=np.full((ix, il), 8, np.int32),
                                                     precnv=np.full((ix, il), -1.0, dt), se=bc(se)),
                     condensation=data.condensation.copy(precls=np.full((ix, il), 4.0, dt)),
                     shortwave_rad=data.shortwave_rad.copy(compute_shortwave=True),
                     date=so.Struct(tyear=dt(0.6), model_year=1950, model_step=0))
    return st, data, g


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_shortwave_composite(dtype):
    """shortwave_radiation_test.py:145-225 (clouds -> zonal fields -> SW; atol 1e-4)."""
    st, data, g = sw_composite(dtype)
    p = params(dtype)
    f = forcing_const((96, 48), dtype)
    _, data = so.get_clouds(st, data, p, f, g)
    data = so.get_zonal_average_fields(st, data, f, g)
    _, data = so.get_shortwave_rad_fluxes(st, data, p, f, g)
    rsds = [0., 0., 0., 0., 1.08102491, 7.9856262, 17.54767508, 28.67351887, 40.8631746, 53.79605732,
            67.22801389, 80.95422179, 94.79448489, 108.58701854, 122.18603817, 135.46087123, 148.29548103,
            160.58828119, 172.25138545, 183.21006299, 193.40177528, 202.77492961, 211.28786499, 218.90753726,
            225.60832278, 231.37096758, 236.18142325, 240.03003908, 242.91081466, 244.82085417, 245.76014002,
            245.7315415, 244.74127921, 242.79984604, 239.92358203, 236.13704304, 231.47654032, 225.99538369,
            219.77196135, 212.92314683, 205.62864786, 198.18231101, 191.12290959, 185.73622544, 185.85603776,
            186.12903619, 185.31120169, 183.42677496]
    ftop = [0., 0., 0., 0., 1.93599586, 13.84635135, 29.51685016, 46.89146027, 65.11718871, 83.73023451,
            102.44168978, 121.0533787, 139.41874296, 157.42199198, 174.96630874, 191.96679879, 208.34607385,
            224.03188495, 238.95538028, 253.05068109, 266.25471774, 278.50725748, 289.75158834, 299.9350031,
            309.0094909, 316.93257921, 323.66789347, 329.1860073, 333.46512965, 336.4917604, 338.26143437,
            338.77936105, 338.06122342, 336.13410495, 333.03776805, 328.82654748, 323.57244264, 317.37036698,
            310.3474896, 302.680774, 294.63184471, 286.62427534, 279.44504545, 275.01597523, 279.19413113,
            284.20954594, 288.1834261, 291.08877534]
    dfabs_lat = [0., 0., 0., 0., 0.10687137, 0.73259064, 1.49614688, 2.27724268, 3.03175176, 3.74177215,
                 4.40170949, 5.01239461, 5.57803226, 6.10437168, 6.59753382, 7.06324094, 7.5063241, 7.93045047,
                 8.33799935, 8.73007726, 9.10661781, 9.46654098, 9.80796542, 10.12843323, 10.42514601, 10.69520145,
                 10.93580878, 11.14449603, 11.31928937, 11.45886328, 11.56266179, 11.63097744, 11.66499303,
                 11.66678236, 11.63927325, 11.58618806, 11.51198779, 11.42187291, 11.32194103, 11.2197034,
                 11.12539961, 11.05524554, 11.04026698, 11.15996872, 11.66726167, 12.26006372, 12.85902805,
                 13.45775005]
    dfabs_lev = [3.82887045, 7.81598669, 14.17718547, 5.65627818, 7.80939064, 12.48949685, 8.5056334, 5.21519786]
    sw = data.shortwave_rad
    np.testing.assert_allclose(sw.rsds[0, :], rsds, atol=1e-4, rtol=0)
    np.testing.assert_allclose(sw.rsns[0, :], rsds, atol=1e-4, rtol=0)
    np.testing.assert_allclose(sw.ftop[0, :], ftop, atol=1e-4, rtol=0)
    np.testing.assert_allclose(np.mean(sw.dfabs, axis=0)[0, :], dfabs_lat, atol=1e-4, rtol=0)
    np.testing.assert_allclose(np.mean(sw.dfabs, axis=2)[:, 0], dfabs_lev, atol=1e-4, rtol=0)


def test_q1_scatter_matches_advanced_indexing():
    """Q1: the bitmask level set reproduces the reference's mixed advanced-index
    scatter (shortwave_radiation.py:62-70), emulated here with NumPy's own
    advanced-indexing rules (which JAX follows) on heterogeneous cloud tops."""
    rng = np.random.default_rng(7)
    kx, ix, il = 8, 5, 4
    for trial in range(20):
        icltop = rng.integers(1, kx + 2, size=(ix, il)).astype(np.int32)
        cloudc = rng.uniform(0, 1, (ix, il)).astype(F32)
        albcl, albcls = F32(0.43), F32(0.5)
        clstr = rng.uniform(0, 1, (ix, il)).astype(F32)
        tau2 = np.zeros((kx, ix, il, 4), F32)
        mask = icltop < kx
        cl = np.clip(icltop, 0, kx - 1)
        tau2[cl, :, :, 2] = np.where(mask, albcl * cloudc, tau2[cl, :, :, 2])   # literal reference form
        tau2[kx - 1, :, :, 2] = albcls * clstr
        S = so.cloud_level_set(icltop, kx)
        mine = np.zeros((kx, ix, il), F32)
        for k in range(kx):
            if (S >> k) & 1:
                mine[k] = np.where(mask, albcl * cloudc, 0)
        mine[kx - 1] = albcls * clstr
        np.testing.assert_array_equal(mine, tau2[..., 2])


# ----------------------------------------------------------- surface fluxes
def sflux_case(ta, phi0, ua=1.0, qa=5., rh=0.8, phi=5000., sst=290., forcing_ones=False):
    ix, il, kx = 96, 48, 8
    z, xy = (kx, ix, il), (ix, il)
    g = grid_geometry(ix, il, kx, fmask=0.5, orog=phi0 / so.grav)
    if callable(phi):
        phi_a = phi()
    else:
        phi_a = np.full(z, phi, F32)
    st = state_zeros(z, u_wind=np.full(z, ua), v_wind=np.full(z, ua), temperature=np.full(z, ta),
                     specific_humidity=np.full(z, qa), geopotential=phi_a,
                     normalized_surface_pressure=np.ones(xy))
    data = so.physics_data_zeros(xy, kx)
    data = data.copy(surface_flux=data.surface_flux.copy(rlds=np.full(xy, 400., F32)),
                     humidity=data.humidity.copy(rh=np.full(z, rh, F32)),
                     shortwave_rad=data.shortwave_rad.copy(rsds=np.full(xy, 400., F32)))
    f = forcing_const(xy, value=1.0 if forcing_ones else 0.0, lfluxland=True,
                      sea_surface_temperature=np.full(xy, sst), soilw_am=np.full(xy, 0.5),
                      stl_am=np.full(xy, 288.))
    return so.get_surface_fluxes(st, data, params(), f, g)


def test_surface_flux_updated_points():
    """surface_flux_test.py:67-106."""
    phi = lambda: (np.ones((8, 96, 48), F32) * np.arange(8)[::-1][:, None, None]).astype(F32)
    _, d = sflux_case(290., 0.0, qa=1.0, rh=0.5, phi=phi, sst=292., forcing_ones=True)
    s = d.surface_flux
    np.testing.assert_allclose(s.ustr[0, 0], [-0.01493673, -0.00900353, -0.01197013], atol=1e-4)
    np.testing.assert_allclose(s.vstr[0, 0], [-0.01493673, -0.00900353, -0.01197013], atol=1e-4)
    np.testing.assert_allclose(s.shf[0, 0], [81.73508, 16.271175, 49.003124], atol=1e-4)
    np.testing.assert_allclose(s.evap[0, 0], [0.06291558, 0.10244954, 0.08268256], atol=1e-4)
    np.testing.assert_allclose(s.rlus[0, 0], [459.7182, 403.96204, 431.84012], atol=1e-4)
    np.testing.assert_allclose(s.hfluxn[0, 0], [101.19495, 668.53546], atol=1e-4)
    assert abs(float(s.tsfc[0, 0]) - 290.0) < 1e-4
    assert abs(float(s.tskin[0, 0]) - 297.22821044921875) < 1e-4
    assert abs(float(s.u0[0, 0]) - 0.949999988079071) < 1e-4
    assert abs(float(s.t0[0, 0]) - 290.0) < 1e-4


SFLUX_TRIPLES = {   # surface_flux_test.py:108-354: (ta, phi0) -> [max, min, mean] rows
    (288., 500.): [[-4.18139994e-03, -4.18139994e-03, 1.08220810e+02, 4.80042472e-02, 4.87866394e+02, 4.80595490e+02,
                    2.89e+02, 2.98854797e+02, 9.49999988e-01, 9.49999988e-01, 2.88e+02],
                   [-1.50404554e-02, -1.50404554e-02, 7.55662489e+00, 2.64080837e-02, 3.93007751e+02, 1.06054558e+02,
                    2.89e+02, 2.96575317e+02, 9.49999988e-01, 9.49999988e-01, 2.88e+02],
                   [-9.61105898e-03, -9.61105898e-03, 5.54379463e+01, 3.52742635e-02, 4.32339783e+02, 2.97601044e+02,
                    2.89e+02, 2.97186432e+02, 9.50001001e-01, 9.50001001e-01, 2.88e+02]],
    (288., -10.): [[-4.18139994e-03, -4.18139994e-03, 1.05182373e+02, 4.66440842e-02, 4.92244934e+02, 4.80595490e+02,
                    2.89e+02, 2.99263367e+02, 9.49999988e-01, 9.49999988e-01, 2.88e+02],
                   [-1.50404554e-02, -1.50404554e-02, 7.55662489e+00, 2.64080837e-02, 3.93007751e+02, 1.09651382e+02,
                    2.89e+02, 2.96832245e+02, 9.49999988e-01, 9.49999988e-01, 2.88e+02],
                   [-9.61105898e-03, -9.61105898e-03, 5.36600761e+01, 3.45076099e-02, 4.33961243e+02, 2.99674957e+02,
                    2.89e+02, 2.97482452e+02, 9.50001001e-01, 9.50001001e-01, 2.88e+02]],
    (300., 500.): [[-1.98534015e-03, -1.98534015e-03, 3.40381584e+01, 2.68966686e-02, 5.22806824e+02, 4.20411835e+02,
                    2.89e+02, 3.02115173e+02, 9.49999988e-01, 9.49999988e-01, 3.0e+02],
                   [-1.44388555e-02, -1.44388555e-02, -1.79395313e+01, 1.25386305e-02, 3.93007751e+02, 1.78033081e+02,
                    2.89e+02, 3.01716644e+02, 9.49999988e-01, 9.49999988e-01, 3.0e+02],
                   [-8.21213890e-03, -8.21213890e-03, 7.37131262e+00, 1.92672648e-02, 4.57831177e+02, 3.00025909e+02,
                    2.89e+02, 3.01831757e+02, 9.50001001e-01, 9.50001001e-01, 3.0e+02]],
    (285., 500.): [[-6.3609974e-03, -6.3609974e-03, 1.5656566e+02, 5.3803049e-02, 4.6281613e+02, 5.3620532e+02,
                    2.89e+02, 2.9651727e+02, 9.4999999e-01, 9.4999999e-01, 2.85e+02],
                   [-1.5198796e-02, -1.5198796e-02, 2.8738983e+01, 4.0173572e-02, 3.9300775e+02, 7.0954407e+01,
                    2.89e+02, 2.9406818e+02, 9.4999999e-01, 9.4999999e-01, 2.85e+02],
                   [-1.0780082e-02, -1.0780082e-02, 8.8576477e+01, 4.5306068e-02, 4.1897797e+02, 3.0835028e+02,
                    2.89e+02, 2.9474951e+02, 9.5000100e-01, 9.5000100e-01, 2.85e+02]],
}


@pytest.mark.parametrize("case", list(SFLUX_TRIPLES))
def test_surface_flux_triples(case):
    """surface_flux_test.py:108-354: [max, min, mean] per field at rtol 2e-5."""
    ta, phi0 = case
    _, d = sflux_case(ta, phi0)
    s = d.surface_flux
    vars_ = [s.ustr, s.vstr, s.shf, s.evap, s.rlus, s.hfluxn, s.tsfc, s.tskin, s.u0, s.v0, s.t0]
    got = np.array([[np.max(v), np.min(v), np.mean(v, dtype=np.float64)] for v in vars_])
    np.testing.assert_allclose(got, np.array(SFLUX_TRIPLES[case]).T, rtol=2e-5)


def test_orog_drag():
    """surface_flux_test.py:355-364."""
    f = so.get_orog_land_sfc_drag(np.full((96, 48), 500., F32), F32(2000.))
    assert abs(float(np.max(f)) - 1.0000012824780082) < 1e-7
    assert abs(float(np.min(f)) - 1.0000012824780082) < 1e-7


# ------------------------------------------------------------------ humidity
def test_qsat_properties():
    """humidity_test.py:72-148."""
    g = grid_geometry()
    ps = np.ones((96, 48), F32)
    for t in (273., 100., 350.):
        assert np.all(so.get_qsat(np.full((96, 48), t, F32), ps, F32(4)) >= 0)
    kx = 8
    st = state_zeros((kx, 96, 48), temperature=np.full((kx, 96, 48), 273.), specific_humidity=np.zeros((kx, 96, 48)),
                     normalized_surface_pressure=ps)
    _, d = so.spec_hum_to_rel_hum(st, so.physics_data_zeros((96, 48), kx), params(), None, g)
    assert np.all(d.humidity.rh == 0)
    st.temperature = np.full((kx, 96, 48), 400., F32)
    _, d = so.spec_hum_to_rel_hum(st, d, params(), None, g)
    assert np.all((d.humidity.rh >= 0) & (d.humidity.rh <= 1))
    st.temperature = np.full((kx, 96, 48), 273., F32)
    st.specific_humidity = np.full((kx, 96, 48), 2., F32)
    _, d = so.spec_hum_to_rel_hum(st, d, params(), None, g)
    qa, _ = so.rel_hum_to_spec_hum(st.temperature[0], ps, g.fsg[0], d.humidity.rh[0])
    np.testing.assert_allclose(qa, 2.0, atol=1e-6)
    # closed form (humidity.py:97-102) at one point, in float64
    t = 300.0
    e = 6.108e-3 * np.exp(17.269 * (t - 273.16) / (t - 35.86))
    assert abs(float(so.get_qsat(np.float64(t), np.float64(1.0), 0.5)) - 622 * e / (0.5 - 0.378 * e)) < 1e-12
