"""Pins of the transform oracle (oracle/sht_oracle.py) from the mathematics.

Each check is chosen so that a plausible slip in the oracle fails it: a
missing Condon-Shortley sign or normalisation (scipy's lpmv), a transposed
Fourier basis or wrong weight (orthonormality, closed-form coefficients), a
wrong index in the l +- 1 recurrences or a lost radius factor (closed-form
winds of solid-body rotation and of degree-1/2 harmonics, both directions),
and a wrong hydrostatic weight (isothermal closed form and the trapezoid
identity).  The last test restates physics_interface_test.py:9-44.
"""
import math

import numpy as np
import pytest
from scipy.special import lpmv

from oracle import sht_oracle as so


@pytest.fixture(scope="module")
def t31():
    return so.Grid.truncation(31)


def _lonlat(g):
    lon = 2.0 * np.pi * np.arange(g.nlon) / g.nlon
    return lon[:, None], np.arcsin(g.sin_lat)[None, :]


def test_gauss_quadrature_is_exact_to_degree_2n_minus_1(t31):
    x, w = t31.sin_lat, t31.weights / (2.0 * np.pi / t31.nlon)
    for k in range(2 * t31.nlat):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-13


def test_legendre_matches_scipy_lpmv_with_unit_norm():
    x = np.linspace(-0.99, 0.99, 37)
    p = so.legendre_normalized(12, 14, x)
    for m in range(12):
        for l in range(14):
            if l < m:
                assert np.all(p[m, :, l] == 0.0)
                continue
            norm = math.sqrt((2 * l + 1) / 2.0 * math.factorial(l - m) / math.factorial(l + m))
            np.testing.assert_allclose(p[m, :, l], norm * lpmv(m, l, x), rtol=1e-10, atol=1e-12)


def test_bases_are_orthonormal_under_the_quadrature(t31):
    g = t31
    fw = g.f.T @ g.f * (2.0 * np.pi / g.nlon)
    np.testing.assert_allclose(fw, np.eye(fw.shape[0]), atol=1e-12)
    w = g.weights / (2.0 * np.pi / g.nlon)
    for m in (0, 1, 7, g.n_m - 1):
        pm = g.p[m][:, m:]
        np.testing.assert_allclose(pm.T @ (w[:, None] * pm), np.eye(pm.shape[1]), atol=1e-11)


def test_round_trip_of_band_limited_fields(t31):
    g = t31
    rng = np.random.default_rng(0)
    c = rng.standard_normal((3,) + g.modal_shape) * g.mask
    np.testing.assert_allclose(g.to_modal(g.to_nodal(c)), c, atol=1e-11)


def test_closed_form_coefficients(t31):
    g = t31
    lon, lat = _lonlat(g)
    # cos(lat) cos(lon) = -(2 sqrt(pi) / sqrt(3)) * [cos(lon)/sqrt(pi)] * [-(sqrt(3)/2) cos(lat)]
    c = g.to_modal(np.cos(lat) * np.cos(lon))
    expect = np.zeros(g.modal_shape)
    expect[1, 1] = -2.0 * math.sqrt(np.pi) / math.sqrt(3.0)
    np.testing.assert_allclose(c, expect, atol=1e-12)
    # sin(lat) = sqrt(2/3) * sqrt(2 pi) * [1/sqrt(2 pi)] * [sqrt(3/2) sin(lat)]
    c = g.to_modal(np.sin(lat) * np.ones_like(lon))
    expect = np.zeros(g.modal_shape)
    expect[0, 1] = math.sqrt(2.0 / 3.0) * math.sqrt(2.0 * np.pi)
    np.testing.assert_allclose(c, expect, atol=1e-12)


def test_laplacian_of_a_degree_two_harmonic(t31):
    g = t31
    lon, lat = _lonlat(g)
    f = np.sin(lat) * np.cos(lat) * np.sin(lon)
    lap = g.to_nodal(g.laplacian(g.to_modal(f)))
    np.testing.assert_allclose(lap, -6.0 * f / g.radius ** 2, atol=1e-12 / g.radius ** 2)
    np.testing.assert_allclose(g.to_nodal(g.inverse_laplacian(g.laplacian(g.to_modal(f)))), f, atol=1e-12)


def _winds_cases(g):
    lon, lat = _lonlat(g)
    a = g.radius
    one = np.ones_like(lon) * np.ones_like(lat)
    s, c = np.sin(lat), np.cos(lat)
    u0 = 20.0
    # (vorticity, divergence, u, v) in closed form
    return [
        ("solid body", 2.0 * u0 * s / a * one, 0.0 * one, u0 * c * one, 0.0 * one),
        ("chi = cos(lat) cos(lon)", 0.0 * one, -2.0 * c * np.cos(lon) / a ** 2,
         -np.sin(lon) / a * one, -s * np.cos(lon) / a),
        ("chi = sin cos(lat) sin(lon)", 0.0 * one, -6.0 * s * c * np.sin(lon) / a ** 2,
         s * np.cos(lon) / a, np.cos(2 * lat) * np.sin(lon) / a),
        ("psi = sin cos(lat) sin(lon)", -6.0 * s * c * np.sin(lon) / a ** 2, 0.0 * one,
         -np.cos(2 * lat) * np.sin(lon) / a, s * np.cos(lon) / a),
    ]


@pytest.mark.parametrize("case", range(4))
def test_winds_from_vorticity_and_divergence(t31, case):
    g = t31
    name, vor, div, u, v = _winds_cases(g)[case]
    uu, vv = so.vor_div_to_uv_nodal(g, g.to_modal(vor), g.to_modal(div))
    scale = max(np.abs(u).max(), np.abs(v).max())
    np.testing.assert_allclose(uu, u, atol=1e-11 * scale, err_msg=name)
    np.testing.assert_allclose(vv, v, atol=1e-11 * scale, err_msg=name)


@pytest.mark.parametrize("case", range(4))
def test_vorticity_and_divergence_from_winds(t31, case):
    g = t31
    name, vor, div, u, v = _winds_cases(g)[case]
    zm, dm = so.uv_nodal_to_vor_div_modal(g, u, v)
    scale = max(np.abs(vor).max(), np.abs(div).max())
    np.testing.assert_allclose(g.to_nodal(zm), vor, atol=1e-10 * scale, err_msg=name)
    np.testing.assert_allclose(g.to_nodal(dm), div, atol=1e-10 * scale, err_msg=name)


def test_geopotential_isothermal_closed_form_and_trapezoid():
    for kx in (7, 8):
        sig = so.sigma_centers(_sigma(kx))
        w = so.geopotential_weights(sig)
        t0 = 250.0
        np.testing.assert_allclose(w @ np.full(kx, t0), -so.IDEAL_GAS_CONSTANT * t0 * np.log(sig), rtol=1e-13)
        t = np.random.default_rng(kx).uniform(200.0, 300.0, kx)
        phi = w @ t
        # bottom layer: -R T ln(sigma); between layers: R (T_j + T_j+1)/2 ln(s_j+1 / s_j)
        assert abs(phi[-1] + so.IDEAL_GAS_CONSTANT * t[-1] * math.log(sig[-1])) < 1e-9
        for j in range(kx - 1):
            d = so.IDEAL_GAS_CONSTANT * 0.5 * (t[j] + t[j + 1]) * math.log(sig[j + 1] / sig[j])
            assert abs((phi[j] - phi[j + 1]) - d) < 1e-9


def _sigma(kx):
    return {7: [0.0, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
            8: [0.0, 0.05, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0]}[kx]


def _dynamics(g, kx, seed=0):
    rng = np.random.default_rng(seed)
    orog = g.to_modal(np.clip(rng.standard_normal(g.nodal_shape) * 500.0, 0.0, None))
    orog *= g.mask * (np.arange(g.n_l) < 16)          # a smooth, truncated orography
    return so.Dynamics(grid=g, sigma_boundaries=np.asarray(_sigma(kx)),
                       reference_temperature=np.full(kx, 288.0), orography=orog)


def test_reference_round_trip_recovers_temperature(t31):
    """physics_interface_test.py:9-44: constant state (T 288, u 0.5, v -0.5,
    q 0.5, phi 5000, sp 1) through physics -> dynamics -> physics."""
    g, kx = t31, 8
    dyn = _dynamics(g, kx)
    shape = (kx,) + g.nodal_shape
    temp = np.full(shape, 288.0)
    u, v, q = np.full(shape, 0.5), np.full(shape, -0.5), np.full(shape, 0.5)
    sp = np.ones(shape)
    ds = so.physics_state_to_dynamics_state(u, v, temp, q, sp, dyn)
    _, _, t2, q2, phi2, sp2 = so.dynamics_state_to_physics_state(ds, dyn)
    assert np.allclose(temp, t2)
    np.testing.assert_allclose(q2, q, rtol=1e-12)
    np.testing.assert_allclose(sp2, sp, rtol=1e-12)
    # phi from the recovered temperature: isothermal closed form above the orography
    sig = so.sigma_centers(_sigma(kx))
    expect = g.to_nodal(dyn.orography)[None] * so.GRAVITY - so.IDEAL_GAS_CONSTANT * 288.0 * np.log(sig)[:, None, None]
    np.testing.assert_allclose(phi2, expect, rtol=1e-10)


def test_tendency_conversion_units_and_shapes(t31):
    g, kx = t31, 8
    dyn = _dynamics(g, kx)
    rng = np.random.default_rng(3)
    shape = (kx,) + g.nodal_shape
    du, dv, dt, dq = (rng.standard_normal(shape) for _ in range(4))
    out = so.physics_tendency_to_dynamics_tendency(du, dv, dt, dq, dyn)
    assert out.vorticity.shape == (kx,) + g.modal_shape
    assert out.log_surface_pressure.shape == g.modal_shape and not out.log_surface_pressure.any()
    np.testing.assert_allclose(out.tracers["specific_humidity"], g.to_modal(dq) * 1e-3, rtol=1e-12, atol=1e-18)
    assert out.sim_time == 0.0
