"""CPU oracle (test infrastructure only) for the dynamics <-> physics transforms.

Float64 NumPy restatement of what jcm/physics_interface.py:196-303 computes:

  dynamics_state_to_physics_state        physics_interface.py:196-234
  physics_state_to_dynamics_state        physics_interface.py:237-273
  physics_tendency_to_dynamics_tendency  physics_interface.py:275-303

Those functions call into dinosaur (requirements.txt: ``dinosaur``, no version
pinned; not present in the reference tree): ``Grid.T<N>`` with
``RealSphericalHarmonics`` (utils.py:28-54), ``vor_div_to_uv_nodal``,
``uv_nodal_to_vor_div_modal``, ``compute_diagnostic_state`` and
``get_geopotential``.  This module restates dinosaur's published conventions
for those:

* nodes: longitudes 2*pi*i/nlon; latitudes arcsin of the Gauss-Legendre nodes,
  south to north; quadrature weights w_j * 2*pi/nlon;
* modal array [..., 2*M-1, L] (M longitude wavenumbers, L total wavenumbers),
  mode index 0 -> m=0, 2m-1 -> cos(m lon), 2m -> sin(m lon);
* Fourier basis: 1/sqrt(2 pi), cos(m lon)/sqrt(pi), sin(m lon)/sqrt(pi);
* Legendre basis: associated Legendre functions normalised to unit L2 norm on
  [-1, 1], Condon-Shortley phase included, zero for l < m;
* to_modal = Legendre(weights * Fourier^T x); to_nodal = Fourier(Legendre^T x);
* Laplacian eigenvalues -l(l+1)/a^2 (inverse 0 at l=0); the gradient of a
  modal field is (d/dlon, cos(lat) d/dlat)/a with the highest total
  wavenumber clipped; u cos(lat) = (chi_lon - cos d/dlat psi)/a,
  v cos(lat) = (psi_lon + cos d/dlat chi)/a; vorticity / divergence of
  (u, v)/cos(lat) in modal form through sec(lat) d/dlat cos^2(lat), clipped;
* geopotential: surface g*orography plus the hydrostatic trapezoid on
  log(sigma) (the lowest layer integrates from sigma=1 with its own
  temperature);
* SI units (SI_SCALE): specific humidity is g/kg on the physics side and kg/kg
  on the dynamics side.

Every function is pinned from the mathematics in tests/test_sht_oracle.py:
Gauss quadrature exactness, scipy's associated Legendre functions, basis
orthonormality, closed-form winds of solid-body rotation and of low-degree
harmonics, isothermal and linear-in-log(sigma) geopotentials, and the round
trip of physics_interface_test.py:9-44.  Parity unpinned: ``clip`` (the
highest total wavenumber zeroed after each gradient, curl and divergence) is
dinosaur's truncation convention, restated; the closed-form fields have no
energy there, so no test can tell it from its absence.  Only tests/, smoke()
and bench.py's CPU baseline use this module; the library never does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

GRAVITY = 9.80616                     # dinosaur scales.GRAVITY_ACCELERATION (m s-2)
IDEAL_GAS_CONSTANT = 1004.0 * 2.0 / 7.0   # ISOBARIC_HEAT_CAPACITY * KAPPA (J kg-1 K-1)
RADIUS = 6.37122e6                    # scales.RADIUS (m)

# dinosaur Grid.T<N>: (longitude wavenumbers, total wavenumbers, nlon, nlat)
TRUNCATIONS = {21: (22, 23, 64, 32), 31: (32, 33, 96, 48), 42: (43, 44, 128, 64),
               85: (86, 87, 256, 128), 106: (107, 108, 320, 160), 119: (120, 121, 360, 180),
               170: (171, 172, 512, 256), 213: (214, 215, 640, 320), 340: (341, 342, 1024, 512),
               425: (426, 427, 1280, 640)}


def legendre_normalized(n_m: int, n_l: int, x: np.ndarray) -> np.ndarray:
    """p[m, j, l]: associated Legendre functions of degree l and order m at x_j,
    unit L2 norm on [-1, 1], Condon-Shortley phase, zero for l < m.
    Standard three-term recurrence in l (the diagonal from the sectoral seed)."""
    x = np.asarray(x, np.float64)
    y = np.sqrt(1.0 - x * x)
    p = np.zeros((n_m, x.size, n_l))
    diag = np.full(x.size, 1.0 / math.sqrt(2.0))
    for m in range(n_m):
        if m > 0:
            diag = -math.sqrt(1.0 + 1.0 / (2.0 * m)) * y * diag
        if m >= n_l:
            break
        p[m, :, m] = diag
        if m + 1 < n_l:
            p[m, :, m + 1] = math.sqrt(2.0 * m + 3.0) * x * diag
        for l in range(m + 2, n_l):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            p[m, :, l] = a * (x * p[m, :, l - 1] - b * p[m, :, l - 2])
    return p


def fourier_real_basis(n_m: int, nlon: int) -> np.ndarray:
    """f[i, mode]: 1/sqrt(2 pi), then cos(m lon_i)/sqrt(pi), sin(m lon_i)/sqrt(pi)."""
    lon = 2.0 * np.pi * np.arange(nlon) / nlon
    f = np.empty((nlon, 2 * n_m - 1))
    f[:, 0] = 1.0 / math.sqrt(2.0 * np.pi)
    for m in range(1, n_m):
        f[:, 2 * m - 1] = np.cos(m * lon) / math.sqrt(np.pi)
        f[:, 2 * m] = np.sin(m * lon) / math.sqrt(np.pi)
    return f


@dataclass
class Grid:
    """A dinosaur-style spectral grid (Grid.T<N>, RealSphericalHarmonics)."""
    n_m: int
    n_l: int
    nlon: int
    nlat: int
    radius: float = RADIUS

    @classmethod
    def truncation(cls, n: int, radius: float = RADIUS) -> "Grid":
        return cls(*TRUNCATIONS[n], radius=radius)

    def __post_init__(self):
        x, w = np.polynomial.legendre.leggauss(self.nlat)
        self.sin_lat = x
        self.cos_lat = np.sqrt(1.0 - x * x)
        self.weights = w * (2.0 * np.pi / self.nlon)
        self.f = fourier_real_basis(self.n_m, self.nlon)
        self.p = legendre_normalized(self.n_m, self.n_l, x)
        self.mode_m = np.array([0] + [m for mm in range(1, self.n_m) for m in (mm, mm)])
        self.mask = self.mode_m[:, None] <= np.arange(self.n_l)[None, :]
        self.p_mode = self.p[self.mode_m]            # [2M-1, nlat, L]

    @property
    def modal_shape(self):
        return (2 * self.n_m - 1, self.n_l)

    @property
    def nodal_shape(self):
        return (self.nlon, self.nlat)

    # ---- transforms (dinosaur RealSphericalHarmonics.transform / inverse_transform)
    def to_modal(self, x: np.ndarray) -> np.ndarray:
        wx = np.asarray(x, np.float64) * self.weights
        fwx = np.einsum("im,...ij->...mj", self.f, wx)
        return np.einsum("mjl,...mj->...ml", self.p_mode, fwx)

    def to_nodal(self, x: np.ndarray) -> np.ndarray:
        px = np.einsum("mjl,...ml->...mj", self.p_mode, np.asarray(x, np.float64))
        return np.einsum("im,...mj->...ij", self.f, px)

    # ---- modal operators
    def _l(self):
        return np.arange(self.n_l, dtype=np.float64)

    def laplacian(self, x):
        l = self._l()
        return x * (-l * (l + 1.0) / self.radius ** 2)

    def inverse_laplacian(self, x):
        l = self._l()
        inv = np.zeros(self.n_l)
        inv[1:] = -self.radius ** 2 / (l[1:] * (l[1:] + 1.0))
        return x * inv

    def clip(self, x):
        out = np.array(x, np.float64, copy=True)
        out[..., -1] = 0.0
        return out

    def d_dlon(self, x):
        """Longitude derivative: cos slot <- m * sin coefficient, sin slot <- -m * cos."""
        out = np.zeros_like(x, dtype=np.float64)
        for m in range(1, self.n_m):
            out[..., 2 * m - 1, :] = m * x[..., 2 * m, :]
            out[..., 2 * m, :] = -m * x[..., 2 * m - 1, :]
        return out

    def _eps(self):
        """eps[mode, l] = sqrt((l^2 - m^2) / (4 l^2 - 1)) where m <= l, else 0:
        x P_l^m = eps_{l+1} P_{l+1}^m + eps_l P_{l-1}^m for the normalised functions."""
        l = self._l()[None, :]
        m = self.mode_m[:, None].astype(np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            e = np.sqrt(np.maximum(l * l - m * m, 0.0) / (4.0 * l * l - 1.0))
        e[:, 0] = 0.0
        return e

    def cos_lat_d_dlat(self, x):
        """(1 - x^2) d/dx in modal form:
        out_l = (l+2) eps_{l+1} x_{l+1} - (l-1) eps_l x_{l-1}."""
        e = self._eps()
        l = self._l()
        out = np.zeros_like(x, dtype=np.float64)
        out[..., :-1] += ((l[:-1] + 2.0) * e[:, 1:]) * x[..., 1:]
        out[..., 1:] += (-(l[1:] - 1.0) * e[:, 1:]) * x[..., :-1]
        return out

    def sec_lat_d_dlat_cos2(self, x):
        """(1/cos) d/dlat (cos^2 x) in modal form:
        out_l = l eps_{l+1} x_{l+1} - (l+1) eps_l x_{l-1}."""
        e = self._eps()
        l = self._l()
        out = np.zeros_like(x, dtype=np.float64)
        out[..., :-1] += (l[:-1] * e[:, 1:]) * x[..., 1:]
        out[..., 1:] += (-(l[1:] + 1.0) * e[:, 1:]) * x[..., :-1]
        return out

    def cos_lat_grad(self, x):
        return (self.clip(self.d_dlon(x) / self.radius), self.clip(self.cos_lat_d_dlat(x) / self.radius))

    def div_cos_lat(self, u, v):
        return self.clip((self.d_dlon(u) + self.sec_lat_d_dlat_cos2(v)) / self.radius)

    def curl_cos_lat(self, u, v):
        return self.clip((self.d_dlon(v) - self.sec_lat_d_dlat_cos2(u)) / self.radius)


def vor_div_to_uv_nodal(grid: Grid, vorticity, divergence):
    psi = grid.inverse_laplacian(vorticity)
    chi = grid.inverse_laplacian(divergence)
    chi_lon, chi_lat = grid.cos_lat_grad(chi)
    psi_lon, psi_lat = grid.cos_lat_grad(psi)
    u_cos = chi_lon - psi_lat
    v_cos = chi_lat + psi_lon
    return grid.to_nodal(u_cos) / grid.cos_lat, grid.to_nodal(v_cos) / grid.cos_lat


def uv_nodal_to_vor_div_modal(grid: Grid, u, v):
    um = grid.to_modal(np.asarray(u, np.float64) / grid.cos_lat)
    vm = grid.to_modal(np.asarray(v, np.float64) / grid.cos_lat)
    return grid.curl_cos_lat(um, vm), grid.div_cos_lat(um, vm)


def sigma_centers(boundaries) -> np.ndarray:
    b = np.asarray(boundaries, np.float64)
    return (b[1:] + b[:-1]) / 2.0


def geopotential_weights(centers, gas_constant=IDEAL_GAS_CONSTANT) -> np.ndarray:
    """W[j, k]: phi_j = phi_s + sum_k W[j, k] T_k.  alpha_k = (log s_{k+1} - log s_k)/2,
    alpha_{K-1} = -log s_{K-1}; W[j, j] = alpha_j, W[j, k>j] = alpha_k + alpha_{k-1}."""
    ls = np.log(np.asarray(centers, np.float64))
    kx = ls.size
    alpha = np.zeros(kx)
    alpha[:-1] = np.diff(ls) / 2.0
    alpha[-1] = -ls[-1]
    w = np.zeros((kx, kx))
    for j in range(kx):
        w[j, j] = alpha[j]
        for k in range(j + 1, kx):
            w[j, k] = alpha[k] + alpha[k - 1]
    return gas_constant * w


@dataclass
class DynamicsState:
    """dinosaur primitive_equations.State (modal fields)."""
    vorticity: np.ndarray
    divergence: np.ndarray
    temperature_variation: np.ndarray
    log_surface_pressure: np.ndarray
    tracers: Dict[str, np.ndarray] = field(default_factory=dict)
    sim_time: Optional[float] = None


@dataclass
class Dynamics:
    """What the transforms read from dinosaur's PrimitiveEquations."""
    grid: Grid
    sigma_boundaries: np.ndarray
    reference_temperature: np.ndarray          # [kx]
    orography: np.ndarray                      # modal [2M-1, L]
    gravity: float = GRAVITY
    gas_constant: float = IDEAL_GAS_CONSTANT


def dynamics_state_to_physics_state(state: DynamicsState, dyn: Dynamics):
    """physics_interface.py:196-234 -> (u, v, t, q, phi, sp) nodal."""
    g = dyn.grid
    u, v = vor_div_to_uv_nodal(g, state.vorticity, state.divergence)
    ref = np.asarray(dyn.reference_temperature, np.float64)
    t = g.to_nodal(state.temperature_variation) + ref[:, None, None]
    q = g.to_nodal(state.tracers["specific_humidity"]) * 1000.0
    w = geopotential_weights(sigma_centers(dyn.sigma_boundaries), dyn.gas_constant)
    phi = g.to_nodal(dyn.orography)[None] * dyn.gravity + np.einsum("jk,kxy->jxy", w, t)
    sp = np.squeeze(np.exp(g.to_nodal(state.log_surface_pressure)))
    return u, v, t, q, phi, sp


def physics_state_to_dynamics_state(u, v, t, q, sp, dyn: Dynamics) -> DynamicsState:
    """physics_interface.py:237-273."""
    g = dyn.grid
    vor, div = uv_nodal_to_vor_div_modal(g, u, v)
    qm = g.to_modal(np.asarray(q, np.float64) * 1e-3)
    ref = np.asarray(dyn.reference_temperature, np.float64)
    tm = g.to_modal(np.asarray(t, np.float64) - ref[:, None, None])
    lsp = g.to_modal(np.log(np.asarray(sp, np.float64)))
    return DynamicsState(vor, div, tm, lsp, {"specific_humidity": qm})


def physics_tendency_to_dynamics_tendency(du, dv, dt, dq, dyn: Dynamics) -> DynamicsState:
    """physics_interface.py:275-303."""
    g = dyn.grid
    vor, div = uv_nodal_to_vor_div_modal(g, du, dv)
    tm = g.to_modal(dt)
    qm = g.to_modal(np.asarray(dq, np.float64) * 1e-3)
    return DynamicsState(vor, div, tm, np.zeros_like(tm[0]), {"specific_humidity": qm}, sim_time=0.0)
