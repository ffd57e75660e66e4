"""Seeded synthetic SPEEDY inputs for the benchmark and parity workloads.

Recipe (aquaplanet; ``land=True`` adds the land surface):
  * Gaussian latitudes (south to north), ix longitudes, kx sigma levels.
  * T = 288 * sigma**(R*Gamma/g) + N(0, 2 K) with R = 286.857, Gamma = 6.5 K/km,
    g = 9.81, floored at 200 K (the lapse-rate profile alone reaches 143 K at
    sigma = 0.025, below the range the scheme's tables cover).
  * q = RH * qsat(T, ps, sigma) [g/kg] with RH ~ U(0.3, 0.95).
  * u, v ~ N(0, 10 m/s); normalised surface pressure ps = exp(N(0, 0.01)).
  * geopotential by hydrostatic integration in log-sigma from phi_s = g*orog:
    phi[kx-1] = phi_s - R*T[kx-1]*ln(sigma[kx-1]),
    phi[k] = phi[k+1] - R*(T[k]+T[k+1])/2 * ln(sigma[k]/sigma[k+1]).
  * SST = 273.15 + 27*cos(1.5*lat)**2 K, clamped >= 271.4 K, zonally symmetric.
  * aquaplanet: fmask = 0, orog = 0, stl_am = SST, soilw = snowc = sice = 0.
  * land: fmask = 1 where Gaussian-smoothed noise exceeds its 70th percentile
    (30 % land), orog ~ |N(0, 500 m)| on land, phis0 = g*orog,
    soilw_am ~ U(0, 1), snow depth ~ U(0, 0.1) m -> snowc_am = 1000*depth/60,
    stl_am = SST - 0.0065*orog + N(0, 1 K) on land.
  * alb0 = 0.2, tyear = 0.25, model_step = 0 (shortwave on), lfluxland = True.
"""
from __future__ import annotations

import numpy as np

from .date import DateData
from .forcing import ForcingData
from .geometry import Geometry, gaussian_latitudes
from .params import Parameters
from .physics_interface import PhysicsState

RGAS = 2.0 / 7.0 * 1004.0
GRAV = 9.81

RESOLUTIONS = {"T21": (64, 32), "T31": (96, 48), "T42": (128, 64), "T85": (256, 128), "T425": (1280, 640)}


def qsat_np(ta, ps, sig):
    """Saturation specific humidity [g/kg] (humidity.py:87-102) in float64."""
    ta = np.asarray(ta, np.float64)
    e = np.where(ta >= 273.16, 6.108e-3 * np.exp(17.269 * (ta - 273.16) / (ta - 35.86)),
                 6.108e-3 * np.exp(21.875 * (ta - 273.16) / (ta - 7.66)))
    return 622.0 * e / (sig * ps - 0.378 * e)


def _smooth_noise(rng, ix, il, sigma=3.0):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(rng.standard_normal((ix, il)), sigma=sigma, mode=("wrap", "nearest"))


def make_case(nlon: int, nlat: int, kx: int = 8, seed: int = 0, land: bool = False, nmember: int = 1,
              member_t_sigma: float = 0.0, tyear: float = 0.25, model_step: int = 0):
    """Returns (state, forcing, geometry, date) with NumPy float32 fields.
    State fields are (kx, ix, il), or (nmember, kx, ix, il) when nmember > 1."""
    rng = np.random.default_rng(seed)
    ix, il = nlon, nlat
    lat = gaussian_latitudes(il)
    geom0 = Geometry.from_latitudes(lat, ix, kx)
    fsg = geom0.fsg.astype(np.float64)
    if land:
        noise = _smooth_noise(rng, ix, il)
        fmask = (noise > np.percentile(noise, 70.0)).astype(np.float64)
        orog = np.abs(rng.normal(0.0, 500.0, (ix, il))) * fmask
        soilw = rng.uniform(0.0, 1.0, (ix, il))
        snowc = 1000.0 * rng.uniform(0.0, 0.1, (ix, il)) / 60.0
    else:
        fmask = np.zeros((ix, il))
        orog = np.zeros((ix, il))
        soilw = np.zeros((ix, il))
        snowc = np.zeros((ix, il))
    sst = np.maximum(273.15 + 27.0 * np.cos(1.5 * lat) ** 2, 271.4)
    sst = np.broadcast_to(sst, (ix, il)).copy()
    stl = sst.copy()
    if land:
        stl = np.where(fmask > 0, sst - 0.0065 * orog + rng.normal(0.0, 1.0, (ix, il)), sst)
    geometry = Geometry.from_latitudes(lat, ix, kx, orography=orog.astype(np.float32),
                                       fmask=fmask.astype(np.float32))
    forcing = ForcingData(alb0=np.full((ix, il), 0.2, np.float32), sice_am=np.zeros((ix, il), np.float32),
                          snowc_am=snowc.astype(np.float32), soilw_am=soilw.astype(np.float32),
                          stl_am=stl.astype(np.float32), sea_surface_temperature=sst.astype(np.float32),
                          lfluxland=True)
    shape = (nmember, kx, ix, il)
    gamma = RGAS * 6.5e-3 / GRAV
    tprof = 288.0 * fsg ** gamma
    base_t = tprof[:, None, None] + rng.normal(0.0, 2.0, (kx, ix, il))
    T = np.broadcast_to(base_t, shape).copy()
    if nmember > 1 and member_t_sigma > 0:
        T = T + rng.normal(0.0, member_t_sigma, shape)
    T = np.maximum(T, 200.0)
    ps = np.exp(rng.normal(0.0, 0.01, (nmember, ix, il)))
    rh = rng.uniform(0.3, 0.95, shape)
    q = rh * qsat_np(T, ps[:, None], fsg[None, :, None, None])
    u = rng.normal(0.0, 10.0, shape)
    v = rng.normal(0.0, 10.0, shape)
    phis = GRAV * orog
    phi = np.empty(shape)
    phi[:, kx - 1] = phis - RGAS * T[:, kx - 1] * np.log(fsg[kx - 1])
    for k in range(kx - 2, -1, -1):
        phi[:, k] = phi[:, k + 1] - RGAS * 0.5 * (T[:, k] + T[:, k + 1]) * np.log(fsg[k] / fsg[k + 1])
    f32 = lambda a: np.ascontiguousarray(a[0] if nmember == 1 else a, np.float32)
    state = PhysicsState(u_wind=f32(u), v_wind=f32(v), temperature=f32(T), specific_humidity=f32(q),
                         geopotential=f32(phi), normalized_surface_pressure=f32(ps))
    date = DateData(tyear=tyear, model_year=1950, model_step=model_step)
    return state, forcing, geometry, date


def draw_vjp_parameters(seed: int = 2) -> Parameters:
    """The five differentiated parameters drawn as in the adjoint workload."""
    rng = np.random.default_rng(seed)
    p = Parameters.default()
    return p.replace(convection=dict(rhbl=float(rng.uniform(0.7, 0.9)), trcnv=float(rng.uniform(4.0, 8.0))),
                     condensation=dict(rhlsc=float(rng.uniform(0.9, 0.98))),
                     forcing=dict(co2_scale=float(rng.uniform(0.8, 1.2))),
                     surface_flux=dict(drag_scale=float(rng.uniform(0.5, 1.5))))
