"""Geometry: mirror of jcm/geometry.py (Geometry, _initialize_vertical,
from_grid_shape, single_column_geometry) without the dycore dependency.

Per-level constants follow geometry.py:62-85 in float32.  Latitudes are the
Gaussian latitudes of the grid (south to north, geometry.py:130), computed with
NumPy's Gauss-Legendre nodes in place of dinosaur's grid.  ``phis0`` defaults
to ``grav * orog`` (the reference spectrally truncates it, geometry.py:128;
that transform belongs to the dycore and is not part of this library).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native

P0 = 1e5
GRAV = 9.81
CP = 1004.0

# physical_constants.py:46-50; 16 levels: this library's own choice (each
# 8-level layer split in two), the reference accepts only 7 and 8.
SIGMA_LAYER_BOUNDARIES = {
    7: [0.0, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
    8: [0.0, 0.05, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
    16: [0.0, 0.025, 0.05, 0.095, 0.14, 0.2, 0.26, 0.34, 0.42, 0.51, 0.6,
         0.685, 0.77, 0.835, 0.9, 0.95, 1.0],
}

# utils.py:12-26 (truncation <-> nodal shape) for the grids this library serves.
TRUNCATION_FOR_NODAL_SHAPE = {(64, 32): 21, (96, 48): 31, (128, 64): 42, (192, 96): 63,
                              (256, 128): 85, (320, 160): 106, (384, 192): 127, (512, 256): 170,
                              (640, 320): 213, (1024, 512): 340, (1280, 640): 425}


def initialize_vertical(kx: int):
    """_initialize_vertical (geometry.py:62-85), float32."""
    if kx not in SIGMA_LAYER_BOUNDARIES:
        raise ValueError(f"Invalid number of vertical levels: {kx}")
    f = np.float32
    hsg = np.asarray(SIGMA_LAYER_BOUNDARIES[kx], f)
    fsg = (hsg[1:] + hsg[:-1]) / f(2.0)
    dhs = np.diff(hsg)
    sigl = np.log(fsg)
    grdsig = f(GRAV) / (dhs * f(P0))
    grdscp = grdsig / f(CP)
    wvi = np.zeros((kx, 2), f)
    wvi[:-1, 0] = f(1.0) / np.diff(sigl)
    wvi[:-1, 1] = (np.log(hsg[1:-1]) - sigl[:-1]) * wvi[:-1, 0]
    wvi[-1, 1] = (np.log(f(0.99)) - sigl[-1]) * wvi[-2, 0]
    return hsg, fsg, dhs, sigl, grdsig, grdscp, wvi


def gaussian_latitudes(il: int) -> np.ndarray:
    """Gaussian latitudes in radians, south to north."""
    x, _ = np.polynomial.legendre.leggauss(il)
    return np.arcsin(x)


@dataclass
class Geometry:
    nodal_shape: tuple            # (kx, ix, il)
    orog: np.ndarray              # (ix, il) m
    phis0: np.ndarray             # (ix, il) surface geopotential
    fmask: np.ndarray             # (ix, il) land fraction
    radang: np.ndarray            # (il,)
    sia: np.ndarray               # (il,)
    coa: np.ndarray               # (il,)
    hsg: np.ndarray
    fsg: np.ndarray
    dhs: np.ndarray
    sigl: np.ndarray
    grdsig: np.ndarray
    grdscp: np.ndarray
    wvi: np.ndarray

    @classmethod
    def from_latitudes(cls, radang, ix, num_levels=8, orography=None, fmask=None, phis0=None,
                       sia=None, coa=None):
        f = np.float32
        il = len(radang)
        orog = np.zeros((ix, il), f) if orography is None else np.asarray(orography, f)
        fm = np.zeros((ix, il), f) if fmask is None else np.asarray(fmask, f)
        if fmask is not None:   # get_terrain rounding (geometry.py:57-58)
            fm = np.where(fm <= 0.1, f(0.0), np.where(fm >= 0.9, f(1.0), fm)).astype(f)
        if phis0 is None:
            phis0 = f(GRAV) * orog
        radang = np.asarray(radang, f)
        hsg, fsg, dhs, sigl, grdsig, grdscp, wvi = initialize_vertical(num_levels)
        return cls((num_levels, ix, il), orog, np.asarray(phis0, f), fm, radang,
                   np.sin(radang) if sia is None else np.asarray(sia, f),
                   np.cos(radang) if coa is None else np.asarray(coa, f),
                   hsg, fsg, dhs, sigl, grdsig, grdscp, wvi)

    @classmethod
    def from_grid_shape(cls, nodal_shape, num_levels=8, orography=None, fmask=None, phis0=None):
        """geometry.py:163-182 with Gaussian latitudes."""
        ix, il = nodal_shape
        return cls.from_latitudes(gaussian_latitudes(il), ix, num_levels, orography, fmask, phis0)

    @classmethod
    def single_column_geometry(cls, radang=0.0, orog=0.0, fmask=0.0, phis0=None, num_levels=8):
        """geometry.py:207-236."""
        f = np.float32
        if phis0 is None:
            phis0 = GRAV * orog
        return cls.from_latitudes(np.array([radang], f), 1, num_levels,
                                  np.full((1, 1), orog, f), None, np.full((1, 1), phis0, f)).replace_fmask(fmask)

    def replace_fmask(self, fmask):
        self.fmask = np.full(self.fmask.shape, fmask, np.float32) if np.isscalar(fmask) else np.asarray(fmask, np.float32)
        return self

    @property
    def kx(self) -> int:
        return self.nodal_shape[0]

    @property
    def ncol(self) -> int:
        return self.nodal_shape[1] * self.nodal_shape[2]

    def vertical_key(self) -> tuple:
        """Every value to_abi packs, as bytes: equal keys give equal ABI structs."""
        return (self.kx,) + tuple(np.asarray(getattr(self, n)).tobytes()
                                  for n in ("hsg", "fsg", "dhs", "sigl", "grdsig", "grdscp", "wvi"))

    def to_abi(self) -> _native.SpeedyVertical:
        v = _native.SpeedyVertical()
        kx = self.kx
        v.kx = kx
        for i in range(kx + 1):
            v.hsg[i] = float(self.hsg[i])
        for name in ("fsg", "dhs", "sigl", "grdsig", "grdscp"):
            arr = getattr(self, name)
            dst = getattr(v, name)
            for i in range(kx):
                dst[i] = float(arr[i])
        for i in range(kx):
            v.wvi[i][0] = float(self.wvi[i, 0])
            v.wvi[i][1] = float(self.wvi[i, 1])
        return v
