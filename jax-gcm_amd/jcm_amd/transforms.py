"""Dynamics <-> physics transforms on the GPU (C ABI: include/speedy_sht.h).

Mirror of jcm/physics_interface.py:196-303 and of the dinosaur pieces those
functions use (Grid.T<N> with RealSphericalHarmonics, SigmaCoordinates,
primitive_equations.State, the reference temperature / orography / constants
of PrimitiveEquations):

  dynamics_state_to_physics_state        physics_interface.py:196-234
  physics_state_to_dynamics_state        physics_interface.py:237-273
  physics_tendency_to_dynamics_tendency  physics_interface.py:275-303

Fields are torch tensors (float32, contiguous, on the plan's device) or NumPy
arrays (uploaded).  Modal fields are [..., 2*n_m-1, n_l] with mode 0 -> m=0,
2m-1 -> cos(m lon), 2m -> sin(m lon); nodal fields are [..., nlon, nlat]
with Gauss-Legendre latitudes south to north.  Units are SI (dinosaur's
SI_SCALE).  There is no CPU path: without the HIP library or a GPU the calls
raise.
"""
from __future__ import annotations

import ctypes
import os
import threading
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from .physics_interface import PhysicsState, PhysicsTendency

LIB_NAME = "libspeedy_sht.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)
KX_MAX = 16

GRAVITY = 9.80616                 # dinosaur scales.GRAVITY_ACCELERATION, m s-2
IDEAL_GAS_CONSTANT = 1004.0 * 2.0 / 7.0   # ISOBARIC_HEAT_CAPACITY * KAPPA, J kg-1 K-1
RADIUS = 6.37122e6                # scales.RADIUS, m

# dinosaur Grid.T<N>: (longitude wavenumbers, total wavenumbers, nlon, nlat)
TRUNCATIONS = {21: (22, 23, 64, 32), 31: (32, 33, 96, 48), 42: (43, 44, 128, 64),
               85: (86, 87, 256, 128), 106: (107, 108, 320, 160), 119: (120, 121, 360, 180),
               170: (171, 172, 512, 256), 213: (214, 215, 640, 320), 340: (341, 342, 1024, 512),
               425: (426, 427, 1280, 640)}

CFP = ctypes.c_void_p


class ShtError(RuntimeError):
    pass


class _DynamicsSpec(ctypes.Structure):
    _fields_ = [("kx", ctypes.c_int32), ("sigma_boundaries", ctypes.POINTER(ctypes.c_double)),
                ("reference_temperature", CFP), ("orography", CFP), ("gravity", ctypes.c_double),
                ("gas_constant", ctypes.c_double)]


class _Spectral(ctypes.Structure):
    _fields_ = [(n, CFP) for n in ("vorticity", "divergence", "temperature_variation", "log_surface_pressure",
                                   "specific_humidity")] + [("nsp", ctypes.c_int32)]


class _Nodal(ctypes.Structure):
    _fields_ = [(n, CFP) for n in ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential",
                                   "normalized_surface_pressure")] + [("nsp", ctypes.c_int32)]


_lib = None
_lib_lock = threading.Lock()


def load(path: Optional[str] = None):
    """Load libspeedy_sht.so once; raise ShtError if it is missing (no fallback)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise ShtError(f"{p} not found: build it with __graft_entry__.build() (there is no CPU fallback)")
        lib = ctypes.CDLL(p)
        i32, sz = ctypes.c_int32, ctypes.c_size_t
        dp, fp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
        lib.speedy_sht_host_basis.restype = ctypes.c_int
        lib.speedy_sht_host_basis.argtypes = [i32, i32, i32, i32, dp, dp, fp, fp]
        lib.speedy_sht_create.restype = ctypes.c_int
        lib.speedy_sht_create.argtypes = [i32, i32, i32, i32, ctypes.c_double, ctypes.POINTER(CFP)]
        lib.speedy_sht_destroy.restype = ctypes.c_int
        lib.speedy_sht_destroy.argtypes = [CFP]
        lib.speedy_sht_workspace_bytes.restype = sz
        lib.speedy_sht_workspace_bytes.argtypes = [CFP, i32]
        for n in ("speedy_sht_to_modal", "speedy_sht_to_nodal"):
            getattr(lib, n).restype = ctypes.c_int
            getattr(lib, n).argtypes = [CFP, CFP, CFP, i32, CFP, sz, CFP]
        lib.speedy_dynamics_to_physics.restype = ctypes.c_int
        lib.speedy_dynamics_to_physics.argtypes = [CFP, ctypes.POINTER(_DynamicsSpec), ctypes.POINTER(_Spectral),
                                                   ctypes.POINTER(_Nodal), CFP, sz, CFP]
        for n in ("speedy_physics_to_dynamics", "speedy_physics_tendency_to_dynamics"):
            getattr(lib, n).restype = ctypes.c_int
            getattr(lib, n).argtypes = [CFP, ctypes.POINTER(_DynamicsSpec), ctypes.POINTER(_Nodal),
                                        ctypes.POINTER(_Spectral), CFP, sz, CFP]
        lib.speedy_sht_last_error.restype = ctypes.c_char_p
        lib.speedy_sht_last_error.argtypes = []
        _lib = lib
        return lib


def _check(rc: int, what: str):
    if rc != 0:
        raise ShtError(f"{what} failed ({rc}): {load().speedy_sht_last_error().decode()}")


def host_basis(n_m: int, n_l: int, nlon: int, nlat: int):
    """(sin_lat, weights, fourier[nlon, 2*n_m-1], legendre[n_m, nlat, n_l]) as the
    library builds them, on the host (no GPU call)."""
    lib = load()
    x, w = np.zeros(nlat), np.zeros(nlat)
    f = np.zeros((nlon, 2 * n_m - 1), np.float32)
    p = np.zeros((n_m, nlat, n_l), np.float32)
    dp, fp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
    _check(lib.speedy_sht_host_basis(n_m, n_l, nlon, nlat, x.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                     f.ctypes.data_as(fp), p.ctypes.data_as(fp)), "speedy_sht_host_basis")
    return x, w, f, p


def _torch():
    import torch
    return torch


class SpectralGrid:
    """dinosaur Grid.T<N> with RealSphericalHarmonics: one device plan per GPU and
    a workspace per (device, stream) that grows with the stacked field count."""

    def __init__(self, longitude_wavenumbers: int, total_wavenumbers: int, longitude_nodes: int,
                 latitude_nodes: int, radius: float = RADIUS):
        self.longitude_wavenumbers = self.n_m = int(longitude_wavenumbers)
        self.total_wavenumbers = self.n_l = int(total_wavenumbers)
        self.longitude_nodes = self.nlon = int(longitude_nodes)
        self.latitude_nodes = self.nlat = int(latitude_nodes)
        self.radius = float(radius)
        self._plans: Dict[int, int] = {}
        self._ws: Dict[tuple, object] = {}
        self._lock = threading.Lock()
        x, _ = np.polynomial.legendre.leggauss(self.nlat)
        self.sin_lat = x
        self.cos_lat = np.sqrt(1.0 - x * x)
        self.latitudes = np.arcsin(x)
        self.longitudes = 2.0 * np.pi * np.arange(self.nlon) / self.nlon

    @classmethod
    def truncation(cls, n: int, radius: float = RADIUS) -> "SpectralGrid":
        """Grid.T<n> (dinosaur spherical_harmonic.Grid.T21 ... T425)."""
        if n not in TRUNCATIONS:
            raise ValueError(f"no Grid.T{n}; known truncations: {sorted(TRUNCATIONS)}")
        return cls(*TRUNCATIONS[n], radius=radius)

    @property
    def modal_shape(self):
        return (2 * self.n_m - 1, self.n_l)

    @property
    def nodal_shape(self):
        return (self.nlon, self.nlat)

    def plan(self, device) -> int:
        torch = _torch()
        dev = torch.device(device)
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        with self._lock:
            if idx not in self._plans:
                h = CFP()
                with torch.cuda.device(idx):
                    _check(load().speedy_sht_create(self.n_m, self.n_l, self.nlon, self.nlat, self.radius,
                                                    ctypes.byref(h)), "speedy_sht_create")
                self._plans[idx] = h.value
            return self._plans[idx]

    def workspace(self, device, nfield: int):
        torch = _torch()
        plan = self.plan(device)
        need = load().speedy_sht_workspace_bytes(plan, nfield)
        stream = torch.cuda.current_stream(device)
        key = (torch.device(device).index, stream.cuda_stream)
        with self._lock:
            ws = self._ws.get(key)
            if ws is None or ws.numel() < need:
                if torch.cuda.is_current_stream_capturing():
                    raise ShtError("transform workspace must exist before graph capture: call once uncaptured")
                ws = torch.empty(need, dtype=torch.uint8, device=device)
                self._ws[key] = ws
        return plan, ws, stream

    def __del__(self):
        try:
            lib = load()
            for h in self._plans.values():
                lib.speedy_sht_destroy(h)
        except Exception:
            pass

    # ---- Grid.to_modal / Grid.to_nodal over any leading axes
    def to_modal(self, x, device=None):
        t = _as_device(x, device)
        lead = tuple(t.shape[:-2])
        if tuple(t.shape[-2:]) != self.nodal_shape:
            raise ValueError(f"to_modal: trailing shape {tuple(t.shape[-2:])} != {self.nodal_shape}")
        nf = int(np.prod(lead)) if lead else 1
        out = _torch().empty(lead + self.modal_shape, dtype=t.dtype, device=t.device)
        plan, ws, s = self.workspace(t.device, nf)
        _check(load().speedy_sht_to_modal(plan, t.data_ptr(), out.data_ptr(), nf, ws.data_ptr(), ws.numel(),
                                          s.cuda_stream), "speedy_sht_to_modal")
        return out

    def to_nodal(self, x, device=None):
        t = _as_device(x, device)
        lead = tuple(t.shape[:-2])
        if tuple(t.shape[-2:]) != self.modal_shape:
            raise ValueError(f"to_nodal: trailing shape {tuple(t.shape[-2:])} != {self.modal_shape}")
        nf = int(np.prod(lead)) if lead else 1
        out = _torch().empty(lead + self.nodal_shape, dtype=t.dtype, device=t.device)
        plan, ws, s = self.workspace(t.device, nf)
        _check(load().speedy_sht_to_nodal(plan, t.data_ptr(), out.data_ptr(), nf, ws.data_ptr(), ws.numel(),
                                          s.cuda_stream), "speedy_sht_to_nodal")
        return out


@dataclass
class SigmaCoordinates:
    """dinosaur sigma_coordinates.SigmaCoordinates (layer boundaries, top to surface)."""
    boundaries: np.ndarray

    def __post_init__(self):
        self.boundaries = np.asarray(self.boundaries, np.float64)

    @property
    def centers(self):
        return (self.boundaries[1:] + self.boundaries[:-1]) / 2.0

    @property
    def layers(self) -> int:
        return self.boundaries.size - 1


@dataclass
class CoordinateSystem:
    horizontal: SpectralGrid
    vertical: SigmaCoordinates


@dataclass
class DynamicsState:
    """dinosaur primitive_equations.State: modal fields [kx, 2*n_m-1, n_l]
    (log_surface_pressure [1, ...] or [...])."""
    vorticity: object
    divergence: object
    temperature_variation: object
    log_surface_pressure: object
    tracers: Dict[str, object] = field(default_factory=dict)
    sim_time: Optional[float] = None


@dataclass
class PrimitiveEquations:
    """What the conversions read from dinosaur's PrimitiveEquations: the reference
    temperature [kx] (K), the truncated modal orography [2*n_m-1, n_l] (m), the
    coordinates, and the constants in SI units."""
    reference_temperature: object
    orography: object
    coords: CoordinateSystem
    gravity: float = GRAVITY
    gas_constant: float = IDEAL_GAS_CONSTANT

    def spec(self, device):
        """(ctypes spec, objects that must stay alive during the call) for `device`:
        device tensors are used in place, NumPy arrays are uploaded per call."""
        ref = _as_device(self.reference_temperature, device)
        orog = _as_device(self.orography, device)
        sig = np.ascontiguousarray(self.coords.vertical.boundaries, np.float64)
        kx = self.coords.vertical.layers
        if kx < 1 or kx > KX_MAX or ref.numel() != kx or sig.size != kx + 1:
            raise ValueError(f"need 1..{KX_MAX} levels and one reference temperature per level")
        if tuple(orog.shape) != self.coords.horizontal.modal_shape:
            raise ValueError("orography must be modal [2*n_m-1, n_l]")
        s = _DynamicsSpec(kx, sig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ref.data_ptr(), orog.data_ptr(),
                          float(self.gravity), float(self.gas_constant))
        return s, (ref, orog, sig)


def _as_device(x, device=None):
    torch = _torch()
    if torch.is_tensor(x):
        t = x
        if device is not None and t.device != torch.device(device):
            t = t.to(device)
        if t.device.type != "cuda":
            raise ShtError("transform fields must be on a GPU (there is no CPU path)")
        if t.dtype != torch.float32:
            t = t.float()
        return t.contiguous()
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    return torch.as_tensor(np.ascontiguousarray(x, np.float32), device=dev)


def _device_of(*xs):
    torch = _torch()
    for x in xs:
        if torch.is_tensor(x) and x.device.type == "cuda":
            return x.device
    return torch.device("cuda", torch.cuda.current_device())


def dynamics_state_to_physics_state(state: DynamicsState, dynamics: PrimitiveEquations) -> PhysicsState:
    """physics_interface.py:196-234 on the GPU: winds from vorticity/divergence,
    temperature plus the reference, humidity in g/kg, geopotential from
    orography and temperature, surface pressure exp(log_surface_pressure)."""
    torch = _torch()
    g = dynamics.coords.horizontal
    dev = _device_of(state.vorticity, state.divergence, state.temperature_variation)
    vor, div, tv = (_as_device(x, dev) for x in (state.vorticity, state.divergence, state.temperature_variation))
    q = _as_device(state.tracers["specific_humidity"], dev)
    lsp = _as_device(state.log_surface_pressure, dev)
    kx = dynamics.coords.vertical.layers
    for name, t in (("vorticity", vor), ("divergence", div), ("temperature_variation", tv),
                    ("specific_humidity", q)):
        if tuple(t.shape) != (kx,) + g.modal_shape:
            raise ValueError(f"{name}: shape {tuple(t.shape)} != {(kx,) + g.modal_shape}")
    nsp = 1 if lsp.dim() == 2 else int(lsp.shape[0])
    if lsp.dim() not in (2, 3) or tuple(lsp.shape[-2:]) != g.modal_shape or nsp > kx:
        raise ValueError(f"log_surface_pressure: shape {tuple(lsp.shape)}")
    spec, keep = dynamics.spec(dev)
    shape = (kx,) + g.nodal_shape
    u, v, t, qn, phi = (torch.empty(shape, device=dev) for _ in range(5))
    sp = torch.empty((nsp,) + g.nodal_shape, device=dev)
    nfield = 5 * kx + nsp
    plan, ws, s = g.workspace(dev, nfield)
    ins = _Spectral(vor.data_ptr(), div.data_ptr(), tv.data_ptr(), lsp.data_ptr(), q.data_ptr(), nsp)
    outs = _Nodal(u.data_ptr(), v.data_ptr(), t.data_ptr(), qn.data_ptr(), phi.data_ptr(), sp.data_ptr(), nsp)
    _check(load().speedy_dynamics_to_physics(plan, ctypes.byref(spec), ctypes.byref(ins), ctypes.byref(outs),
                                             ws.data_ptr(), ws.numel(), s.cuda_stream),
           "speedy_dynamics_to_physics")
    return PhysicsState(u, v, t, qn, phi, torch.squeeze(sp))


def _nodal_to_spectral(u, v, t, q, sp, dynamics, tendency):
    torch = _torch()
    g = dynamics.coords.horizontal
    kx = dynamics.coords.vertical.layers
    dev = _device_of(u, v, t, q)
    u, v, t, q = (_as_device(x, dev) for x in (u, v, t, q))
    for name, x in (("u", u), ("v", v), ("temperature", t), ("specific_humidity", q)):
        if tuple(x.shape) != (kx,) + g.nodal_shape:
            raise ValueError(f"{name}: shape {tuple(x.shape)} != {(kx,) + g.nodal_shape}")
    if tendency:
        nsp, sp_ptr, lsp_shape = 1, None, g.modal_shape
    else:
        sp = _as_device(sp, dev)
        if sp.dim() not in (2, 3) or tuple(sp.shape[-2:]) != g.nodal_shape:
            raise ValueError(f"normalized_surface_pressure: shape {tuple(sp.shape)}")
        nsp = 1 if sp.dim() == 2 else int(sp.shape[0])
        if nsp > kx:
            raise ValueError("normalized_surface_pressure: more fields than levels")
        sp_ptr, lsp_shape = sp.data_ptr(), tuple(sp.shape[:-2]) + g.modal_shape
    spec, keep = dynamics.spec(dev)
    mshape = (kx,) + g.modal_shape
    vor, div, tm, qm = (torch.empty(mshape, device=dev) for _ in range(4))
    lsp = torch.empty(lsp_shape, device=dev)
    nfield = 4 * kx + (0 if tendency else nsp)
    plan, ws, s = g.workspace(dev, nfield)
    ins = _Nodal(u.data_ptr(), v.data_ptr(), t.data_ptr(), q.data_ptr(), None, sp_ptr, 0 if tendency else nsp)
    outs = _Spectral(vor.data_ptr(), div.data_ptr(), tm.data_ptr(), lsp.data_ptr(), qm.data_ptr(), nsp)
    fn = "speedy_physics_tendency_to_dynamics" if tendency else "speedy_physics_to_dynamics"
    _check(getattr(load(), fn)(plan, ctypes.byref(spec), ctypes.byref(ins), ctypes.byref(outs), ws.data_ptr(),
                               ws.numel(), s.cuda_stream), fn)
    return vor, div, tm, lsp, qm


def physics_state_to_dynamics_state(physics_state: PhysicsState, dynamics: PrimitiveEquations) -> DynamicsState:
    """physics_interface.py:237-273 on the GPU (the geopotential is not read)."""
    vor, div, tm, lsp, qm = _nodal_to_spectral(physics_state.u_wind, physics_state.v_wind, physics_state.temperature,
                                               physics_state.specific_humidity,
                                               physics_state.normalized_surface_pressure, dynamics, False)
    return DynamicsState(vor, div, tm, lsp, {"specific_humidity": qm})


def physics_tendency_to_dynamics_tendency(physics_tendency: PhysicsTendency,
                                          dynamics: PrimitiveEquations) -> DynamicsState:
    """physics_interface.py:275-303 on the GPU: the log-surface-pressure
    tendency is zeros shaped like one level of the temperature tendency."""
    vor, div, tm, lsp, qm = _nodal_to_spectral(physics_tendency.u_wind, physics_tendency.v_wind,
                                               physics_tendency.temperature, physics_tendency.specific_humidity,
                                               None, dynamics, True)
    return DynamicsState(vor, div, tm, lsp, {"specific_humidity": qm}, sim_time=0.0)
