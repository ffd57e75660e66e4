"""SpeedyPhysics on MI355X: drop-in for jcm/physics/speedy/speedy_physics.py:34-133.

``SpeedyPhysics.compute_tendencies(state, forcing, geometry, date)`` keeps the
reference signature and return value ``(PhysicsTendency, data)``
(physics_interface.py:138); the eleven terms run fused in HIP kernels
(libspeedy_hip.so, C ABI include/speedy_abi.h).  ``data`` is the PhysicsData
subset the kernel writes when the object is built with
``with_diagnostics=True`` (otherwise None, which avoids ~560 B/column of
writes).  There is no CPU path: without the library or a GPU the call raises.
"""
from __future__ import annotations

import copy
import ctypes
from types import SimpleNamespace
from typing import Any, Tuple

import numpy as np
import torch

from . import _native
from .date import DateData
from .forcing import FORCING_FIELDS, ForcingData
from .geometry import Geometry
from .params import VJP_PARAMS, Parameters
from .physics_interface import Physics, PhysicsState, PhysicsTendency

STATE_FIELDS = ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential")
VERIFY_FLAGS = _native.FLAG_VERIFY_STATE | _native.FLAG_VERIFY_TENDENCIES


# torch's raw current-stream query (the one its compiled kernels use): about 0.2 us,
# against about 2 us for torch.cuda.current_stream(device).cuda_stream
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_STATE_FIELDS = (("u_wind", "u"), ("v_wind", "v"), ("temperature", "t"), ("specific_humidity", "q"),
                 ("geopotential", "phi"))


def _device(x, dev) -> torch.Tensor:
    """float32 contiguous tensor on `dev`; tensors already there are used as is."""
    if isinstance(x, torch.Tensor):
        if x.device == dev and x.dtype == torch.float32 and x.is_contiguous():
            return x
        return x.to(device=dev, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, np.float32), device=dev)


def climatology_rows(arr, dev) -> torch.Tensor:
    """A forcing climatology (ix, il, ndays) laid out as [ndays][ix * il] on `dev`:
    row d is arr[:, :, d] in the kernels' column order (c = i * il + j), so the day
    slice of a call (speedy_physics.py:119-121) is one contiguous row."""
    t = arr if isinstance(arr, torch.Tensor) else torch.as_tensor(np.asarray(arr, np.float32))
    if t.dim() != 3:
        raise ValueError(f"a climatology is (ix, il, ndays), got shape {tuple(t.shape)}")
    ix, il, nd = t.shape
    return t.to(device=dev, dtype=torch.float32).permute(2, 0, 1).contiguous().view(nd, ix * il)


class PhysicsData:
    """The PhysicsData subset written by the kernel, viewed per struct:
    ``data.shortwave_rad.rsds``, ``data.mod_radcon.tau2`` (kx, ix, il, 4) etc.
    Field names are those of jcm/physics/speedy/physics_data.py."""

    def __init__(self, diag: torch.Tensor, names, kx: int, hshape, date: DateData | None = None):
        self._diag = diag          # (rows, ncol) or (nmember, rows, ncol)
        self._names = names
        self._kx = kx
        self._hshape = tuple(hshape)
        self._index = {n: i for i, n in enumerate(names)}
        self.date = date           # PhysicsData.date (physics_data.py), read by the term calls

    def _row(self, name):
        i = self._index[name]
        d = self._diag[..., i, :]
        return d.reshape(d.shape[:-1] + self._hshape)

    def to_dict(self) -> dict:
        """Flat {"struct.field[.channel]": array} like Physics.data_struct_to_dict."""
        out = {}
        for n in self._names:
            if n.endswith("]"):
                base = n[:n.index("[")]
                if base not in out:
                    out[base] = torch.stack([self._row(f"{base}[{k}]") for k in range(self._kx)], dim=-3)
            else:
                out[n] = self._row(n)
        return out

    def __getattr__(self, group):
        if group.startswith("_"):
            raise AttributeError(group)
        d = self.to_dict()
        sub = {}
        for key, val in d.items():
            parts = key.split(".")
            if parts[0] != group:
                continue
            if len(parts) == 2:
                sub[parts[1]] = val
            else:
                sub.setdefault(parts[1], {})[int(parts[2])] = val
        if not sub:
            raise AttributeError(group)
        ns = {}
        for k, v in sub.items():
            ns[k] = torch.stack([v[i] for i in sorted(v)], dim=-1) if isinstance(v, dict) else v
        return SimpleNamespace(**ns)


class SpeedyTerm:
    """One SPEEDY term with the reference's call shape,
    ``term(state, physics_data, parameters, forcing, geometry) -> (PhysicsTendency,
    PhysicsData)`` (speedy_physics.py:123-125), on the GPU.

    The kernels run the term chain fused, so a call runs the chain on ``state``
    with ``parameters``, ``forcing`` and ``geometry`` at the date of
    ``physics_data`` (``physics_data.date``, or the ``date`` keyword) and
    returns:
      * this term's own tendency: the value the full compute_tendencies adds for
        it (term_mask = 1 << index), bit for bit;
      * the PhysicsData the chain writes (the library's diagnostic subset), which
        holds every term's outputs, those of the terms after this one included.
    The fields of the incoming ``physics_data`` are not read: the terms before
    this one are recomputed from ``state``, which equals the reference's result
    whenever ``physics_data`` came from those terms on the same inputs (the way
    compute_tendencies chains them)."""

    def __init__(self, physics: "SpeedyPhysics", index: int, name: str):
        self.physics, self.index, self.__name__ = physics, index, name

    def __repr__(self):
        return f"SpeedyTerm({self.index}, {self.__name__})"

    def __call__(self, state: PhysicsState, physics_data, parameters: Parameters, forcing: ForcingData,
                 geometry: Geometry, date: DateData | None = None):
        date = date if date is not None else getattr(physics_data, "date", None)
        if date is None:
            raise ValueError(f"{self.__name__}: pass physics_data with a date (get_empty_data(geometry, date)) "
                             "or the date keyword")
        # a shallow copy with these parameters and diagnostics on: it shares the
        # workspaces and caches of the physics object, which stays unchanged
        q = copy.copy(self.physics)
        q.with_diagnostics = True
        if parameters is not None and parameters is not q.parameters:
            q.parameters, q._abi_params, q._abi_params_src = parameters, None, None
        return q.compute_tendencies(state, forcing, geometry, date, term_mask=1 << self.index)


class DiagnosticsCollector:
    """DiagnosticsCollector (model.py:46-64) for the PhysicsData subset this library
    writes, with the averages on the device.  ``data`` holds the stacked averages of
    the outputs (init_data stacks an empty data struct, as model.py:88-89 does);
    accumulate_if_physical_step adds new_data / steps_to_average into output ``i``
    once per physics step (a HIP kernel, speedy_diag_accumulate, in the reference's
    float32 arithmetic) and clears ``physical_step``, which the caller sets again
    before each step (model.py:103)."""

    def __init__(self, steps_to_average: int):
        self.i = 0
        self.physical_step = True
        self.steps_to_average = steps_to_average
        self.data = None

    def init_data(self, empty_data: PhysicsData, outer_steps: int) -> None:
        d = empty_data._diag
        stacked = d.unsqueeze(0).repeat((outer_steps,) + (1,) * d.dim()).contiguous()
        self.data = PhysicsData(stacked, empty_data._names, empty_data._kx, empty_data._hshape)

    def accumulate_if_physical_step(self, new_data: PhysicsData) -> None:
        if not self.physical_step:
            return
        if self.data is None:
            raise _native.SpeedyError("DiagnosticsCollector: init_data first")
        src, dst = new_data._diag, self.data._diag[self.i]
        if tuple(src.shape) != tuple(dst.shape) or src.device != dst.device or not src.is_contiguous():
            raise ValueError(f"diagnostics of shape {tuple(src.shape)} do not match the averages {tuple(dst.shape)}")
        stream = ctypes.c_void_p(torch.cuda.current_stream(src.device).cuda_stream)
        _native.check(_native.load().speedy_diag_accumulate(src.data_ptr(), dst.data_ptr(), src.numel(),
                                                            float(self.steps_to_average), stream),
                      "speedy_diag_accumulate")
        self.physical_step = False


class SpeedyPhysics(Physics):
    """SPEEDY physics (speedy_physics.py:34-133) fused into gfx950 kernels."""

    def __init__(self, parameters: Parameters | None = None, checkpoint_terms: bool = True,
                 with_diagnostics: bool = False, device: str | torch.device | None = None,
                 plain_launch: bool = False, check_barrier: bool = False):
        """plain_launch: uncaptured single-launch calls use a plain launch instead of a
        cooperative one (about 10 us less per call at T85).  Only for a process that
        has the GPU to itself: the grid-wide level-set exchange then relies on all
        workgroups being resident together.  barrier_status() reports a call whose
        exchange gave up (its results are invalid) and resets the workspace.
        check_barrier: after every call, synchronize and raise SpeedyError if the
        call's level-set exchange gave up, resetting the workspace (costs a
        synchronization per call)."""
        self.parameters = parameters if parameters is not None else Parameters.default()
        self.with_diagnostics = with_diagnostics
        self._launch_flags = _native.FLAG_PLAIN_LAUNCH if plain_launch else 0
        self.check_barrier = check_barrier
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device()) \
            if torch.cuda.is_available() else None
        self._lib = _native.load()
        self._workspaces: dict = {}
        self._consts: dict = {}
        self._abi_params = None
        self._abi_params_src = None
        self._verts: dict = {}
        self._fg: dict = {}
        self._climates: dict = {}
        self.last_budgets = None
        self._gpu_ok = False

    # ------------------------------------------------------------------ helpers
    def _require_gpu(self):
        if self._gpu_ok:
            return
        if self.device is None or self.device.type != "cuda" or not torch.cuda.is_available():
            raise _native.SpeedyError("SpeedyPhysics needs a ROCm GPU; there is no CPU path")
        self._gpu_ok = True

    def _params(self) -> _native.SpeedyParams:
        """parameters.to_abi(), rebuilt only when a value changed (also in place)."""
        key = self.parameters.abi_key()
        if self._abi_params is None or self._abi_params_src != key:
            self._abi_params = self.parameters.to_abi()
            self._abi_params_src = key
        return self._abi_params

    def _vertical(self, geometry: Geometry) -> _native.SpeedyVertical:
        """geometry.to_abi(), rebuilt only for another geometry or changed levels."""
        key = geometry.vertical_key()
        hit = self._verts.get(id(geometry))
        if hit is None or hit[0] is not geometry or hit[1] != key:
            if len(self._verts) > 64:
                self._verts.clear()
            hit = self._verts[id(geometry)] = (geometry, key, geometry.to_abi())
        return hit[2]

    def workspace(self, nlon: int, nlat: int, nmember: int, stream=None) -> torch.Tensor:
        """The call workspace of this grid on `stream` (the current stream when None).
        Workspaces are per (grid, stream): calls in flight on different streams never
        share one (speedy_abi.h: concurrent calls need distinct workspaces)."""
        s = stream if stream is not None else self._stream()
        key = (nlon, nlat, nmember, s.value)
        ws = self._workspaces.get(key)
        if ws is None:
            nbytes = self._lib.speedy_workspace_bytes(nlon, nlat, nmember)
            if nbytes == 0:
                raise _native.SpeedyError(f"grid {nlon}x{nlat}x{nmember} is not supported")
            if torch.cuda.is_current_stream_capturing():
                # its initialising fill would be captured and replayed with every call
                raise _native.SpeedyError("a new workspace is needed during graph capture: call once on this "
                                          "stream and grid outside capture first")
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            _native.check(self._lib.speedy_workspace_init(ws.data_ptr(), nbytes, s), "speedy_workspace_init")
            self._workspaces[key] = ws
        return ws

    def _stream(self):
        if _RAW_STREAM is not None:
            idx = self.device.index
            return ctypes.c_void_p(_RAW_STREAM(torch.cuda.current_device() if idx is None else idx))
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _climatology(self, arr) -> torch.Tensor:
        """climatology_rows(arr) on the device, made once per array (per version of a
        tensor, so in-place updates are seen); the calls then read the day's row."""
        key = id(arr)
        ver = arr._version if isinstance(arr, torch.Tensor) else None
        hit = self._climates.get(key)
        if hit is not None and hit[0] is arr and hit[1] == ver:
            return hit[2]
        rows = climatology_rows(arr, self.device)
        if len(self._climates) > 16:
            self._climates.clear()
        self._climates[key] = (arr, ver, rows)
        return rows

    def _const(self, arr) -> torch.Tensor:
        """Device copy of a geometry/forcing array, cached while the array lives."""
        if isinstance(arr, torch.Tensor):
            return _device(arr, self.device)
        key = id(arr)
        hit = self._consts.get(key)
        if hit is not None and hit[0] is arr:
            return hit[1]
        t = _device(arr, self.device)
        if len(self._consts) > 256:
            self._consts.clear()
        self._consts[key] = (arr, t)
        return t

    def _fields(self, state: PhysicsState, forcing: ForcingData, geometry: Geometry, date: DateData,
                keep: list):
        kx, ix, il = geometry.nodal_shape
        t = _device(state.temperature, self.device)
        if t.shape[-3:] != (kx, ix, il):
            raise ValueError(f"state shape {tuple(t.shape)} does not match geometry {geometry.nodal_shape}")
        nmember = 1 if t.dim() == 3 else int(np.prod(t.shape[:-3]))
        f = _native.SpeedyFields()
        for src, dst in _STATE_FIELDS:
            x = _device(getattr(state, src), self.device)
            if x.shape != t.shape:
                raise ValueError(f"{src} has shape {tuple(x.shape)}, expected {tuple(t.shape)}")
            keep.append(x)
            setattr(f, dst, x.data_ptr())
        ps = _device(state.normalized_surface_pressure, self.device)
        if ps.numel() != nmember * ix * il:
            raise ValueError("normalized_surface_pressure has the wrong size")
        keep.append(ps)
        f.ps = ps.data_ptr()
        for dst, x in self._forcing_fields(forcing, geometry, date.model_day(), ix, il):
            keep.append(x)
            setattr(f, dst, x.data_ptr())
        return f, t, nmember

    def _forcing_fields(self, forcing: ForcingData, geometry: Geometry, day: int, ix: int, il: int):
        """(field, device tensor) of the forcing at `day` and the geometry.  Cached for
        the same forcing and geometry objects and day when every array is used by
        identity (NumPy arrays through _const's and _climatology's caches, device
        tensors as they are); broadcasts, host tensors and tensor climatologies (whose
        version is checked) are looked up again on every call."""
        key = (id(forcing), id(geometry), day, ix, il)
        hit = self._fg.get(key)
        if hit is not None and hit[0] is forcing and hit[1] is geometry:
            return hit[2]
        out, by_identity = [], True
        fmap = dict(zip(FORCING_FIELDS, ("alb0", "sice_am", "snowc_am", "soilw_am", "stl_am", "sst")))
        srcs = [(dst, getattr(forcing, src)) for src, dst in fmap.items()]
        srcs += [(src, getattr(geometry, src)) for src in ("fmask", "orog", "phis0", "sia", "coa")]
        for i, (dst, src) in enumerate(srcs):
            arr = src
            if i < len(fmap) and arr.ndim > 2:
                # a climatology: the day's row of its [day][ncol] layout (made once);
                # the day is clamped like the reference's gather
                if tuple(arr.shape[:2]) != (ix, il):
                    raise ValueError(f"forcing {dst} has shape {tuple(arr.shape)}, expected ({ix}, {il}, ndays)")
                rows = self._climatology(arr)
                by_identity &= not isinstance(arr, torch.Tensor)
                out.append((dst, rows[min(max(day, 0), rows.shape[0] - 1)]))
                continue
            if i < len(fmap) and isinstance(arr, np.ndarray) and arr.shape != (ix, il):
                arr = np.broadcast_to(arr, (ix, il))
            x = self._const(arr)
            by_identity &= arr is src and (isinstance(arr, np.ndarray) or x is arr)
            out.append((dst, x))
        if by_identity:
            if len(self._fg) > 64:
                self._fg.clear()
            self._fg[key] = (forcing, geometry, out)
        return out

    # ------------------------------------------------------------------ API
    def compute_tendencies(self, state: PhysicsState, forcing: ForcingData, geometry: Geometry,
                           date: DateData, *, term_mask: int = _native.TERMS_ALL,
                           level_set: int = -1, verify: bool = False,
                           budgets: bool = False) -> Tuple[PhysicsTendency, Any]:
        """SpeedyPhysics.compute_tendencies (speedy_physics.py:88-127).

        term_mask: bit i keeps term i's tendency in the sum (speedy_physics.py:67-80 order);
        level_set: >= 0 overrides the Q1 cloud-top level set (parity tests);
        verify: also apply verify_state before and verify_tendencies after, with
        dt = date.dt_seconds, as get_physical_tendencies does (physics_interface.py:371-375),
        in the same kernel;
        budgets: also reduce the float64 column energy and water budgets (speedy_abi.h,
        SPEEDY_NBUDGET) into ``self.last_budgets``: a float64 tensor (nmember, 6) whose
        columns are named by ``_native.BUDGET_NAMES``."""
        self._require_gpu()
        keep: list = []
        f, t, nmember = self._fields(state, forcing, geometry, date, keep)
        kx, ix, il = geometry.nodal_shape
        out = [torch.empty_like(t) for _ in range(4)]
        f.du, f.dv, f.dt, f.dq = (o.data_ptr() for o in out)
        diag = None
        if self.with_diagnostics:
            rows = self._lib.speedy_diag_rows(kx)
            diag = torch.empty(t.shape[:-3] + (rows, ix * il), dtype=torch.float32, device=self.device)
            f.diag = diag.data_ptr()
        else:
            f.diag = None
        if budgets:
            self.last_budgets = torch.empty((nmember, _native.NBUDGET), dtype=torch.float64, device=self.device)
            f.budget = self.last_budgets.data_ptr()
        stream = self._stream()
        ws = self.workspace(ix, il, nmember, stream)
        flags = ((_native.FLAG_LFLUXLAND if bool(forcing.lfluxland) else 0) | (VERIFY_FLAGS if verify else 0)
                 | self._launch_flags)
        rc = self._lib.speedy_fwd(ctypes.byref(f), ctypes.byref(self._vertical(geometry)), ctypes.byref(self._params()),
                                  ctypes.byref(date.to_abi()), ix, il, nmember, flags, term_mask, level_set,
                                  ws.data_ptr(), ws.numel(), stream)
        _native.check(rc, "speedy_fwd")
        if self.check_barrier:
            self._raise_on_barrier(ix, il, nmember, stream)
        data = PhysicsData(diag, _native.diag_names(kx), kx, (ix, il), date) if diag is not None else None
        return PhysicsTendency(*out), data

    @property
    def terms(self) -> list:
        """The eleven terms in order (speedy_physics.py:67-80), each callable as the
        reference's ``term(state, physics_data, parameters, forcing, geometry)``
        (speedy_physics.py:123-125); see SpeedyTerm."""
        return [SpeedyTerm(self, i, name) for i, name in enumerate(_native.TERMS)]

    def prepare_vjp(self, state: PhysicsState, forcing: ForcingData, geometry: Geometry, date: DateData,
                    cotangent: PhysicsTendency, keep: list):
        """C structs of one fused forward + VJP call and its freshly allocated outputs:
        (SpeedyFields, SpeedyAdjoint, outputs, (nlon, nlat, nmember), flags).  `keep`
        collects the tensors the call reads, which must outlive it."""
        f, t, nmember = self._fields(state, forcing, geometry, date, keep)
        kx, ix, il = geometry.nodal_shape
        out = [torch.empty_like(t) for _ in range(4)]
        f.du, f.dv, f.dt, f.dq = (o.data_ptr() for o in out)
        f.diag = None
        adj = _native.SpeedyAdjoint()
        for name, src in zip(("ct_du", "ct_dv", "ct_dt", "ct_dq"),
                             ("u_wind", "v_wind", "temperature", "specific_humidity")):
            x = _device(getattr(cotangent, src), self.device)
            if x.shape != t.shape:
                raise ValueError(f"cotangent {src} has the wrong shape")
            keep.append(x)
            setattr(adj, name, x.data_ptr())
        grads = [torch.empty_like(t) for _ in range(4)]
        g_ps = torch.empty(t.shape[:-3] + (ix, il), dtype=torch.float32, device=self.device)
        g_par = torch.empty(t.shape[:-3] + (len(VJP_PARAMS),), dtype=torch.float32, device=self.device)
        adj.g_u, adj.g_v, adj.g_t, adj.g_q = (g.data_ptr() for g in grads)
        adj.g_ps = g_ps.data_ptr()
        adj.g_params = g_par.data_ptr()
        flags = _native.FLAG_LFLUXLAND if bool(forcing.lfluxland) else 0
        outputs = dict(tendency=PhysicsTendency(*out),
                       state_grads=PhysicsState(grads[0], grads[1], grads[2], grads[3], None, g_ps),
                       param_grads={f"{g}.{n}": g_par[..., i] for i, (g, n) in enumerate(VJP_PARAMS)})
        return f, adj, outputs, (ix, il, nmember), flags

    def prepare_fwd(self, state: PhysicsState, forcing: ForcingData, geometry: Geometry, date: DateData, keep: list):
        """As prepare_vjp for a forward call without diagnostics."""
        f, t, nmember = self._fields(state, forcing, geometry, date, keep)
        kx, ix, il = geometry.nodal_shape
        out = [torch.empty_like(t) for _ in range(4)]
        f.du, f.dv, f.dt, f.dq = (o.data_ptr() for o in out)
        f.diag = None
        flags = _native.FLAG_LFLUXLAND if bool(forcing.lfluxland) else 0
        return f, dict(tendency=PhysicsTendency(*out)), (ix, il, nmember), flags

    def compute_tendencies_vjp(self, state: PhysicsState, forcing: ForcingData, geometry: Geometry,
                               date: DateData, cotangent: PhysicsTendency, *, level_set: int = -1,
                               verify: bool = False):
        """Fused forward + VJP of compute_tendencies: returns (tendency, state_grads, param_grads)
        where state_grads is a PhysicsState of gradients (geopotential entry None) and
        param_grads maps the five VJP parameters (params.VJP_PARAMS) to per-member sums.
        verify: as in compute_tendencies (the gradients then include the two steps)."""
        self._require_gpu()
        keep: list = []
        f, adj, outputs, (ix, il, nmember), flags = self.prepare_vjp(state, forcing, geometry, date, cotangent, keep)
        flags |= (VERIFY_FLAGS if verify else 0) | self._launch_flags
        stream = self._stream()
        ws = self.workspace(ix, il, nmember, stream)
        rc = self._lib.speedy_fwd_vjp(ctypes.byref(f), ctypes.byref(adj), ctypes.byref(self._vertical(geometry)),
                                      ctypes.byref(self._params()), ctypes.byref(date.to_abi()), ix, il, nmember,
                                      flags, level_set, ws.data_ptr(), ws.numel(), stream)
        _native.check(rc, "speedy_fwd_vjp")
        if self.check_barrier:
            self._raise_on_barrier(ix, il, nmember, stream)
        return outputs["tendency"], outputs["state_grads"], outputs["param_grads"]

    def _raise_on_barrier(self, nlon, nlat, nmember, stream):
        ws = self.workspace(nlon, nlat, nmember, stream)
        st = ctypes.c_uint32(0)
        _native.check(self._lib.speedy_workspace_status(ws.data_ptr(), ws.numel(), ctypes.byref(st), stream),
                      "speedy_workspace_status")
        if st.value:
            raise _native.SpeedyError("the call's level-set exchange gave up (not all workgroups were resident); "
                                      "its results are invalid and the workspace has been reset")

    def level_sets(self, nlon: int, nlat: int, nmember: int = 1) -> list:
        """Q1 level set of each member in the last call on this grid and the current stream."""
        ws = self.workspace(nlon, nlat, nmember)
        buf = (ctypes.c_uint32 * nmember)()
        _native.check(self._lib.speedy_read_level_sets(ws.data_ptr(), ws.numel(), nmember, buf, self._stream()),
                      "speedy_read_level_sets")
        return list(buf)

    def barrier_status(self, nlon: int, nlat: int, nmember: int = 1) -> int:
        # (a set flag also re-zeroes the workspace: see speedy_workspace_status)
        """Nonzero if a call on this grid and the current stream since the last query
        gave up waiting for the level set (its results are invalid); reading clears
        the flag."""
        ws = self.workspace(nlon, nlat, nmember)
        st = ctypes.c_uint32(0)
        _native.check(self._lib.speedy_workspace_status(ws.data_ptr(), ws.numel(), ctypes.byref(st), self._stream()),
                      "speedy_workspace_status")
        return int(st.value)

    def get_empty_data(self, geometry: Geometry, date: DateData | None = None):
        """speedy_physics.py:129-133: an all-zero data struct for accumulating averages
        (``date`` as PhysicsData.zeros takes it)."""
        kx, ix, il = geometry.nodal_shape
        rows = self._lib.speedy_diag_rows(kx)
        z = torch.zeros((rows, ix * il), dtype=torch.float32, device=self.device)
        return PhysicsData(z, _native.diag_names(kx), kx, (ix, il), date)
