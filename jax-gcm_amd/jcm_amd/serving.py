"""Continuous-batching request loop for SPEEDY physics on one GPU.

A request is one forecast member's physics evaluation: its device-resident
state (torch tensors on the GPU, owned by the client between requests), its
forcing, geometry and date, and the kind -- forward (tendencies), or forward +
VJP when a cotangent is given.  Members may have different resolutions (T21,
T31, T42, T85, ...).  The loop:

  * admits waiting requests in arrival order while their HBM footprint (the
    outputs plus the call workspace; states are already resident) fits in the
    budget beside every request on the device, and while the batch stays under
    a column and a segment limit;
  * launches each admitted group as ONE batched call (``speedy_fwd_batch``:
    level sets, one column kernel over all of the batch's columns, parameter
    sums) on one of ``n_streams`` HIP streams, so a batch can start while the
    previous one still runs and their work overlaps on the device;
  * completes a request when its batch's event has fired, releasing its budget
    so that waiting requests are admitted into the next batch.

Latency is measured on the host clock from the request's arrival to the
moment the loop observes its batch's completion.  ``batching=False`` runs the
same stream of requests with one call per request (the plain per-call path) for
comparison.

No counterpart in the reference (jcm calls the physics once per model step).
"""
from __future__ import annotations

import ctypes
import math
import operator
import time
import weakref
from collections import deque
from dataclasses import dataclass, field
from typing import Any, Callable, Optional

import numpy as np
import torch

from . import _native
from .forcing import FORCING_FIELDS
from .params import VJP_PARAMS
from .physics_interface import PhysicsState, PhysicsTendency
from .speedy_physics import SpeedyPhysics

KIND_FWD = "fwd"
KIND_VJP = "fwd_vjp"
_VJP_NAMES = tuple(f"{g}.{n}" for g, n in VJP_PARAMS)
_STATE_NAMES = ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential", "normalized_surface_pressure")


@dataclass
class Request:
    """One member's physics call.  Results are filled in when it completes: views
    of the output block its batch shares with the batch's other requests, so a
    result that is kept keeps that block allocated (copy it to keep one result
    alone)."""
    state: PhysicsState
    forcing: Any
    geometry: Any
    date: Any
    cotangent: Optional[PhysicsTendency] = None
    arrival: float = 0.0          # seconds on the server clock
    rid: int = 0
    # filled by the server
    level_set: Optional[list] = None   # per member, when the server records level sets
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    batch: int = -1
    # (tendency, state_grads, param_grads), or the batch output block they are views of
    _results: Any = field(default=None, repr=False)

    @property
    def kind(self) -> str:
        return KIND_VJP if self.cotangent is not None else KIND_FWD

    @property
    def columns(self) -> int:
        c = self.__dict__.get("_ncols")
        if c is None:
            kx, ix, il = self.geometry.nodal_shape
            t = self.state.temperature
            nmember = 1 if len(t.shape) == 3 else int(np.prod(t.shape[:-3]))
            c = self._ncols = ix * il * nmember
        return c

    def _get(self, i):
        r = self._results
        if isinstance(r, _Outputs):
            self._results = r = r.views()
        return None if r is None else r[i]

    def _set(self, i, v):
        self._get(0)   # a lazy block becomes its views first
        r = list(self._results) if self._results is not None else [None, None, None]
        r[i] = v
        self._results = tuple(r)

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.arrival

    # the outputs: views of the batch's output block, made when first read
    tendency = property(lambda self: self._get(0), lambda self, v: self._set(0, v))
    state_grads = property(lambda self: self._get(1), lambda self, v: self._set(1, v))
    param_grads = property(lambda self: self._get(2), lambda self, v: self._set(2, v))


class _Outputs:
    """One request's outputs inside a batch's output block (float32, device):
    du, dv, dT, dq [, dL/du, dL/dv, dL/dT, dL/dq] of the state's shape, then for
    the VJP dL/dps and the five parameter sums per member."""
    __slots__ = ("buf", "off", "tshape", "vjp")

    def __init__(self, buf, off, tshape, vjp):
        self.buf, self.off, self.tshape, self.vjp = buf, off, tshape, vjp

    def views(self):
        tshape, n3 = self.tshape, math.prod(self.tshape)
        lead, (ix, il) = tshape[:-3], tshape[-2:]
        nout = 8 if self.vjp else 4
        blk = self.buf[self.off:self.off + nout * n3].view((nout,) + tshape).unbind(0)
        tend = PhysicsTendency(*blk[:4])
        if not self.vjp:
            return tend, None, None
        n2 = n3 // tshape[-3]
        o = self.off + 8 * n3
        g_ps = self.buf[o:o + n2].view(lead + (ix, il))
        g_par = self.buf[o + n2:o + n2 + 5 * (n2 // (ix * il))].view(lead + (5,))
        return tend, PhysicsState(blk[4], blk[5], blk[6], blk[7], None, g_ps), dict(zip(_VJP_NAMES, g_par.unbind(-1)))


def _output_floats(tshape, vjp: bool) -> int:
    """Floats of one request's outputs (see _Outputs), rounded up to 64 (256 bytes)."""
    n3 = math.prod(tshape)
    n = (8 * n3 + n3 // tshape[-3] + 5 * (n3 // (tshape[-3] * tshape[-2] * tshape[-1]))) if vjp else 4 * n3
    return (n + 63) // 64 * 64


def request_footprint(req: Request, kx: int, workspace_bytes: int) -> int:
    """HBM bytes a request holds while it is on the device: its outputs
    (tendencies; for the VJP also state and parameter gradients) and its call
    workspace.  States and cotangents are resident before admission."""
    ix, il = req.geometry.nodal_shape[1:]
    nmember = req.columns // (ix * il)
    n3 = kx * req.columns * 4                          # one 3-D float field of the request
    out = 4 * n3                                       # du, dv, dT, dq
    if req.kind == KIND_VJP:
        out += 4 * n3 + req.columns * 4 + nmember * 5 * 4   # dL/d(u, v, T, q), dL/dps, parameter sums
    return out + workspace_bytes


@dataclass
class _Slot:
    stream: Any
    desc: Optional[torch.Tensor] = None       # device descriptor buffer
    staging: Optional[torch.Tensor] = None    # pinned host staging
    event: Any = None                         # recorded after each batch, reused
    ready: Any = None                         # the client stream's point a batch waits for, reused
    batch: list = field(default_factory=list)
    keep: list = field(default_factory=list)
    busy: bool = False


class PhysicsServer:
    """Continuous-batching server over one GPU (see the module docstring)."""

    def __init__(self, physics: SpeedyPhysics | None = None, hbm_budget: int = 8 << 30,
                 max_batch_columns: int = 1 << 20, max_segments: int = 64, n_streams: int = 2,
                 batching: bool = True, clock: Callable[[], float] = time.perf_counter,
                 record_level_sets: bool = False):
        self.phys = physics or SpeedyPhysics()
        self.dev = self.phys.device
        self._dev_index = self.dev.index if self.dev.index is not None else torch.cuda.current_device()
        self.lib = self.phys._lib
        self.hbm_budget = int(hbm_budget)
        self.max_batch_columns = int(max_batch_columns)
        self.max_segments = min(int(max_segments), _native.BATCH_MAX_SEGMENTS)
        self.batching = batching
        self.clock = clock
        self.record_level_sets = record_level_sets   # read each request's Q1 level sets at completion
        self.slots = [_Slot(stream=torch.cuda.Stream(device=self.dev)) for _ in range(max(1, n_streams))]
        self.queue: deque = deque()
        self.in_flight_bytes = 0
        self.peak_bytes = 0
        self.done: list = []
        self.n_batches = 0
        self.batch_sizes: list = []
        self._ws_free: dict = {}      # (nlon, nlat, nmember) -> [workspace tensors]
        self._ws_bytes: dict = {}
        self._members: dict = {}
        self._verts: dict = {}

    # ------------------------------------------------------------ resources
    def _workspace(self, key, stream) -> torch.Tensor:
        free = self._ws_free.setdefault(key, [])
        if free:
            return free.pop()
        nbytes = self.lib.speedy_workspace_bytes(*key)
        if nbytes == 0:
            raise _native.SpeedyError(f"grid {key} is not supported")
        ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        ws.record_stream(stream)
        _native.check(self.lib.speedy_workspace_init(ws.data_ptr(), nbytes, ctypes.c_void_p(stream.cuda_stream)),
                      "speedy_workspace_init")
        return ws

    def _wait_client(self, slot: _Slot):
        """Order the batch after the work the client's stream holds.  When that
        stream has finished everything (the usual case: member states written
        long before the request), there is nothing to wait for and the event
        record + stream wait (about 20 us of host time) is skipped."""
        cur = torch.cuda.current_stream(self.dev)
        if cur.query():
            return
        if slot.ready is None:
            slot.ready = torch.cuda.Event()
        slot.ready.record(cur)
        slot.stream.wait_event(slot.ready)

    def footprint(self, req: Request) -> int:
        fp = getattr(req, "_fp", None)
        if fp is None:
            kx, ix, il = req.geometry.nodal_shape
            key = (ix, il, req.columns // (ix * il))
            wsb = self._ws_bytes.get(key)
            if wsb is None:
                wsb = self._ws_bytes[key] = self.lib.speedy_workspace_bytes(*key)
            fp = req._fp = request_footprint(req, kx, wsb)
        return fp

    def _member(self, req: Request):
        """Input half of the call structs for a member's (state, forcing, geometry, day).
        Built once and cached only when every state field is a contiguous float32
        tensor on the device: the kernels then read the client's memory at launch, so
        in-place updates are seen; each hit checks that the state still holds the
        same tensors at the same addresses.  States with NumPy or host fields, and
        forcings with tensor climatologies, are rebuilt on every request (their
        device copies are made per call).  Forcing and geometry NumPy arrays are
        treated as constants, like the reference's JAX arrays."""
        st = req.state
        fields = tuple(getattr(st, n) for n in _STATE_NAMES)
        key = (id(st), id(req.forcing), id(req.geometry), req.date.model_day())
        hit = self._members.get(key)
        if hit is not None:
            f, tshape, grid, flags, keep, refs, objs, ptrs = hit
            if not (refs[0]() is st and refs[1]() is req.forcing and refs[2]() is req.geometry
                    and all(a is b for a, b in zip(objs, fields))
                    and all(a.data_ptr() == p for a, p in zip(fields, ptrs))):
                hit = None   # an id reused by a new object, or a field replaced
        if hit is not None:
            return hit
        keep: list = []
        f, t, nm = self.phys._fields(st, req.forcing, req.geometry, req.date, keep)
        kx, ix, il = req.geometry.nodal_shape
        flags = _native.FLAG_LFLUXLAND if bool(req.forcing.lfluxland) else 0
        cacheable = all(isinstance(x, torch.Tensor) and x.dtype is torch.float32 and x.is_contiguous()
                        and x.get_device() == self._dev_index for x in fields)
        cacheable &= not any(isinstance(getattr(req.forcing, n, None), torch.Tensor)
                             and getattr(req.forcing, n).dim() > 2 for n in FORCING_FIELDS)
        hit = (f, tuple(t.shape), (ix, il, nm), flags, keep, None, fields, tuple(x.data_ptr() for x in fields)
               if cacheable else ())
        if cacheable:
            # weak references: the entry (and the device tensors it keeps) goes
            # when the client drops the state, so states that are new with every
            # request are not held here
            members = self._members
            drop = lambda _ref, k=key, d=members: d.pop(k, None)
            refs = (weakref.ref(st, drop), weakref.ref(req.forcing), weakref.ref(req.geometry))
            hit = hit[:5] + (refs,) + hit[6:]
            if len(members) > 4096:
                members.clear()
            members[key] = hit
        return hit

    def _prepare(self, req: Request, seg, buf: torch.Tensor, off: int):
        """Fill one batch segment for req, its outputs at float offset `off` of the
        batch's output block (views are made when the client first reads them)."""
        f_in, tshape, (ix, il, nm), flags, member_keep = self._member(req)[:5]
        seg.fields = f_in
        f = seg.fields                     # a view of the segment's memory
        n3 = math.prod(tshape)
        vjp = req.cotangent is not None
        base, step = buf.data_ptr() + 4 * off, 4 * n3
        f.du, f.dv, f.dt, f.dq = base, base + step, base + 2 * step, base + 3 * step
        if vjp:
            adj = seg.adj
            adj.ct_du, adj.ct_dv, adj.ct_dt, adj.ct_dq = self._cotangent_ptrs(req.cotangent, tshape, req)
            n2 = nm * ix * il
            adj.g_u, adj.g_v, adj.g_t, adj.g_q = base + 4 * step, base + 5 * step, base + 6 * step, base + 7 * step
            adj.g_ps, adj.g_params = base + 8 * step, base + 8 * step + 4 * n2
        req._results = _Outputs(buf, off, tshape, vjp)
        return (ix, il, nm), flags, member_keep

    def _cotangent_ptrs(self, ct: PhysicsTendency, tshape, req: Request | None = None) -> tuple:
        """Device addresses of a cotangent's four fields, after checking them.  The
        check made at submit is kept on the request and reused at launch while the
        cotangent still holds the same tensors at the same addresses."""
        cts = (ct.u_wind, ct.v_wind, ct.temperature, ct.specific_humidity)
        hit = getattr(req, "_ct", None) if req is not None else None
        if hit is not None and all(map(operator.is_, hit[0], cts)):
            ptrs = tuple(c.data_ptr() for c in cts)
            if ptrs == hit[1]:
                return ptrs
        for c in cts:
            if not (isinstance(c, torch.Tensor) and c.dtype is torch.float32 and c.shape == tshape
                    and c.get_device() == self._dev_index and c.is_contiguous()):
                raise ValueError("cotangents must be contiguous float32 device tensors of the state's shape")
        ptrs = tuple(c.data_ptr() for c in cts)
        if req is not None:
            req._ct = (cts, ptrs)
        return ptrs

    # ------------------------------------------------------------ loop
    def submit(self, req: Request):
        """Queue a request.  Everything that can be wrong with it is checked here,
        before it is queued, so a request is never dropped after admission."""
        if self.footprint(req) > self.hbm_budget:
            raise _native.SpeedyError(f"request {req.rid} needs more HBM than the whole budget")
        kx, ix, il = req.geometry.nodal_shape
        shape = tuple(req.state.temperature.shape)
        if shape[-3:] != (kx, ix, il):
            raise ValueError(f"request {req.rid}: state shape {shape} does not match geometry {(kx, ix, il)}")
        if req.cotangent is not None:
            self._cotangent_ptrs(req.cotangent, shape, req)
        self.queue.append(req)

    def _free_slot(self) -> Optional[_Slot]:
        for s in self.slots:
            if not s.busy:
                return s
        return None

    def _admit(self) -> list:
        batch, cols = [], 0
        limit = self.max_segments if self.batching else 1
        while self.queue and len(batch) < limit:
            req = self.queue[0]
            fp = self.footprint(req)
            if self.in_flight_bytes + fp > self.hbm_budget:
                break
            if batch and cols + req.columns > self.max_batch_columns:
                break
            if batch and req.geometry.nodal_shape[0] != batch[0].geometry.nodal_shape[0]:
                break   # a batch shares one vertical grid; this request opens the next batch
            self.queue.popleft()
            self.in_flight_bytes += fp
            self.peak_bytes = max(self.peak_bytes, self.in_flight_bytes)
            req.t_admit = self.clock()
            batch.append(req)
            cols += req.columns
        return batch

    def _launch(self, slot: _Slot, batch: list):
        phys, lib = self.phys, self.lib
        keep: list = []
        segs = (_native.SpeedySegment * len(batch))()
        s_ptr = ctypes.c_void_p(slot.stream.cuda_stream)
        # the states were written on the client's stream
        self._wait_client(slot)
        # one output block for the whole batch, allocated on the caller's stream (where
        # the results are read); the batch stream's use is recorded so the block is not
        # reused before the batch is done
        offs, total = [], 0
        for req in batch:
            offs.append(total)
            total += _output_floats(tuple(req.state.temperature.shape), req.cotangent is not None)
        buf = torch.empty(total, dtype=torch.float32, device=self.dev)
        buf.record_stream(slot.stream)
        keep.append(buf)
        for i, req in enumerate(batch):
            seg = segs[i]
            (ix, il, nm), flags, member_keep = self._prepare(req, seg, buf, offs[i])
            keep.append(member_keep)   # device copies of a NumPy state live until the batch completes
            ws = self._workspace((ix, il, nm), slot.stream)
            req._ws = ((ix, il, nm), ws)
            seg.date = req.date.to_abi()
            seg.nlon, seg.nlat, seg.nmember = ix, il, nm
            seg.flags, seg.level_set = flags, -1
            seg.workspace, seg.workspace_bytes = ws.data_ptr(), ws.numel()
        gkey = id(batch[0].geometry)
        hit = self._verts.get(gkey)
        if hit is None or hit[0] is not batch[0].geometry:
            hit = self._verts[gkey] = (batch[0].geometry, batch[0].geometry.to_abi())
        vert = hit[1]
        nbytes = lib.speedy_batch_desc_bytes(segs, len(batch))
        if nbytes == 0:
            raise _native.SpeedyError("speedy_batch_desc_bytes failed: " + lib.speedy_last_error().decode())
        if slot.desc is None or slot.desc.numel() < nbytes:
            if slot.event is not None:
                slot.event.synchronize()
            cap = max(nbytes, 1 << 16)
            slot.desc = torch.empty(cap, dtype=torch.uint8, device=self.dev)
            slot.desc.record_stream(slot.stream)
            slot.staging = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
        rc = lib.speedy_fwd_batch(segs, len(batch), ctypes.byref(vert), ctypes.byref(phys._params()),
                                  slot.desc.data_ptr(), slot.desc.numel(), slot.staging.data_ptr(), s_ptr)
        _native.check(rc, "speedy_fwd_batch")
        if slot.event is None:
            slot.event = torch.cuda.Event()
        slot.event.record(slot.stream)
        slot.batch, slot.keep, slot.busy = batch, keep, True
        for req in batch:
            req.batch = self.n_batches
        self.n_batches += 1
        self.batch_sizes.append(len(batch))

    def _launch_single(self, slot: _Slot, req: Request):
        """The per-call path: one speedy_fwd / speedy_fwd_vjp call for one request."""
        self._wait_client(slot)
        with torch.cuda.stream(slot.stream):
            if req.kind == KIND_VJP:
                req.tendency, req.state_grads, req.param_grads = self.phys.compute_tendencies_vjp(
                    req.state, req.forcing, req.geometry, req.date, req.cotangent)
            else:
                req.tendency, _ = self.phys.compute_tendencies(req.state, req.forcing, req.geometry, req.date)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        req._ws = None
        slot.event, slot.batch, slot.keep, slot.busy = ev, [req], [], True
        req.batch = self.n_batches
        self.n_batches += 1
        self.batch_sizes.append(1)

    def poll(self) -> int:
        """Complete every batch whose event has fired; returns how many requests completed."""
        n = 0
        for slot in self.slots:
            if slot.busy and slot.event.query():
                now = self.clock()
                for req in slot.batch:
                    req.t_done = now
                    self.in_flight_bytes -= self.footprint(req)
                    if req._ws is not None:
                        key, ws = req._ws
                        if self.record_level_sets:
                            req.level_set = self._read_level_sets(ws, key[2], slot.stream)
                        self._ws_free.setdefault(key, []).append(ws)
                    req._ws = None
                    self.done.append(req)
                    n += 1
                slot.batch, slot.keep, slot.busy = [], [], False
        return n

    def _read_level_sets(self, ws: torch.Tensor, nmember: int, stream) -> list:
        buf = (ctypes.c_uint32 * nmember)()
        _native.check(self.lib.speedy_read_level_sets(ws.data_ptr(), ws.numel(), nmember, buf,
                                                      ctypes.c_void_p(stream.cuda_stream)),
                      "speedy_read_level_sets")
        return list(buf)

    def step(self) -> bool:
        """Launch one batch if a stream is free and a request can be admitted."""
        slot = self._free_slot()
        if slot is None or not self.queue:
            return False
        batch = self._admit()
        if not batch:
            return False
        if self.batching:
            self._launch(slot, batch)
        else:
            self._launch_single(slot, batch[0])
        return True

    def run(self, requests: list, realtime: bool = True, timeout: float = 120.0) -> list:
        """Serve `requests` (sorted by arrival, seconds relative to the start).  With
        realtime, a request enters the queue only once the clock reaches its arrival
        time; otherwise all are queued at once.  Returns the completed requests."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = self.clock()
        for r in pending:
            r.arrival = r.arrival + t0 if realtime else t0
        while pending or self.queue or any(s.busy for s in self.slots):
            now = self.clock()
            while pending and (not realtime or pending[0].arrival <= now):
                self.submit(pending.popleft())
            self.poll()
            while self.step():
                pass
            if now - t0 > timeout:
                raise _native.SpeedyError("serving loop timed out")
        return self.done

    def reset_stats(self):
        """Forget completed requests and batch counts (after a warm-up run); the
        workspaces, descriptors and member structs built so far are kept."""
        self.done = []
        self.n_batches = 0
        self.batch_sizes = []
        self.peak_bytes = self.in_flight_bytes

    # ------------------------------------------------------------ report
    def stats(self, requests: list | None = None) -> dict:
        reqs = requests if requests is not None else self.done
        lat = np.array([r.latency for r in reqs if r.latency is not None])
        if len(lat) == 0:
            return {}
        t_first = min(r.arrival for r in reqs)
        t_last = max(r.t_done for r in reqs)
        cols = sum(r.columns for r in reqs)
        return {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "columns_per_s": cols / max(t_last - t_first, 1e-9), "batches": self.n_batches,
                "mean_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
                "peak_hbm_bytes": int(self.peak_bytes), "hbm_budget_bytes": self.hbm_budget}
