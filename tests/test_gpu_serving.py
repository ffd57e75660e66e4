"""The continuous-batching server (jcm_amd.serving) on the GPU.

* A mix of requests (T21, T31, T42, T85; forward and forward + VJP; one and
  several members) served together gives, for every request, the result of the
  same request called alone: tendencies, state gradients and parameter
  gradients equal within 1e-6 relative (measured: identical bits), and the same
  level set.
* A state written on the client's stream just before the request, while that
  stream is still busy, is the state the batch reads.
* Admission respects the HBM budget: with a budget that holds only a few
  requests, the loop runs several batches, never holds more than the budget,
  and admits requests that arrive while earlier batches are on the GPU.
"""
import numpy as np
import pytest

import helpers as H
from jcm_amd import Parameters, synthetic
from jcm_amd.physics_interface import PhysicsTendency

pytestmark = pytest.mark.gpu
RES = {"T21": (64, 32), "T31": (96, 48), "T42": (128, 64), "T85": (256, 128)}


def _device_case(nx, ny, seed, nmember=1):
    import torch
    st, fo, ge, da = synthetic.make_case(nx, ny, 8, seed=seed, land=True, nmember=nmember, member_t_sigma=0.5)
    dev = torch.device("cuda", 0)
    st = st.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), device=dev)
                    for k in ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential",
                              "normalized_surface_pressure")})
    return st, fo, ge, da


def _mix(n, seed=0, vjp_every=2):
    import torch
    from jcm_amd.serving import Request
    rng = np.random.default_rng(seed)
    reqs = []
    names = list(RES)
    for i in range(n):
        nx, ny = RES[names[i % len(names)]]
        nm = 2 if i % 5 == 3 else 1
        st, fo, ge, da = _device_case(nx, ny, seed + i, nm)
        da = da.copy(tyear=float(rng.uniform(0, 1)), model_step=int(i % 3))
        cot = None
        if i % vjp_every == 1:
            shp = st.temperature.shape
            cot = PhysicsTendency(*[torch.randn(shp, device=st.temperature.device) for _ in range(4)])
        reqs.append(Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot, rid=i))
    return reqs


def _same(a, b, what):
    import torch
    a, b = a.float(), b.float()
    den = float(b.abs().max()) or 1.0
    err = float((a - b).abs().max()) / den
    assert err <= 1e-6, f"{what}: {err:.2e}"
    return err == 0.0


def test_batched_results_equal_alone():
    import torch
    from jcm_amd.serving import PhysicsServer
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    reqs = _mix(10)
    server = PhysicsServer(phys, n_streams=2)
    done = server.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    assert server.n_batches < len(reqs)        # requests were actually batched
    identical = True
    alone = SpeedyPhysics(parameters=phys.parameters)
    for r in reqs:
        if r.cotangent is None:
            t, _ = alone.compute_tendencies(r.state, r.forcing, r.geometry, r.date)
        else:
            t, sg, pg = alone.compute_tendencies_vjp(r.state, r.forcing, r.geometry, r.date, r.cotangent)
        torch.cuda.synchronize()
        for f in ("u_wind", "v_wind", "temperature", "specific_humidity"):
            identical &= _same(getattr(r.tendency, f), getattr(t, f), f"req {r.rid} {f}")
        if r.cotangent is not None:
            for f in ("u_wind", "v_wind", "temperature", "specific_humidity", "normalized_surface_pressure"):
                identical &= _same(getattr(r.state_grads, f), getattr(sg, f), f"req {r.rid} grad {f}")
            for k in pg:
                identical &= _same(r.param_grads[k], pg[k], f"req {r.rid} {k}")
    print(f"batched == alone bit for bit: {identical}; batches {server.n_batches}", flush=True)


def test_batched_ragged_members_equal_alone():
    """Segments whose members do not fill whole workgroups (100 x 50 and 90 x 45
    columns, 1-3 members, forward and forward + VJP) in one batch: each result equals
    the same request called alone."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    specs = [((100, 50), 3, False), ((100, 50), 3, True), ((90, 45), 2, False), ((90, 45), 2, True),
             ((64, 32), 1, True), ((100, 50), 1, False)]
    reqs = []
    for i, ((nx, ny), nm, vjp) in enumerate(specs):
        st, fo, ge, da = _device_case(nx, ny, 40 + i, nm)
        cot = None
        if vjp:
            cot = PhysicsTendency(*[torch.randn(st.temperature.shape, device=st.temperature.device) for _ in range(4)])
        reqs.append(Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot, rid=i))
    server = PhysicsServer(phys, n_streams=1)
    done = server.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert len(done) == len(reqs) and server.n_batches == 1
    alone = SpeedyPhysics(parameters=phys.parameters)
    for r in reqs:
        if r.cotangent is None:
            t, _ = alone.compute_tendencies(r.state, r.forcing, r.geometry, r.date)
        else:
            t, sg, pg = alone.compute_tendencies_vjp(r.state, r.forcing, r.geometry, r.date, r.cotangent)
        torch.cuda.synchronize()
        for f in ("u_wind", "v_wind", "temperature", "specific_humidity"):
            _same(getattr(r.tendency, f), getattr(t, f), f"req {r.rid} {f}")
        if r.cotangent is not None:
            for f in ("u_wind", "v_wind", "temperature", "specific_humidity", "normalized_surface_pressure"):
                _same(getattr(r.state_grads, f), getattr(sg, f), f"req {r.rid} grad {f}")
            for k in pg:
                _same(r.param_grads[k], pg[k], f"req {r.rid} {k}")


@pytest.mark.parametrize("batching", [True, False])
def test_batch_waits_for_client_stream_writes(batching):
    """A state written on the client's stream just before the request (behind
    about ten milliseconds of other work on that stream) is what the server reads:
    the batch is ordered after the client's stream when that stream is busy."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    st, fo, ge, da = _device_case(256, 128, 77)
    shp = st.temperature.shape
    cot = PhysicsTendency(*[torch.randn(shp, device=st.temperature.device) for _ in range(4)])
    new_t = st.temperature + 3.0 * torch.sin(torch.arange(st.temperature.numel(), device=st.temperature.device)
                                              .view(shp).float() * 0.37)
    old_t = st.temperature.clone()
    st.temperature.copy_(new_t)
    want_f, _ = phys.compute_tendencies(st, fo, ge, da)
    want_v, want_g, _ = phys.compute_tendencies_vjp(st, fo, ge, da, cot)
    want_f = want_f.temperature.clone()
    want_v, want_g = want_v.temperature.clone(), want_g.temperature.clone()
    st.temperature.copy_(old_t)
    a = torch.randn(4096, 4096, device=st.temperature.device)
    torch.cuda.synchronize()
    for _ in range(8):
        a = torch.mm(a, a) * 1e-3
    st.temperature.copy_(new_t)
    busy = not torch.cuda.current_stream().query()
    reqs = [Request(state=st, forcing=fo, geometry=ge, date=da, rid=0),
            Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot, rid=1)]
    server = PhysicsServer(phys, n_streams=2, batching=batching)
    server.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert busy, "the client stream finished before the request was served; the wait was not exercised"
    _same(reqs[0].tendency.temperature, want_f, "forward dT")
    _same(reqs[1].tendency.temperature, want_v, "forward + VJP dT")
    _same(reqs[1].state_grads.temperature, want_g, "dL/dT")


def test_member_cache_releases_dropped_states():
    """States that are new with every request (a model stepping forward) are not
    held by the server once the client drops them: the member cache keeps weak
    references to the state, forcing and geometry."""
    import gc
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    server = PhysicsServer(phys, n_streams=1)
    st, fo, ge, da = _device_case(64, 32, 5)
    fields = ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential", "normalized_surface_pressure")
    for i in range(3):
        s_i = st.copy(**{k: getattr(st, k).clone() for k in fields})
        reqs = [Request(state=s_i, forcing=fo, geometry=ge, date=da, rid=i)]
        server.run(reqs, realtime=False)
        torch.cuda.synchronize()
        assert len(server._members) == 1
        del reqs, s_i
        server.reset_stats()
        gc.collect()
        assert len(server._members) == 0, "a dropped state is still held by the server"


def test_admission_respects_hbm_budget_and_overlaps():
    import torch
    from jcm_amd.serving import PhysicsServer
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    reqs = _mix(16, seed=5)
    server = PhysicsServer(phys, n_streams=2)
    budget = 3 * max(server.footprint(r) for r in reqs)
    server = PhysicsServer(phys, hbm_budget=budget, n_streams=2)
    # arrivals spread over ~20 ms so that requests arrive while batches run
    for i, r in enumerate(reqs):
        r.arrival = 0.00125 * i
    done = server.run(reqs, realtime=True)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    assert server.peak_bytes <= budget
    assert server.n_batches >= 3
    st = server.stats()
    print(st, flush=True)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    late = [r for r in reqs if r.batch > 0 and r.arrival > min(x.t_admit for x in reqs)]
    assert late, "no request was admitted after the first batch had started"


def test_per_request_two_states_same_grid_two_streams():
    """batching=False with two slot streams: two requests on the same grid with
    different states are in flight together; each equals the same request alone,
    bit for bit (each stream has its own workspace: level-set words, generation,
    parameter partials)."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2))
    reqs = []
    for i in range(6):
        st, fo, ge, da = _device_case(256, 128, 90 + i)
        if i % 2:   # wetter states: other level sets than the dry ones
            st = st.copy(specific_humidity=st.specific_humidity * 1.12)
        cot = PhysicsTendency(*[torch.randn(st.temperature.shape, device=st.temperature.device) for _ in range(4)]) \
            if i % 3 == 2 else None
        reqs.append(Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot, rid=i))
    server = PhysicsServer(phys, n_streams=2, batching=False)
    server.run(reqs, realtime=False)
    torch.cuda.synchronize()
    alone = SpeedyPhysics(parameters=phys.parameters)
    for r in reqs:
        if r.cotangent is None:
            t, _ = alone.compute_tendencies(r.state, r.forcing, r.geometry, r.date)
        else:
            t, sg, pg = alone.compute_tendencies_vjp(r.state, r.forcing, r.geometry, r.date, r.cotangent)
        torch.cuda.synchronize()
        for f in ("u_wind", "v_wind", "temperature", "specific_humidity"):
            assert torch.equal(getattr(r.tendency, f), getattr(t, f)), f"req {r.rid} {f}"
        if r.cotangent is not None:
            for f in ("temperature", "specific_humidity", "normalized_surface_pressure"):
                assert torch.equal(getattr(r.state_grads, f), getattr(sg, f)), f"req {r.rid} grad {f}"
            for k in pg:
                assert torch.equal(r.param_grads[k], pg[k]), f"req {r.rid} {k}"


@pytest.mark.parametrize("batching", [True, False])
def test_numpy_state_edited_in_place_between_requests(batching):
    """A NumPy state is copied to the device for each request (it is not cached),
    so an in-place edit between two requests is seen by the second."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    st, fo, ge, da = synthetic.make_case(64, 32, 8, seed=12, land=True)
    server = PhysicsServer(phys, n_streams=1, batching=batching)
    r0 = Request(state=st, forcing=fo, geometry=ge, date=da, rid=0)
    server.run([r0], realtime=False)
    torch.cuda.synchronize()
    st.temperature += np.float32(2.0)                    # in place, same object
    r1 = Request(state=st, forcing=fo, geometry=ge, date=da, rid=1)
    server.run([r1], realtime=False)
    torch.cuda.synchronize()
    want, _ = phys.compute_tendencies(st, fo, ge, da)
    torch.cuda.synchronize()
    assert torch.equal(r1.tendency.temperature, want.temperature)
    assert not torch.equal(r0.tendency.temperature, r1.tendency.temperature)


def test_mixed_level_counts_in_one_stream():
    """Requests with 8 and 16 levels interleaved: each batch holds one vertical
    grid, every request is served, and each result equals the request alone."""
    import torch
    from jcm_amd.serving import PhysicsServer, Request
    from jcm_amd.speedy_physics import SpeedyPhysics
    phys = SpeedyPhysics()
    reqs = []
    for i in range(8):
        kx = 8 if i % 2 == 0 else 16
        st, fo, ge, da = synthetic.make_case(64, 32, kx, seed=60 + i, land=True)
        dev = torch.device("cuda", 0)
        st = st.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(st, k)), device=dev)
                        for k in H.STATE_FIELDS})
        reqs.append(Request(state=st, forcing=fo, geometry=ge, date=da, rid=i))
    server = PhysicsServer(phys, n_streams=2)
    done = server.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    assert server.in_flight_bytes == 0
    for r in reqs:
        t, _ = phys.compute_tendencies(r.state, r.forcing, r.geometry, r.date)
        torch.cuda.synchronize()
        assert torch.equal(r.tendency.temperature, t.temperature), r.rid
