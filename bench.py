"""Benchmark: SPEEDY physics columns/s on one MI355X at T85 x 8 levels (FP32).

Workload: T85 Gaussian grid, 256 x 128 = 32768 columns, 8 sigma levels, the
land recipe of jcm_amd.synthetic (seed 1), inputs resident in HBM.

  value      forward: one step is one SpeedyPhysics.compute_tendencies call (the
             eleven SPEEDY terms, one fused kernel launch).  Steps are replayed
             from HIP graphs of --graph-calls calls so that host launch overhead
             does not hide device time; K steps are bracketed by
             torch.cuda.synchronize() on both sides, and HIP events on the
             replay stream time them (untimed replays queued just before keep
             the GPU's clock ramp after setup out of short measurements).
  fwd_vjp    the same for compute_tendencies_vjp (fused forward + VJP w.r.t. T, q,
             u, v, ps and five parameters, cotangents N(0,1); two launches: the
             column kernel and the parameter-gradient sums).
  per_call   both calls made from Python one by one without graphs (the plain
             per-call path: host overhead included), with the default
             cooperative launch and with plain_launch=True.
  serving    a Poisson stream of requests of mixed resolution (T21, T31, T42,
             T85) and kind (forward, forward + VJP) through the continuous-
             batching loop (jcm_amd.serving) and through one call per request:
             p50/p99 latency and columns/s.

Prints one JSON line.  `roofline` is for the dominant kernel (the fused forward
kernel, one launch per step): its average duration is the graph-replayed step
time measured with HIP events on the replay stream; `achieved` counts the path's
algorithmic bytes (328 B per column at 8 levels: u, v, T, q, phi at every
level), `used_bytes` what the kernel needs (272 B: u and v at the lowest level
only).  `cpu_baseline` times the kernel's own column routine built for the host
with OpenMP on the host's cores (forward and forward + VJP, T85 and T31), and
the float32 NumPy oracle on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "jax-gcm_amd")]

METRIC = "SPEEDY physics columns/s fwd & fwd+VJP at T85 vs plain; p50/p99 req latency"
HBM_PEAK_GBS = 8000.0


def FWD_BYTES_PER_COL(kx):
    """Algorithmic bytes of the forward per column, as the path's contract counts
    them: reads u, v, T, q, phi at all levels and ps (5*kx + 1 floats), 6 forcing
    and 3 geometry fields; writes the 4 tendency fields (4*kx): (9*kx + 10) * 4,
    328 B at 8 levels."""
    return (9 * kx + 10) * 4


def VJP_BYTES_PER_COL(kx):
    """Algorithmic bytes of the forward + VJP per column: the forward's reads plus
    the 4 cotangent fields (4*kx); writes the 4 tendency fields, the 4 state
    gradients (8*kx) and dL/dps (1): (17*kx + 11) * 4, 588 B at 8 levels."""
    return (17 * kx + 11) * 4


def FWD_USED_BYTES_PER_COL(kx):
    """Bytes the forward kernel actually needs: u and v are read at the lowest
    level only (only the surface stress uses them): (3*kx + 12 + 4*kx) * 4,
    272 B at 8 levels."""
    return (3 * kx + 12 + 4 * kx) * 4


def VJP_USED_BYTES_PER_COL(kx):
    """As FWD_USED_BYTES_PER_COL for the forward + VJP (cotangents du, dv at the
    lowest level only): ((3*kx + 12) + (2*kx + 2) + (8*kx + 1)) * 4, 476 B."""
    return ((3 * kx + 12) + (2 * kx + 2) + (8 * kx + 1)) * 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--graph-calls", type=int, default=50)
    ap.add_argument("--per-call-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rates", type=str, default="2000,20000",
                    help="Poisson arrival rates (requests/s) of the serving runs")
    ap.add_argument("--profile-only", action="store_true",
                    help="replay --steps forward and fwd+VJP calls, run a copy of --copy-mib MiB, and exit")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--transforms-only", action="store_true",
                    help="time only the T85 dynamics <-> physics transform step (--steps steps) and print its JSON")
    return ap.parse_args()


def to_device(case, dev):
    import numpy as np
    import torch
    state, forcing, geometry, date = case
    st = state.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(state, k)), device=dev)
                       for k in ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential",
                                 "normalized_surface_pressure")})
    fo = forcing.copy(**{k: torch.as_tensor(np.ascontiguousarray(getattr(forcing, k)), device=dev)
                         for k in ("alb0", "sice_am", "snowc_am", "soilw_am", "stl_am", "sea_surface_temperature")})
    return st, fo, geometry, date


def _timed(fn, seconds, max_n=100000):
    t0 = time.perf_counter()
    n = 0
    while True:
        fn()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= max_n:
            return n, el


def _host_cpus():
    """(threads the OpenMP leg uses, CPUs the host shows, CPU model)."""
    ncpu = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", ncpu))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, ncpu, model


def cpu_baseline(case_t85, case_t31, seconds):
    """CPU restatements of the path, timed on this host: the kernel's own column
    routine compiled for the host with OpenMP over the columns (all threads the
    host allows, OMP_NUM_THREADS) -- forward and forward + VJP at T85 and T31 --
    and the float32 NumPy oracle on one core."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers
    from jcm_amd import synthetic
    threads, ncpu, model = _host_cpus()
    omp = ctypes.CDLL(os.path.join(ROOT, "tests", "host_column", "libcolumn_host_omp.so"))
    helpers.build_host_column = lambda: omp      # the OpenMP build of the same harness
    legs = {}
    params = synthetic.draw_vjp_parameters(2)
    for name, case in (("T85", case_t85), ("T31", case_t31)):
        kx, ix, il = case[2].nodal_shape
        ncol = ix * il
        rng = np.random.default_rng(2)
        cot = [rng.standard_normal((kx, ix, il)).astype(np.float32) for _ in range(4)]
        n, el = _timed(lambda: helpers.run_host_column(case, params), seconds / 4)
        legs[f"fwd_{name}"] = ncol * n / el
        n, el = _timed(lambda: helpers.run_host_vjp(case, params, cot, dtype=np.float32), seconds / 4)
        legs[f"fwd_vjp_{name}"] = ncol * n / el
    n, el = _timed(lambda: helpers.run_oracle(*case_t85), seconds / 2, max_n=1000)
    numpy = case_t85[2].ncol * n / el
    return {"value": legs["fwd_T85"], "unit": "columns/s", "cores": threads, "kind": "port",
            "sample": f"the kernel's column routine built for the host (g++ -O2 -fopenmp, float32), T85x8 forward "
                      f"calls for {seconds / 4:.1f} s on {threads} OpenMP threads of a {ncpu}-CPU host ({model})",
            "legs_columns_per_s": legs, "host_cpus": ncpu, "cpu_model": model,
            "numpy_oracle": {"value": numpy, "unit": "columns/s", "cores": 1,
                             "sample": f"{n} T85x8 forward calls of the float32 NumPy oracle, {el:.1f} s"}}


class GraphStep:
    """`calls` invocations of fn captured in one HIP graph on a side stream."""

    def __init__(self, fn, calls, dev):
        import torch
        self.calls = calls
        self.stream = stream = torch.cuda.Stream(device=dev)
        stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(stream):
            fn()                                  # warm the side stream
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            for _ in range(calls):
                fn()
        torch.cuda.synchronize()
        self.graph.replay()
        torch.cuda.synchronize()

    def queue(self, warm_ms):
        """Queue untimed replays worth about warm_ms of device work and return at once:
        the GPU's clocks are up and the stream is busy when the next work starts."""
        import torch
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        self.graph.replay()
        torch.cuda.synchronize()
        one = max(time.perf_counter() - w0, 1e-6)
        for _ in range(max(1, int(warm_ms * 1e-3 / one))):
            self.graph.replay()

    def time(self, reps, warm_ms=20.0):
        """ms per call over reps replays, HIP events on the replaying stream.  Untimed
        replays (about warm_ms of device work) are queued right ahead of the first
        event, so a short measurement (one replay of a 20-call graph) sees the same
        per-call time as a long one instead of the clock ramp after setup."""
        import torch
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        self.queue(warm_ms)
        t0.record()
        for _ in range(reps):
            self.graph.replay()
        t1.record()
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / (reps * self.calls)


def per_call_ms(fn, steps):
    import torch
    for _ in range(10):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def serving_runs(phys, dev, n_requests, rates, seed=7):
    """Poisson streams of mixed requests through the batching loop and through one
    call per request."""
    import numpy as np
    import torch
    from jcm_amd import synthetic
    from jcm_amd.physics_interface import PhysicsTendency
    from jcm_amd.serving import PhysicsServer, Request
    res = [(64, 32), (96, 48), (128, 64), (256, 128)]
    # device-resident member states, one per resolution, reused by the requests
    pool = []
    for i, (nx, ny) in enumerate(res):
        st, fo, ge, da = to_device(synthetic.make_case(nx, ny, 8, seed=100 + i, land=True), dev)
        cot = PhysicsTendency(*[torch.randn(st.temperature.shape, device=dev) for _ in range(4)])
        pool.append((st, fo, ge, da, cot))
    out = {"workload": f"{n_requests} requests, resolution uniform over T21/T31/T42/T85 x 8 levels, "
                       "50/50 forward and forward+VJP, Poisson arrivals (and all at once: saturation), 2 streams, "
                       "4 GiB HBM budget; measured after one untimed pass of the same stream",
           "runs": []}
    # the Poisson streams, then saturation (rate None: every request queued at once,
    # realtime=False), where throughput is no longer capped by the offered load
    for rate in list(rates) + [None]:
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / (rate or 1e9), n_requests))
        picks = rng.integers(0, len(res), n_requests)
        kinds = rng.integers(0, 2, n_requests)
        realtime = rate is not None
        run = {"rate_per_s": rate if realtime else "saturation"}
        for mode, batching in (("batched", True), ("per_request", False)):
            def stream():
                reqs = []
                for i in range(n_requests):
                    st, fo, ge, da, cot = pool[picks[i]]
                    reqs.append(Request(state=st, forcing=fo, geometry=ge, date=da,
                                        cotangent=cot if kinds[i] else None, arrival=float(arrivals[i]), rid=i))
                return reqs
            server = PhysicsServer(phys, hbm_budget=4 << 30, n_streams=2, batching=batching)
            # warm-up: the same stream once, untimed, so that workspaces, batch
            # descriptors and member structs exist before the measured run
            server.run(stream(), realtime=realtime, timeout=60.0)
            torch.cuda.synchronize()
            server.reset_stats()
            reqs = stream()
            server.run(reqs, realtime=realtime, timeout=60.0)
            torch.cuda.synchronize()
            run[mode] = server.stats(reqs)
        b, p = run["batched"].get("columns_per_s"), run["per_request"].get("columns_per_s")
        run["throughput_batched_over_per_request"] = b / p if b and p else None
        out["runs"].append(run)
    return out


def profile_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/<name>_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", name + "_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def profile_valu(kernel):
    """rocprofv3 VALUBusy / VALUUtilization (percent) of a kernel from the committed
    summary profiles/speedy_kernels_valu.json, or None."""
    path = os.path.join(ROOT, "profiles", "speedy_kernels_valu.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        k = json.load(fh)["kernels"].get(kernel)
    return None if k is None else {"valu_busy_pct": k["VALUBusy"], "valu_utilization_pct": k["VALUUtilization"]}


def large_leg(phys, dev, reps=5):
    """T425 (1280 x 640 = 819200 columns, 8 levels) forward + VJP, and the forward
    alone, through the batched entry: kernel time from the library's HIP-event timers
    (level-set pass, forward (+ VJP), parameter reduction), after one untimed call each."""
    import torch
    from jcm_amd import _native, synthetic
    from jcm_amd.physics_interface import PhysicsTendency
    from jcm_amd.serving import PhysicsServer, Request
    st, fo, ge, da = to_device(synthetic.make_case(1280, 640, 8, seed=1, land=True), dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    cot = PhysicsTendency(*[torch.randn(st.temperature.shape, device=dev, generator=gen) for _ in range(4)])
    server = PhysicsServer(phys, n_streams=1)
    lib = _native.load()
    server.run([Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot)], realtime=False)
    server.run([Request(state=st, forcing=fo, geometry=ge, date=da, rid=-1)], realtime=False)
    torch.cuda.synchronize()
    lib.speedy_timer_enable(16 * reps)
    for i in range(reps):
        server.run([Request(state=st, forcing=fo, geometry=ge, date=da, cotangent=cot, rid=i + 1)], realtime=False)
    torch.cuda.synchronize()
    tm = _native.timer_read()
    lib.speedy_timer_disable()
    ncol = 1280 * 640
    per_kind = {k: v[0] / reps for k, v in tm.items() if v[1]}
    ms = sum(per_kind.values())
    kx = 8
    achieved = VJP_BYTES_PER_COL(kx) * ncol / (per_kind["fwd_vjp"] * 1e-3) / 1e9
    used = VJP_USED_BYTES_PER_COL(kx) * ncol / (per_kind["fwd_vjp"] * 1e-3) / 1e9
    out = {"workload": "T425 1280x640 = 819200 columns x 8 levels, forward + VJP, batched entry (one request)",
           "columns_per_s": ncol / (ms * 1e-3), "ms_per_call": ms, "kernel_ms": per_kind,
           "roofline": {"bound": "hbm", "kernel": "speedy::fwd_batch_kernel<8>", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "used_bytes_frac": used / HBM_PEAK_GBS,
                        "traffic": profile_traffic("fwd_batch_kernel_T425"),
                        "valu": profile_valu("speedy::fwd_batch_kernel<8>")}}
    # the forward alone (forward-only batch kernel)
    lib.speedy_timer_enable(16 * reps)
    for i in range(reps):
        server.run([Request(state=st, forcing=fo, geometry=ge, date=da, rid=100 + i)], realtime=False)
    torch.cuda.synchronize()
    tm = _native.timer_read()
    lib.speedy_timer_disable()
    per_kind = {k: v[0] / reps for k, v in tm.items() if v[1]}
    ms = sum(per_kind.values())
    achieved = FWD_BYTES_PER_COL(kx) * ncol / (per_kind["fwd"] * 1e-3) / 1e9
    used = FWD_USED_BYTES_PER_COL(kx) * ncol / (per_kind["fwd"] * 1e-3) / 1e9
    out["fwd"] = {"columns_per_s": ncol / (ms * 1e-3), "ms_per_call": ms, "kernel_ms": per_kind,
                  "roofline": {"bound": "hbm", "kernel": "speedy::fwd_plain_batch_kernel<8>", "achieved": achieved,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                               "used_bytes_frac": used / HBM_PEAK_GBS,
                               "traffic": profile_traffic("fwd_plain_batch_kernel_T425"),
                               "valu": profile_valu("speedy::fwd_plain_batch_kernel<8>")}}
    return out


MFMA_FP32_PEAK_TFS = 157.3


def _fourier_factors(nlon):
    """nlon = n1 * n2 with n1 the largest divisor <= sqrt(nlon), as the library splits it."""
    n1 = max(d for d in range(1, int(nlon ** 0.5) + 1) if nlon % d == 0)
    return n1, nlon // n1


def transform_work(g, kx):
    """(flops, bytes) of one transform step: dynamics -> physics state plus physics
    tendency -> dynamics at the grid g (5*kx+1 fields synthesised: winds,
    temperature, humidity, log surface pressure, geopotential; 4*kx analysed).
    Flops, 2 per multiply-add, as the kernels compute them: the Legendre
    contraction over half the latitudes (equatorial symmetry), and the Fourier
    stages' two passes per latitude line (nlon = n1 * n2; analysis: n2-point real
    DFTs, twiddles, n1-point DFTs for the n_m wavenumbers kept; synthesis the
    reverse).  Bytes: modal inputs and outputs, nodal inputs and outputs, and the
    Legendre basis and the roots of unity read once."""
    mm, nl, nlon, nlat, n_m = 2 * g.n_m - 1, g.n_l, g.nlon, g.nlat, g.n_m
    n1, n2 = _fourier_factors(nlon)
    nh, k1 = n2 // 2 + 1, -(-n_m // n2)
    legendre = lambda nf: 2.0 * nf * mm * nlat * (nl / 2.0)
    analysis_line = 4.0 * n1 * nh * n2 + 6.0 * n1 * n2 + 8.0 * n_m * n1
    synthesis_line = n2 * n1 * (8.0 * k1 + 6.0) + 4.0 * n1 * n2 * n2
    flops = (legendre(5 * kx + 1) + (5 * kx + 1) * nlat * synthesis_line
             + legendre(4 * kx) + 4 * kx * nlat * analysis_line)
    modal, nodal = mm * nl * 4, nlon * nlat * 4
    bases = g.n_m * nlat * nl * 4 + nlon * 8
    nbytes = ((4 * kx + 2) * modal + (5 * kx + 1) * nodal      # dynamics -> physics
              + 4 * kx * nodal + (4 * kx + 1) * modal           # tendency -> dynamics
              + 2 * bases)
    return flops, nbytes


def transforms_leg(dev, trunc=85, kx=8, reps=None, calls=20, profile_reps=0):
    """T85 x 8 levels: dynamics_state_to_physics_state + physics_tendency_to_dynamics_tendency
    per step (the conversions around one physics call), replayed from a HIP graph."""
    import numpy as np
    import torch
    from jcm_amd import transforms as T
    from jcm_amd.geometry import SIGMA_LAYER_BOUNDARIES
    from jcm_amd.physics_interface import PhysicsTendency
    g = T.SpectralGrid.truncation(trunc)
    rng = np.random.default_rng(11)
    mode_m = np.array([0] + [m for mm in range(1, g.n_m) for m in (mm, mm)])
    mask = (mode_m[:, None] <= np.arange(g.n_l)[None, :]).astype(np.float32)
    modal = lambda lead, scale: torch.as_tensor(
        (rng.standard_normal(tuple(lead) + g.modal_shape) * scale * mask).astype(np.float32), device=dev)
    dyn = T.PrimitiveEquations(reference_temperature=torch.linspace(220.0, 288.0, kx, device=dev),
                               orography=modal((), 300.0),
                               coords=T.CoordinateSystem(g, T.SigmaCoordinates(SIGMA_LAYER_BOUNDARIES[kx])))
    st = T.DynamicsState(modal((kx,), 1e-5), modal((kx,), 1e-6), modal((kx,), 5.0), modal((1,), 0.02),
                         {"specific_humidity": modal((kx,), 1e-3)})
    shape = (kx,) + g.nodal_shape
    tend = PhysicsTendency(*(torch.randn(shape, device=dev) * s for s in (1e-4, 1e-4, 1e-5, 1e-6)))
    fn = lambda: (T.dynamics_state_to_physics_state(st, dyn), T.physics_tendency_to_dynamics_tendency(tend, dyn))
    step = GraphStep(fn, calls, dev)
    if profile_reps:
        for _ in range(profile_reps):
            step.graph.replay()
        torch.cuda.synchronize()
        return None
    ms = step.time(reps)
    flops, nbytes = transform_work(g, kx)
    tfs = flops / (ms * 1e-3) / 1e12
    return {"workload": f"T{trunc} {g.nlon}x{g.nlat} x {kx} levels: dynamics_state_to_physics_state + "
                        "physics_tendency_to_dynamics_tendency per step, graph replay",
            "ms_per_step": ms, "steps_per_s": 1e3 / ms,
            "roofline": {"bound": "mfma", "achieved": tfs, "peak": MFMA_FP32_PEAK_TFS, "unit": "TFLOP/s",
                         "frac": tfs / MFMA_FP32_PEAK_TFS, "flops_per_step": flops, "dtype": "f32",
                         "hbm_bytes_per_step": nbytes,
                         "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the SPEEDY path shards by columns with no exchange but a per-member level-set OR)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from jcm_amd import _native, synthetic
    from jcm_amd.physics_interface import PhysicsTendency
    from jcm_amd.speedy_physics import SpeedyPhysics

    dev = torch.device("cuda", 0)
    if args.transforms_only:
        print(json.dumps(transforms_leg(dev, reps=max(1, args.steps // 20))), flush=True)
        return
    case_np = synthetic.make_case(256, 128, 8, seed=1, land=True)
    case = to_device(case_np, dev)
    kx, ix, il = case_np[2].nodal_shape
    ncol = ix * il
    # The profile run launches every call plainly (as graph replays do): under
    # rocprofv3, a process that made cooperative launches ends in the tool's exit
    # handlers with SIGSEGV after the results are written.
    plain = args.profile_only
    phys = SpeedyPhysics(device=dev, plain_launch=plain)
    vphys = SpeedyPhysics(parameters=synthetic.draw_vjp_parameters(2), device=dev, plain_launch=plain)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    cot = PhysicsTendency(*[torch.randn(case[0].temperature.shape, device=dev, generator=gen) for _ in range(4)])
    fwd = lambda: phys.compute_tendencies(*case)
    vjp = lambda: vphys.compute_tendencies_vjp(*case, cot)

    for _ in range(args.warmup):
        fwd()
        vjp()
    torch.cuda.synchronize()
    G = max(1, min(args.graph_calls, args.steps))
    reps = max(1, (args.steps + G - 1) // G)
    steps = reps * G
    g_fwd = GraphStep(fwd, G, dev)
    g_vjp = GraphStep(vjp, G, dev)

    if args.profile_only:
        for _ in range(reps):
            g_fwd.graph.replay()
        for _ in range(reps):
            g_vjp.graph.replay()
        torch.cuda.synchronize()
        large_leg(vphys, dev, reps=3)
        transforms_leg(dev, profile_reps=max(1, steps // 100))
        # A copy of known size (torch.neg: reads and writes copy_mib MiB each) whose
        # FETCH_SIZE / WRITE_SIZE counters calibrate the counter units per byte.
        n = (args.copy_mib << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(f"profile run: {steps} forward and {steps} forward+VJP calls at T85 replayed from graphs of {G}; "
              f"4 batched T425 forward+VJP calls; {max(1, steps // 100) * 20} T85 transform steps; "
              f"3 x neg of {args.copy_mib} MiB", flush=True)
        return

    # ---- timed region (headline): K forward steps replayed from graphs
    ms_per_step = g_fwd.time(reps)
    value = ncol / (ms_per_step * 1e-3)
    ms_vjp = g_vjp.time(reps)

    # ---- launches per call (library timer on plain calls: counts only)
    lib = _native.load()
    lib.speedy_timer_enable(64)
    fwd()
    vjp()
    torch.cuda.synchronize()
    tm = _native.timer_read()
    lib.speedy_timer_disable()
    launches = {k: v[1] for k, v in tm.items() if v[1]}

    # per-kernel device time of the forward + VJP call (HIP events around each of
    # its two kernels): the VJP kernel without the parameter sums.  The calls use
    # plain launches, as the graph replays do (a cooperative launch waits for an
    # idle device), and queue behind ~20 ms of graph replays on the same stream, so
    # each event pair brackets its kernel and not the host's launch gaps.
    p_phys = SpeedyPhysics(device=dev, plain_launch=True)
    p_vphys = SpeedyPhysics(parameters=vphys.parameters, device=dev, plain_launch=True)
    with torch.cuda.stream(g_vjp.stream):
        p_vphys.compute_tendencies_vjp(*case, cot)
        g_vjp.queue(20.0)
        lib.speedy_timer_enable(64)
        for _ in range(20):
            p_vphys.compute_tendencies_vjp(*case, cot)
    torch.cuda.synchronize()
    tm_v = _native.timer_read()
    lib.speedy_timer_disable()
    vjp_kernel_ms = {k: v[0] / v[1] for k, v in tm_v.items() if v[1]}

    bytes_fwd = FWD_BYTES_PER_COL(kx) * ncol
    achieved = bytes_fwd / (ms_per_step * 1e-3) / 1e9
    used_fwd = FWD_USED_BYTES_PER_COL(kx) * ncol / (ms_per_step * 1e-3) / 1e9
    bytes_vjp = VJP_BYTES_PER_COL(kx) * ncol
    achieved_vjp = bytes_vjp / (ms_vjp * 1e-3) / 1e9
    used_vjp = VJP_USED_BYTES_PER_COL(kx) * ncol / (ms_vjp * 1e-3) / 1e9
    traffic = profile_traffic("fwd_kernel_T85")

    out = {
        "metric": METRIC, "value": value, "unit": "columns/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "SPEEDY physics forward, T85 256x128 = 32768 columns x 8 levels, land recipe seed 1",
                   "columns": ncol, "levels": kx, "pass": "fwd", "graph_calls": G},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "speedy::fwd_kernel<8, false, true, false, false>", "kernel_ms": ms_per_step,
                     "valu": profile_valu("speedy::fwd_kernel<8, false, true, false, false>"),
                     "algorithmic_bytes_per_launch": bytes_fwd,
                     "bytes_per_column": FWD_BYTES_PER_COL(kx),
                     "used_bytes": {"bytes_per_column": FWD_USED_BYTES_PER_COL(kx), "achieved": used_fwd,
                                    "frac": used_fwd / HBM_PEAK_GBS}},
        "fwd_vjp": {"value": ncol / (ms_vjp * 1e-3), "unit": "columns/s", "ms_per_step": ms_vjp,
                    "launches_per_step": 2, "kernel_ms": vjp_kernel_ms,
                    "roofline": {"bound": "hbm", "achieved": achieved_vjp, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": achieved_vjp / HBM_PEAK_GBS, "algorithmic_bytes_per_step": bytes_vjp,
                                 "bytes_per_column": VJP_BYTES_PER_COL(kx),
                                 "used_bytes": {"bytes_per_column": VJP_USED_BYTES_PER_COL(kx), "achieved": used_vjp,
                                                "frac": used_vjp / HBM_PEAK_GBS},
                                 "traffic": profile_traffic("fwd_vjp_kernel_T85"),
                                 "valu": profile_valu("speedy::fwd_kernel<8, false, true, true, true>"),
                                 "kernel": "speedy::fwd_kernel<8, false, true, true, true> + param_reduce_kernel"}},
        "launches_per_call": launches,
    }
    # ---- plain per-call path (Python, no graphs)
    pc_f = per_call_ms(fwd, args.per_call_steps)
    pc_v = per_call_ms(vjp, args.per_call_steps)
    # the same calls with plain instead of cooperative launches (plain_launch=True)
    pc_pf = per_call_ms(lambda: p_phys.compute_tendencies(*case), args.per_call_steps)
    pc_pv = per_call_ms(lambda: p_vphys.compute_tendencies_vjp(*case, cot), args.per_call_steps)
    out["per_call"] = {"fwd_columns_per_s": ncol / (pc_f * 1e-3), "fwd_ms": pc_f,
                       "fwd_vjp_columns_per_s": ncol / (pc_v * 1e-3), "fwd_vjp_ms": pc_v,
                       "speedup_graph_over_per_call": {"fwd": pc_f / ms_per_step, "fwd_vjp": pc_v / ms_vjp},
                       "plain_launch": {"fwd_ms": pc_pf, "fwd_vjp_ms": pc_pv,
                                        "fwd_columns_per_s": ncol / (pc_pf * 1e-3),
                                        "fwd_vjp_columns_per_s": ncol / (pc_pv * 1e-3)}}
    out["large"] = large_leg(vphys, dev)
    out["transforms"] = transforms_leg(dev, reps=max(1, steps // 20))
    if not args.no_serving:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["serving"] = serving_runs(vphys, dev, args.serve_requests, rates)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case_np, synthetic.make_case(96, 48, 8, seed=0, land=False),
                                           args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
