"""Benchmark: SPEEDY physics columns/s on one MI355X at T85 x 8 levels (FP32).

Workload: T85 Gaussian grid, 256 x 128 = 32768 columns, 8 sigma levels, the
land recipe of jcm_amd.synthetic (seed 1), inputs resident in HBM.  One step is
one SpeedyPhysics.compute_tendencies call (the eleven SPEEDY terms; two kernel
launches).  Steps are replayed from HIP graphs of --graph-calls calls so that
host launch overhead does not hide the device time; K steps are bracketed by
torch.cuda.synchronize() on both sides.

Prints one JSON line.  `roofline` is for the dominant kernel (the fused
forward kernel), timed live with HIP events the library records around its
launches; `cpu_baseline` times the float32 NumPy oracle (the CPU restatement of
the reference) on the same workload on one host core.
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
# Algorithmic bytes per column of the forward (kx = 8): reads T, q, phi at all
# levels, u and v at the lowest level, ps, 6 forcing and 3 geometry fields
# (3*kx + 12 floats); writes the 4 tendency fields (4*kx floats).
FWD_BYTES_PER_COL = lambda kx: (3 * kx + 12 + 4 * kx) * 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--graph-calls", type=int, default=50)
    ap.add_argument("--timer-steps", type=int, default=500)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="replay --steps calls, run a copy of --copy-mib MiB, and exit (for rocprofv3)")
    ap.add_argument("--copy-mib", type=int, default=256)
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


def cpu_baseline(case, seconds):
    """Float32 NumPy oracle on the same T85 workload, one host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers
    ncol = case[2].ncol
    t0 = time.perf_counter()
    n = 0
    while True:
        helpers.run_oracle(*case)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 1000:
            break
    return {"value": ncol * n / el, "unit": "columns/s", "cores": 1, "kind": "port",
            "sample": f"{n} full T85x8 forward calls ({n * ncol} columns) of the float32 NumPy oracle, {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the SPEEDY path shards by columns with no exchange but a level-set OR; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from jcm_amd import _native, synthetic
    from jcm_amd.speedy_physics import SpeedyPhysics

    dev = torch.device("cuda", 0)
    case_np = synthetic.make_case(256, 128, 8, seed=1, land=True)
    case = to_device(case_np, dev)
    kx, ix, il = case_np[2].nodal_shape
    ncol = ix * il
    phys = SpeedyPhysics(device=dev)
    lib = _native.load()

    def graph_of(calls):
        stream = torch.cuda.Stream(device=dev)
        stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(stream):
            phys.compute_tendencies(*case)     # warm the side stream
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            for _ in range(calls):
                phys.compute_tendencies(*case)
        torch.cuda.synchronize()
        return g

    for _ in range(args.warmup):
        phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    G = max(1, min(args.graph_calls, args.steps))
    reps = max(1, (args.steps + G - 1) // G)
    steps = reps * G
    graph = graph_of(G)
    graph.replay()
    torch.cuda.synchronize()

    if args.profile_only:
        for _ in range(reps):
            graph.replay()
        torch.cuda.synchronize()
        # A copy of known size (torch.neg: reads and writes copy_mib MiB each) whose
        # FETCH_SIZE / WRITE_SIZE counters calibrate the counter units per byte.
        n = (args.copy_mib << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(f"profile run: {steps} forward calls at T85 replayed from graphs of {G}; "
              f"3 x neg of {args.copy_mib} MiB", flush=True)
        return

    # ---- timed region: K steps replayed from graphs, bracketed by synchronize
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record()
    for _ in range(reps):
        graph.replay()
    t_end.record()
    torch.cuda.synchronize()
    ms = t_start.elapsed_time(t_end)
    ms_per_step = ms / steps
    value = ncol / (ms_per_step * 1e-3)

    # ---- kernel launches per step (library timer on plain calls, counts only)
    lib.speedy_timer_enable(64)
    for _ in range(4):
        phys.compute_tendencies(*case)
    torch.cuda.synchronize()
    tm = _native.timer_read()
    lib.speedy_timer_disable()
    launches = {k: v[1] // 4 for k, v in tm.items() if v[1]}
    # The dominant (forward) kernel's average launch time inside the timed region:
    # HIP events around the replays of the same graph, per launch.
    fwd_ms = ms_per_step if launches.get("levelset", 0) == 0 else None
    if fwd_ms is None:
        # two-launch path: time the forward launches alone is not separable from the
        # graph; fall back to the event timer around each launch on plain calls
        lib.speedy_timer_enable(2 * 200 + 8)
        for _ in range(200):
            phys.compute_tendencies(*case)
        torch.cuda.synchronize()
        tm = _native.timer_read()
        lib.speedy_timer_disable()
        fwd_ms = tm["fwd"][0] / max(1, tm["fwd"][1])
    bytes_launch = FWD_BYTES_PER_COL(kx) * ncol
    achieved = bytes_launch / (fwd_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fwd_kernel_T85_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    out = {
        "metric": METRIC, "value": value, "unit": "columns/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "SPEEDY physics forward, T85 256x128 = 32768 columns x 8 levels, land recipe seed 1",
                   "columns": ncol, "levels": kx, "pass": "fwd", "graph_calls": G},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "speedy::fwd_kernel<8, false, true>", "kernel_ms": fwd_ms,
                     "launches_per_step": launches, "algorithmic_bytes_per_launch": bytes_launch},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case_np, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
