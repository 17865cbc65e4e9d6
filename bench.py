#!/usr/bin/env python3
"""Benchmark: FP64 Karma–Rappel Model C on one MI355X, Gcell-steps/s.

Workload (the headline configuration): 10080² cells as 80×80 blocks of 126²
(101,606,400 interior cells), h = 0.4, dt = 0.008, λ = 3.19, D = 2.0, ε₄ = 0.05,
u = −0.55, 64 seeds of radius 5h, mirror faces, Euler–Maruyama with b = 0 (the
plain path).  One step = one fused RHS + stage launch over every block; state
and inputs are resident in HBM before the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--noise B] [--no-cpu]

Prints ONE JSON line (rank 0).  Extra diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Model C Gcell-steps/s (FP64, % HBM roofline); batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STEP = 32  # phi, u read + phi, u written, f64


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_gpu(args, C):
    import numpy as np
    import torch

    import legendre_amd as lg

    stream = torch.cuda.Stream()  # default priority
    ctx = lg.Context(0, stream.cuda_stream)
    ctx.set_tuning("modelc_kernel", {"auto": 0, "percell": 1, "stream": 2}[args.kernel])
    ctx.set_tuning("modelc_strips", args.strips)
    ctx.set_tuning("modelc_prefetch", args.prefetch)
    grid = lg.CartesianGrid(ctx, [C.cells] * 2, [C.block] * 2, [0.0, 0.0], [C.h, C.h])
    params = lg.ModelCParams(C.eps4, C.lambda_, C.d_thermal, C.noise)
    sim = lg.Simulation.model_c(grid, params, wiener=C.noise > 0.0, seed=C.seed,
                                mode=lg.LGD_MODE_FUSED if args.mode == "fused" else lg.LGD_MODE_PER_INTENT)
    t0 = time.time()
    sim.initialize_grains(C.grains(), C.delta)
    log(f"[bench] initial condition on {C.cells}^2 in {time.time() - t0:.1f}s")
    sim.step(C.dt, args.warmup)
    ctx.sync()
    torch.cuda.synchronize()
    events = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for s in range(args.steps):
        events[s].record(stream)
        sim.step(C.dt, 1)
    events[args.steps].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = events[0].elapsed_time(events[args.steps])
    per = [events[s].elapsed_time(events[s + 1]) for s in range(args.steps)]
    phi = sim.field(0).download(0, 4 * 128 * 128)
    if not np.all(np.isfinite(phi)):
        raise RuntimeError("non-finite state after the timed steps")
    sim.close()
    if args.calibrate:
        # A kernel of known traffic for scaling profiler byte counters: one
        # axpy over 2^27 f64 reads 2 GiB and writes 1 GiB.
        n = 1 << 27
        x, y = lg.DeviceBuffer(ctx, n), lg.DeviceBuffer(ctx, n)
        grid.axpy(x, y, 0.5)
        ctx.sync()
        log(f"[bench] calibration axpy: n={n}, algorithmic fetch {16 * n} B, write {8 * n} B")
        x.free()
        y.free()
    ctx.close()
    return total_ms, per, wall


def read_traffic(profile: str):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (each scaled by a copy kernel's reading)."""
    if not os.path.exists(profile):
        return None
    with open(profile) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def run_cpu_baseline(C, budget_s: float):
    """The oracle (a C++ port of the reference's serial semantics, OpenMP over
    blocks like the RayonScheduler) on this host, on a bounded number of steps
    of the same workload."""
    import numpy as np

    import oracle

    cores = len(os.sched_getaffinity(0))
    threads = max(1, min(16, cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    g = oracle.grid([C.cells] * 2, [C.block] * 2, [0.0, 0.0], [C.h, C.h])
    p = oracle.params(C.eps4, C.lambda_, C.d_thermal, C.noise)
    n = oracle.field_len(g, 1)
    phi = np.zeros(n)
    u = np.zeros(n)
    oracle.initialize_grains(g, phi, u, C.grains(), C.delta)
    t = time.perf_counter()
    tt = oracle.modelc_run(g, p, phi, u, 1, C.dt, seed=C.seed, wiener=C.noise > 0, nthreads=threads)
    first = time.perf_counter() - t
    steps = int(max(1, min(50, budget_s // max(first, 1e-3))))
    t = time.perf_counter()
    oracle.modelc_run(g, p, phi, u, steps, C.dt, t0=tt, seed=C.seed, wiener=C.noise > 0, nthreads=threads)
    el = time.perf_counter() - t
    return {
        "value": C.interior_cells * steps / el / 1e9,
        "unit": "Gcell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (C++ restatement of SerialScheduler semantics, OpenMP over blocks) on the "
                  f"same {C.cells}^2 Model C state, {steps} steps after 1 warm-up step, "
                  f"{el:.1f}s, {threads} threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--noise", type=float, default=0.0)
    ap.add_argument("--mode", choices=["fused", "per-intent"], default="fused")
    ap.add_argument("--cells", type=int, default=10080)
    ap.add_argument("--kernel", choices=["auto", "percell", "stream"], default="auto")
    ap.add_argument("--strips", type=int, default=0, help="row strips per block (0 = default)")
    ap.add_argument("--prefetch", type=int, default=0, choices=[0, 1, 2],
                    help="rows the streaming kernel loads ahead (0 = default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one axpy of known traffic (for profiler byte-counter scaling)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this backend "
              "runs on a single MI355X (multi-GPU sharding is a design note, not a build)",
              file=sys.stderr)
        return 2

    from legendre_amd.workloads import ModelCConfig

    C = ModelCConfig(cells=args.cells, noise=args.noise)
    total_ms, per, wall = run_gpu(args, C)
    ms_per_step = total_ms / args.steps
    value = C.interior_cells * args.steps / (total_ms / 1e3) / 1e9
    avg_launch_ms = sum(per) / len(per)
    bytes_per_launch = BYTES_PER_CELL_STEP * C.interior_cells
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9
    log(f"[bench] {args.steps} steps: {total_ms:.3f} ms on-stream, {wall * 1e3:.3f} ms wall; "
        f"per launch min {min(per):.4f} max {max(per):.4f} ms")
    traffic = read_traffic(os.path.join(ROOT, "profiles", "modelc_stream_kernel_10080sq.md"))
    if args.cells != 10080 or args.noise != 0.0 or args.kernel == "percell" or args.mode != "fused":
        traffic = None  # the committed profile covers the default workload only
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(C, args.cpu_budget)
        log(f"[bench] cpu baseline: {cpu}")
    workload = f"model_c_{C.cells}sq_{C.block}blocks_b{C.noise:g}_{args.mode}"
    if args.kernel != "auto" or args.strips or args.prefetch:
        workload += f"_{args.kernel}_s{args.strips}_p{args.prefetch}"
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "cells": C.interior_cells,
            "grid": f"{C.cells}x{C.cells}",
            "blocks": f"{C.cells // C.block}x{C.cells // C.block} of {C.block}x{C.block}",
            "h": C.h,
            "dt": C.dt,
            "lambda": C.lambda_,
            "D": C.d_thermal,
            "eps4": C.eps4,
            "u0": -C.delta,
            "seeds": C.nseeds,
            "noise_b": C.noise,
            "integrator": "euler_maruyama",
            "parallelism": "single_gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "modelc_step_kernel" if args.kernel == "percell" else "modelc_stream_kernel",
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_launch_ms, 5),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
