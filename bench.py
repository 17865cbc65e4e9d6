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


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_gpu(args, C):
    import numpy as np
    import torch

    import legendre_amd as lg

    stream = torch.cuda.Stream()  # default priority
    ctx = lg.Context(0, stream.cuda_stream)
    ctx.set_tuning("modelc_kernel", {"auto": 0, "percell": 1, "stream": 2, "pair": 3}[args.kernel])
    ctx.set_tuning("modelc_strips", args.strips)
    ctx.set_tuning("modelc_prefetch", args.prefetch)
    grid = lg.CartesianGrid(ctx, [C.cells] * 2, [C.block] * 2, [0.0, 0.0], [C.h, C.h])
    params = lg.ModelCParams(C.eps4, C.lambda_, C.d_thermal, C.noise)
    sim = lg.Simulation.model_c(grid, params, wiener=C.noise > 0.0, seed=C.seed,
                                mode=lg.LGD_MODE_FUSED if args.mode == "fused" else lg.LGD_MODE_PER_INTENT)
    if args.graph:
        sim.set_graph_steps(args.graph)
    if args.integrator != "euler":
        sim.set_integrator(args.integrator)
    t0 = time.time()
    if args.oriented:
        # Multi-grain Model C: each seed a grain of its own θ₀ ∈ U[−π/4, π/4), the same
        # counter-keyed draws as the positions.
        from legendre_amd.workloads import mix_key_np, unit_open_np

        g3 = C.grains()
        th = np.array([(unit_open_np(mix_key_np(C.seed, [i, 2]))[0] - 0.5) * (np.pi / 2)
                       for i in range(len(g3))])
        sim.initialize_grains_oriented(np.column_stack([g3, th]), C.delta)
    else:
        sim.initialize_grains(C.grains(), C.delta)
    log(f"[bench] initial condition on {C.cells}^2 in {time.time() - t0:.1f}s")
    if args.autotune:
        r = ctx.autotune_modelc_step(grid, params, sim.field(0), sim.field(1), C.dt,
                                     wiener=C.noise > 0.0, reps=5)
        log(f"[bench] autotuned launch shape: {r}")
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
    ceiling = measure_ceiling(lg, ctx, stream)
    log(f"[bench] streaming ceiling (copy, 16-byte nontemporal): {ceiling}")
    ctx.close()
    return total_ms, per, wall, ceiling


def measure_ceiling(lg, ctx, stream, n: int = 1 << 28, reps: int = 5):
    """The device's streaming rate: lgd_stream_copy over n f64 (reads 8n B, writes 8n
    B) with 1, 2, 4 and 8 16-byte vectors in flight per lane, each the best of `reps`
    launches timed by HIP events on the launch stream.  One launch of the kVec = 4
    form also serves profiler byte-counter calibration (known traffic)."""
    import ctypes as C

    import torch

    from legendre_amd import _abi

    x, y = lg.DeviceBuffer(ctx, n), lg.DeviceBuffer(ctx, n)
    rates = {}
    for vec in (1, 2, 4, 8):
        best = None
        for r in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            _abi.check(ctx._lib.lgd_stream_copy(ctx.handle, x.handle, y.handle, vec), ctx.handle)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            if r > 0:  # the first launch warms the pages
                best = ms if best is None else min(best, ms)
        rates[vec] = round(16 * n / (best / 1e3) / 1e9, 1)
    x.free()
    y.free()
    vec = max(rates, key=rates.get)
    return {"gbs": rates[vec], "vec": vec, "by_vec": rates, "bytes": 16 * n}


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


def roofline(nbytes: float, seconds: float, kernel: str, unit_bytes: str) -> dict:
    """HBM roofline of a whole workload: algorithmic bytes over wall seconds."""
    achieved = nbytes / seconds / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
            "algorithmic_bytes": unit_bytes}


def cpu_threads() -> int:
    """Every core of this process's affinity mask, within OMP_NUM_THREADS when the
    host sets it (the CPU share the machine grants this process)."""
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(cores, int(env))) if env.isdigit() and int(env) > 0 else cores


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle():
    """Load the oracle built for this host's CPU (-march=native, same bits) for timing."""
    import tempfile

    import oracle

    t = time.perf_counter()
    path = oracle.build_native(tempfile.mkdtemp(prefix="lgd_oracle_native_"))
    oracle.load(path)
    log(f"[bench] oracle built with -march=native in {time.perf_counter() - t:.1f}s")
    return oracle


def run_extras(args, plain_value: float) -> dict:
    """The metric's "batched & adjoint vs plain" legs at reduced length, each with its
    own roofline: batched AMR (60 coarse steps, one per-patch chunk), the ensemble
    (100 steps, 2 per-block steps) and the adjoint gradient (200 steps, 45 slots)."""
    import argparse

    base = dict(vars(args))

    def ns(**kw):
        d = dict(base)
        d.update(kw)
        return argparse.Namespace(**d)

    legs = [
        ("batched_amr", run_amr, ns(amr_steps=60, per_block_chunks=1, graph=0, kernel="auto")),
        ("batched_ensemble", run_ensemble,
         ns(ens_steps=100, per_block_steps=2, no_cpu=True, kernel="auto", strips=0, prefetch=0)),
        ("adjoint", run_adjoint, ns(adj_steps=200, snapshots=0, no_cpu=True, adj_kernel="auto", adj_rows=0)),
    ]
    out = {}
    for name, fn, a in legs:
        t = time.perf_counter()
        line = fn(a)
        ratio = {"batched_amr": ("speedup_vs_per_patch_launches", line.get("amr", {})),
                 "batched_ensemble": ("speedup_vs_per_block", line.get("ensemble", {})),
                 "adjoint": ("gradient_over_forward", line.get("adjoint", {}))}[name]
        out[name] = {
            "workload": line["config"]["workload"],
            "value": line["value"],
            "unit": line["unit"],
            "ms_per_step": line["ms_per_step"],
            "vs_plain": round(line["value"] / plain_value, 4),
            ratio[0]: ratio[1].get(ratio[0]),
            "roofline": {k: line["roofline"][k] for k in ("achieved", "peak", "unit", "frac", "kernel",
                                                           "algorithmic_bytes")},
            "seconds": round(time.perf_counter() - t, 1),
        }
        log(f"[bench] {name}: {out[name]}")
    return out


def run_cpu_baseline(C, budget_s: float):
    """The oracle (a C++ port of the reference's serial semantics, OpenMP over
    blocks like the RayonScheduler) on this host, on a bounded number of steps
    of the same workload, in both of its builds (-march=x86-64-v3, and -march=native
    built here); the faster is the baseline."""
    import numpy as np

    import oracle

    threads = cpu_threads()
    builds = [("x86-64-v3", oracle.LIB_PATH)]
    try:
        import tempfile

        t = time.perf_counter()
        builds.append(("native", oracle.build_native(tempfile.mkdtemp(prefix="lgd_oracle_native_"))))
        log(f"[bench] oracle built with -march=native in {time.perf_counter() - t:.1f}s")
    except Exception as e:  # noqa: BLE001 - the portable build still times the baseline
        log(f"[bench] -march=native build failed ({e}); timing the portable build only")
    g = oracle.grid([C.cells] * 2, [C.block] * 2, [0.0, 0.0], [C.h, C.h])
    p = oracle.params(C.eps4, C.lambda_, C.d_thermal, C.noise)
    n = oracle.field_len(g, 1)
    runs = {}
    for name, path in builds:
        oracle.load(path)
        phi = np.zeros(n)
        u = np.zeros(n)
        oracle.initialize_grains(g, phi, u, C.grains(), C.delta)
        t = time.perf_counter()
        tt = oracle.modelc_run(g, p, phi, u, 1, C.dt, seed=C.seed, wiener=C.noise > 0, nthreads=threads)
        first = time.perf_counter() - t
        steps = int(max(1, min(50, (budget_s / len(builds)) // max(first, 1e-3))))
        t = time.perf_counter()
        oracle.modelc_run(g, p, phi, u, steps, C.dt, t0=tt, seed=C.seed, wiener=C.noise > 0, nthreads=threads)
        el = time.perf_counter() - t
        runs[name] = (C.interior_cells * steps / el / 1e9, steps, el)
        log(f"[bench] cpu baseline, {name} build: {runs[name][0]:.4f} Gcell-steps/s ({steps} steps, {el:.1f}s)")
    best = max(runs, key=lambda k: runs[k][0])
    value, steps, el = runs[best]
    others = ", ".join(f"{k}: {v[0]:.3f}" for k, v in runs.items() if k != best)
    return {
        "value": value,
        "unit": "Gcell-steps/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"oracle (C++ restatement of SerialScheduler semantics, OpenMP over blocks, -O3 "
                  f"-march={best}) on the same {C.cells}^2 Model C state, {steps} steps after 1 "
                  f"warm-up step, {el:.1f}s, {threads} threads on {cpu_model()}"
                  + (f"; other build {others} Gcell-steps/s" if others else ""),
    }


def run_adjoint(args) -> dict:
    """Calibration gradient (the adjoint workload): L = ½‖φ_T − φ_target‖² after
    2000 steps on 1024² (16×16 blocks of 64²), dL/d(λ, D, ε₄, φ₀, u₀) by the device
    adjoint with binomial checkpointing, timed against one plain forward run of the
    same steps; checked against central differences of the device loss and the
    oracle's tangent-linear run (four directions carried at once)."""
    import numpy as np
    import torch

    import legendre_amd as lg
    from legendre_amd.workloads import AdjointConfig

    A = AdjointConfig(steps=args.adj_steps or AdjointConfig.steps,
                      snapshots=args.snapshots or AdjointConfig.snapshots)
    ctx = lg.Context(0)
    ctx.set_tuning("modelc_kernel", {"auto": 0, "percell": 1, "stream": 2, "pair": 3}[args.kernel])
    ctx.set_tuning("modelc_strips", args.strips)
    ctx.set_tuning("adjoint_kernel", {"auto": 0, "passes": 1, "stream": 2}[args.adj_kernel])
    ctx.set_tuning("adjoint_rows", args.adj_rows)
    grid = lg.CartesianGrid(ctx, [A.cells] * 2, [A.block] * 2, [0.0, 0.0], [A.h, A.h])
    p = lg.ModelCParams(A.eps4, A.lambda_, A.d_thermal)
    pt = lg.ModelCParams(A.target_eps4, A.target_lambda, A.target_d)
    grains = A.grains()

    def initial(params):
        sim = lg.Simulation.model_c(grid, params)
        sim.initialize_grains(grains, A.delta)
        return sim

    sim0 = initial(p)
    phi0, u0 = grid.alloc_field(), grid.alloc_field()
    phi0.copy_from(sim0.field(0))
    u0.copy_from(sim0.field(1))
    simt = initial(pt)
    simt.step(A.dt, A.steps)
    target = grid.alloc_field()
    target.copy_from(simt.field(0))
    simt.close()
    ctx.sync()
    if args.autotune:
        r = ctx.autotune_modelc_step(grid, p, phi0, u0, A.dt, reps=20)
        log(f"[adjoint] autotuned launch shape: {r}")
    # One plain forward run (warm-up run first).
    sim0.step(A.dt, 20)
    ctx.sync()
    fwd = []
    for _ in range(2):
        sim = initial(p)
        ctx.sync()
        t = time.perf_counter()
        sim.step(A.dt, A.steps)
        ctx.sync()
        fwd.append(time.perf_counter() - t)
        sim.close()
    t_fwd = min(fwd)
    log(f"[adjoint] forward {A.steps} steps: {t_fwd * 1e3:.2f} ms")
    cal = lg.ModelCCalibration(grid, A.steps, A.dt, snapshots=A.snapshots)
    gphi, gu = grid.alloc_field(), grid.alloc_field()
    runs, bwd = [], []
    for _ in range(2):
        ctx.sync()
        t = time.perf_counter()
        loss = cal.forward(p, phi0, u0, target)
        tb = time.perf_counter()
        grad = cal.backward(gphi, gu)
        ctx.sync()
        runs.append(time.perf_counter() - t)
        bwd.append(time.perf_counter() - tb)
    t_grad = min(runs)
    t_bwd = min(bwd)
    nfwd, nrev = cal.counts()
    log(f"[adjoint] gradient: {t_grad * 1e3:.2f} ms ({nfwd} forward + {nrev} reverse steps; "
        f"backward sweep {t_bwd * 1e3:.2f} ms), loss {loss:.10e}, grad {grad}")
    # Central differences of the device loss.
    base = dict(eps4=A.eps4, lambda_=A.lambda_, d_thermal=A.d_thermal)
    fd = {}
    for k, name in enumerate(["lambda_", "d_thermal", "eps4"]):
        errs = {}
        for h in (1e-6, 1e-7):
            vals = []
            for sgn in (1.0, -1.0):
                q = dict(base)
                q[name] += sgn * h
                vals.append(cal.forward(lg.ModelCParams(q["eps4"], q["lambda_"], q["d_thermal"]),
                                        phi0, u0, target))
            d = (vals[0] - vals[1]) / (2 * h)
            errs[f"h={h:g}"] = abs(d - grad[k]) / abs(grad[k])
        fd[name.rstrip("_")] = errs
    log(f"[adjoint] finite-difference relative errors: {fd}")
    tlm = None
    crossings = None
    if not args.no_cpu:
        import oracle

        g = oracle.grid([A.cells] * 2, [A.block] * 2, [0.0, 0.0], [A.h, A.h])
        op = oracle.params(A.eps4, A.lambda_, A.d_thermal)
        ph, uu, tg = phi0.download(), u0.download(), target.download()
        ph0, uu0 = ph.copy(), uu.copy()
        n = ph.size
        rng = np.random.default_rng(0)
        vu = rng.standard_normal(n)
        dth = np.zeros((4, 3))
        dth[0, 0] = dth[1, 1] = dth[2, 2] = 1.0
        dp, du = np.zeros((4, n)), np.zeros((4, n))
        du[3] = vu
        t = time.perf_counter()
        done, chunk = 0, 250
        while done < A.steps:
            m = min(chunk, A.steps - done)
            oracle.modelc_tlm_run(g, op, ph, uu, dp, du, dth, m, A.dt)
            done += m
            log(f"[adjoint] oracle tangent-linear run: {done}/{A.steps} steps, "
                f"{time.perf_counter() - t:.0f}s")
        mask = np.zeros(n, dtype=bool)
        b = A.block + 2
        mask.reshape(-1, b, b)[:, 1:-1, 1:-1] = True
        r = (ph - tg)[mask]
        dl = np.array([float(np.sum(r * dp[k][mask])) for k in range(4)])
        gu_h = gu.download()
        adj = np.array([grad[0], grad[1], grad[2], float(np.sum(gu_h[mask] * vu[mask]))])
        tlm = {
            "rel_err": [float(abs(adj[k] - dl[k]) / abs(dl[k])) for k in range(4)],
            "directions": ["lambda", "D", "eps4", "u0 (random field)"],
            "cores": cpu_threads(),
            "seconds": round(time.perf_counter() - t, 1),
        }
        log(f"[adjoint] tangent-linear check: {tlm}")
        # Face evaluations whose flat branch |∇φ|⁴ <= tol differs between the θ − h
        # and θ + h runs (finite_volume.rs:129-142): a central difference with any
        # carries a jump of L, whatever h is.
        crossings = {}
        t = time.perf_counter()
        cores = tlm["cores"]
        for name in ("lambda_", "d_thermal", "eps4"):
            per = {}
            for h in (1e-6, 1e-7):
                lo, hi = dict(base), dict(base)
                lo[name] -= h
                hi[name] += h
                per[f"h={h:g}"] = oracle.modelc_branch_crossings(
                    g, oracle.params(lo["eps4"], lo["lambda_"], lo["d_thermal"]),
                    oracle.params(hi["eps4"], hi["lambda_"], hi["d_thermal"]), ph0, uu0, A.steps, A.dt,
                    cores)
                log(f"[adjoint] branch crossings {name} h={h:g}: {per[f'h={h:g}']} "
                    f"({time.perf_counter() - t:.0f}s)")
            crossings[name.rstrip("_")] = per
    cal.close()
    sim0.close()
    ctx.close()
    value = A.interior_cells * A.steps / t_grad / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": A.steps,
        "warmup": 1,
        "ms_per_step": round(t_grad / A.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"modelc_adjoint_{A.cells}sq_{A.block}blocks_{A.steps}steps_{A.snapshots}snapshots",
            "cells": A.interior_cells,
            "theta": [A.lambda_, A.d_thermal, A.eps4],
            "target_theta": [A.target_lambda, A.target_d, A.target_eps4],
            "seeds": A.nseeds,
        },
        "roofline": roofline(32.0 * A.interior_cells * nfwd + 48.0 * A.interior_cells * nrev, t_grad,
                             "the gradient (forward steps and adj_stream_kernel reverse steps)",
                             "32 B per forward cell-step, 48 B per reverse cell-step"),
        "adjoint": {
            "forward_ms": round(t_fwd * 1e3, 3),
            "gradient_ms": round(t_grad * 1e3, 3),
            "gradient_over_forward": round(t_grad / t_fwd, 3),
            "backward_ms": round(t_bwd * 1e3, 3),
            "reverse_kernel": args.adj_kernel,
            "forward_steps": nfwd,
            "reverse_steps": nrev,
            "loss": loss,
            "grad": [float(x) for x in grad],
            "fd_rel_err": fd,
            "fd_branch_crossings": crossings,
            "tlm": tlm,
        },
    }
    return line


def run_ensemble(args) -> dict:
    """Batched stochastic ensemble: 512 Model C realizations of uneven sizes
    (blocks of 64²) advanced by one launch per step over every block of every
    realization, against one launch per block; solid-fraction statistics checked
    against the oracle on an evenly spaced subset of the same realizations."""
    import numpy as np

    import legendre_amd as lg
    from legendre_amd.workloads import EnsembleConfig

    E = EnsembleConfig(steps=args.ens_steps or EnsembleConfig.steps)
    members = E.members()
    ctx = lg.Context(0)
    ctx.set_tuning("modelc_kernel", {"auto": 0, "percell": 1, "stream": 2, "pair": 3}[args.kernel])
    if args.strips:
        ctx.set_tuning("modelc_strips", args.strips)
    if args.prefetch:
        ctx.set_tuning("modelc_prefetch", args.prefetch)
    reals = [lg.Realization((m["side"], m["side"]), m["eps4"], E.lambda_, E.d_thermal, m["b"],
                            seed=m["seed"]) for m in members]
    cells = sum(m["side"] ** 2 for m in members)

    def build():
        ens = lg.Ensemble(ctx, reals, block_cells=(E.block, E.block), spacing=(E.h, E.h))
        for r, m in enumerate(members):
            ens.initialize_grains(r, E.grains(m["side"]), m["delta"])
        ctx.sync()
        return ens

    t = time.perf_counter()
    ens = build()
    log(f"[ensemble] {len(reals)} realizations, {cells / 1e6:.1f} M cells, {ens.num_blocks} blocks, "
        f"set up in {time.perf_counter() - t:.1f}s")
    ctx.sync()
    t = time.perf_counter()
    done = 0
    while done < E.steps:
        m = min(250, E.steps - done)
        ens.step(E.dt, m)
        ctx.sync()
        done += m
        log(f"[ensemble] batched: {done}/{E.steps} steps, {time.perf_counter() - t:.1f}s")
    t_batched = (time.perf_counter() - t) / E.steps
    frac = ens.solid_fraction()
    ens.close()
    # One launch per block (the per-block dispatch of the reference's scheduler).
    ens = build()
    ens.step(E.dt, 1, per_block=True)
    ctx.sync()
    pb = max(1, args.per_block_steps)
    t = time.perf_counter()
    ens.step(E.dt, pb, per_block=True)
    ctx.sync()
    t_per_block = (time.perf_counter() - t) / pb
    ens.close()
    log(f"[ensemble] per step: batched {t_batched * 1e3:.3f} ms, per-block launches "
        f"{t_per_block * 1e3:.3f} ms ({t_per_block / t_batched:.1f}x)")
    mean, var = float(np.mean(frac)), float(np.var(frac, ddof=1))
    se = (var / len(frac)) ** 0.5
    cpu = None
    if not args.no_cpu:
        import oracle

        idx = list(range(0, len(members), max(1, len(members) // args.cpu_members)))[: args.cpu_members]
        t = time.perf_counter()
        ref = []
        for n_done, r in enumerate(idx):
            m = members[r]
            g = oracle.grid([m["side"]] * 2, [E.block] * 2, [0.0, 0.0], [E.h, E.h])
            p = oracle.params(m["eps4"], E.lambda_, E.d_thermal, m["b"])
            n = oracle.field_len(g, 1)
            phi, u = np.zeros(n), np.zeros(n)
            oracle.initialize_grains(g, phi, u, E.grains(m["side"]), m["delta"])
            oracle.modelc_run(g, p, phi, u, E.steps, E.dt, seed=m["seed"], wiener=True,
                              nthreads=cpu_threads())
            b = E.block + 2
            inner = phi.reshape(-1, b, b)[:, 1:-1, 1:-1]
            ref.append(float(np.count_nonzero(inner > 0.0)) / m["side"] ** 2)
            log(f"[ensemble] oracle realization {r} ({m['side']}^2): {n_done + 1}/{len(idx)}, "
                f"{time.perf_counter() - t:.0f}s")
        ref = np.array(ref)
        sub = frac[idx]
        sub_se = (float(np.var(sub, ddof=1)) / len(sub)) ** 0.5 if len(sub) > 1 else 0.0
        cpu = {
            "members": idx,
            "max_abs_diff": float(np.max(np.abs(sub - ref))),
            "mean_gpu": float(np.mean(sub)),
            "mean_oracle": float(np.mean(ref)),
            "var_gpu": float(np.var(sub, ddof=1)) if len(sub) > 1 else 0.0,
            "var_oracle": float(np.var(ref, ddof=1)) if len(ref) > 1 else 0.0,
            "within_3se": bool(abs(float(np.mean(sub)) - float(np.mean(ref))) <= 3 * sub_se),
            "seconds": round(time.perf_counter() - t, 1),
        }
        log(f"[ensemble] oracle check: {cpu}")
    ctx.close()
    value = cells / t_batched / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": E.steps,
        "warmup": 0,
        "ms_per_step": round(t_batched * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"modelc_ensemble_{len(reals)}x_{E.block}blocks_{E.steps}steps",
            "cells": cells,
            "blocks": int(sum((m["side"] // E.block) ** 2 for m in members)),
            "noise_b": "U[0.005,0.02]",
        },
        "roofline": roofline(32.0 * cells, t_batched, "modelc_stream_kernel (block pairs, device domain table)",
                             "32 B per cell-step"),
        "ensemble": {
            "batched_ms_per_step": round(t_batched * 1e3, 4),
            "per_block_ms_per_step": round(t_per_block * 1e3, 4),
            "speedup_vs_per_block": round(t_per_block / t_batched, 2),
            "solid_fraction_mean": mean,
            "solid_fraction_var": var,
            "solid_fraction_se": se,
            "oracle": cpu,
        },
    }
    return line


def run_jump(args) -> int:
    """The jump-diffusion model of the reference's jump-driver tests
    (WienerJump<1, 1>: dw = μ dt + σ dW; on a fire count += 1, jsum += payoff, at a
    per-path intensity λ) on 2^26 paths of one cell each, blocks of 2^16 paths,
    100 Euler–Maruyama steps. A step is one lgd_apply_drivers launch: Time, Wiener(0)
    and Jump(0) over the three state fields. The vector fields are constant and are
    written once before the timed region, as are the intensities (λ ∈ U[0.5, 5))."""
    import numpy as np
    import torch

    import legendre_amd as lg

    paths, block, dt, seed = 1 << 26, 1 << 16, 0.01, 42
    steps = args.jump_steps or 100
    mu, sigma, payoff = 0.1, 0.5, 2.0
    stream = torch.cuda.Stream()
    ctx = lg.Context(0, stream.cuda_stream)
    grid = lg.CartesianGrid(ctx, [paths], [block], [0.0], [1.0])
    lam = 0.5 + 4.5 * np.random.default_rng(seed).random(paths)

    def const(v):
        b = lg.DeviceBuffer(ctx, paths)
        b.upload(np.full(paths, v))
        return b

    w, count, jsum = (lg.DeviceBuffer(ctx, paths) for _ in range(3))
    drift, amp, one, pay = const(mu), const(sigma), const(1.0), const(payoff)
    rate = lg.DeviceBuffer(ctx, paths)
    rate.upload(lam)
    drivers = [lg.Driver(lg.LGD_DRIVER_TIME, 0, [drift, None, None]),
               lg.Driver(lg.LGD_DRIVER_WIENER, 0, [amp, None, None]),
               lg.Driver(lg.LGD_DRIVER_JUMP, 0, [None, one, pay], rate)]
    fields = [w, count, jsum]
    t = 0.0
    for _ in range(args.warmup):
        grid.apply_drivers(fields, [0, 0, 0], drivers, dt, seed, lg.time_bits(t))
        t = t + dt
    torch.cuda.synchronize()
    events = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    for s in range(steps):
        events[s].record(stream)
        grid.apply_drivers(fields, [0, 0, 0], drivers, dt, seed, lg.time_bits(t))
        t = t + dt
    events[steps].record(stream)
    torch.cuda.synchronize()
    total_ms = events[0].elapsed_time(events[steps])
    per = [events[s].elapsed_time(events[s + 1]) for s in range(steps)]
    c = count.download()
    fired = float(c.mean())
    ctx.close()
    avg_ms = sum(per) / len(per)
    bytes_per_launch = 88 * paths  # 3 fields read + written, 5 vector-field / intensity reads
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu:
        oracle = native_oracle()
        threads = cpu_threads()
        sample = 1 << 22
        g = oracle.grid([sample], [block], [0.0], [1.0])
        z = [np.zeros(sample) for _ in range(3)]
        tt = time.perf_counter()
        oracle.jump_diffusion_run(g, lam[:sample].copy(), *z, 1, dt, seed, mu, sigma, payoff,
                                  nthreads=threads)
        first = time.perf_counter() - tt
        n = int(max(1, min(100, 10.0 // max(first, 1e-3))))
        tt = time.perf_counter()
        oracle.jump_diffusion_run(g, lam[:sample].copy(), *z, n, dt, seed, mu, sigma, payoff,
                                  nthreads=threads)
        el = time.perf_counter() - tt
        cpu = {"value": sample * n / el / 1e9, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"oracle jump-diffusion (-march=native), {sample} paths x {n} steps, {el:.1f}s, "
                         f"{threads} threads"}
    line = {
        "metric": METRIC,
        "value": round(paths * steps / (total_ms / 1e3) / 1e9, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"jump_diffusion_{paths}paths_{block}blocks_{steps}steps",
                   "paths": paths, "dt": dt, "mu": mu, "sigma": sigma, "payoff": payoff,
                   "lambda": "U[0.5,5)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "drivers_kernel", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": round(avg_ms, 5)},
        "jump": {"mean_count": fired},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_observe(args) -> int:
    """The observation seam on the headline workload: the 10080² Model C run with a
    ParquetObserver (phi, u) on an asynchronous ring of 3 pinned buffers, snapshots on
    step 1 and every 10 steps. The same steps are timed without and with the observer
    (the solver-side cost: the gather kernel on the simulation's stream; the copy runs
    on a copy stream and the files on a writer thread); then the writer is drained and
    its throughput reported. Files go to a temporary directory removed afterwards."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    import legendre_amd as lg
    from legendre_amd.workloads import ModelCConfig

    C = ModelCConfig(cells=args.cells, block=args.block, noise=args.noise)
    steps = args.steps
    stream = torch.cuda.Stream()
    ctx = lg.Context(0, stream.cuda_stream)
    grid = lg.CartesianGrid(ctx, [C.cells] * 2, [C.block] * 2, [0.0, 0.0], [C.h, C.h])
    params = lg.ModelCParams(C.eps4, C.lambda_, C.d_thermal, C.noise)
    sim = lg.Simulation.model_c(grid, params, wiener=C.noise > 0.0, seed=C.seed)
    sim.initialize_grains(C.grains(), C.delta)
    sim.step(C.dt, args.warmup)

    def timed(n):
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
        for k in range(n):
            ev[k].record(stream)
            sim.step(C.dt, 1)
        ev[n].record(stream)
        torch.cuda.synchronize()
        return [ev[k].elapsed_time(ev[k + 1]) for k in range(n)]

    plain = timed(steps)
    out = tempfile.mkdtemp(prefix="lgd_observe_")
    try:
        obs = lg.ParquetObserver(ctx, out, [("phi", 0), ("u", 1)], every=10, ring=3)
        sim.attach_observer(obs)
        t0 = time.perf_counter()
        observed = timed(steps)
        solver_s = time.perf_counter() - t0
        obs.flush()
        drain_s = time.perf_counter() - t0
        st = obs.stats()
        files = sorted(os.listdir(out))
        nbytes = sum(os.path.getsize(os.path.join(out, f)) for f in files)
        obs.close()
    finally:
        shutil.rmtree(out, ignore_errors=True)
    sim.close()
    ctx.close()
    first = args.warmup + steps  # the step index before the observed steps
    snap_steps = [k for k in range(steps) if first + k + 1 == 1 or (first + k + 1) % 10 == 0]
    # Column bytes the writer encoded: the static file's x, y, level, patch, then two
    # fields per snapshot file.
    col_bytes = 8 * C.interior_cells * (4 + 2 * (st["written"] - 1)) if st["written"] else 0
    base_ms = float(np.median(plain))
    snap_ms = [observed[k] for k in snap_steps] if snap_steps else []
    other_ms = [observed[k] for k in range(steps) if k not in snap_steps]
    cols = 2
    line = {
        "metric": METRIC,
        "value": round(C.interior_cells * steps / (sum(observed) / 1e3) / 1e9, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(sum(observed) / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"model_c_{C.cells}sq_observed_every10_ring3", "cells": C.interior_cells,
                   "fields": ["phi", "u"]},
        "observe": {
            "plain_ms_per_step": round(sum(plain) / steps, 5),
            "observed_ms_per_step": round(sum(observed) / steps, 5),
            "median_plain_step_ms": round(base_ms, 5),
            "snapshot_step_ms": [round(x, 4) for x in snap_ms],
            "median_other_step_ms": round(float(np.median(other_ms)), 5) if other_ms else None,
            # The gather reads and writes 16 B per cell and field; its time is estimated
            # as a snapshot step's excess over the median plain step (the static
            # file's extra columns make the first snapshot step longer).
            "gather_GBps_est": [round(32 * C.interior_cells / ((x - float(np.median(other_ms))) / 1e3) / 1e9, 1)
                                for x in snap_ms[1:] if other_ms and x > float(np.median(other_ms)) * 1.2],
            "snapshots": len(snap_steps),
            "written_files": st["written"],
            "dropped": st["dropped"],
            "failed": st["failed"],
            "column_bytes_per_snapshot": cols * 8 * C.interior_cells,
            "file_bytes_total": nbytes,
            "writer_seconds": round(st["write_seconds"], 3),
            "writer_GBps": round(col_bytes / st["write_seconds"] / 1e9, 3) if st["write_seconds"] else None,
            "solver_wall_s": round(solver_s, 3),
            "drain_wall_s": round(drain_s, 3),
        },
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_heat3d(args) -> int:
    """The 3-D heat configuration: 128³ as 4×4×4 blocks of 32³, periodic, 7-point
    Laplacian, forward Euler, one fused launch per step, 100 steps after the warm-up.
    16 B per cell-step (u read and written); the 16.8 MB field stays in the 256 MB
    Infinity Cache, so the HBM fraction is reported, not a bound."""
    import numpy as np
    import torch

    import legendre_amd as lg
    from legendre_amd.workloads import HEAT3D as C, to_slabs

    cells, block = [C.cells] * 3, [C.block] * 3
    stream = torch.cuda.Stream()
    ctx = lg.Context(0, stream.cuda_stream)
    grid = lg.CartesianGrid(ctx, cells, block, [0.0] * 3, [C.dx] * 3, [True] * 3)
    u0 = to_slabs(C.initial_global(), cells, block, 1)
    sim = lg.Simulation.heat(grid, C.kappa)
    sim.field(0).upload(u0)
    sim.step(C.dt, args.warmup)
    torch.cuda.synchronize()
    steps = C.steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    for k in range(steps):
        ev[k].record(stream)
        sim.step(C.dt, 1)
    ev[steps].record(stream)
    torch.cuda.synchronize()
    per = [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]
    total = ev[0].elapsed_time(ev[steps])
    sim.close()
    ctx.close()
    n = C.cells ** 3
    avg = sum(per) / len(per)
    cpu = None
    if not args.no_cpu:
        oracle = native_oracle()
        threads = cpu_threads()
        og = oracle.grid(cells, block, [0.0] * 3, [C.dx] * 3, [True] * 3)
        u = u0.copy()
        t0 = time.perf_counter()
        oracle.heat_run(og, C.kappa, u, steps, C.dt, nthreads=threads)
        el = time.perf_counter() - t0
        cpu = {"value": n * steps / el / 1e9, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"oracle heat run (-march=native), the whole 128^3 workload ({steps} steps), "
                         f"{el:.2f}s, {threads} threads"}
    line = {
        "metric": METRIC,
        "value": round(n * steps / (total / 1e3) / 1e9, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "heat3d_128cubed_32blocks_periodic", "cells": n, "dt": C.dt},
        "roofline": {"bound": "hbm", "achieved": round(16 * n / (avg / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(16 * n / (avg / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "heat_step_kernel", "bytes_per_launch": 16 * n,
                     "avg_launch_ms": round(avg, 5),
                     "note": "the field fits the Infinity Cache; not HBM-bound"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_amr(args) -> dict:
    """Batched AMR: subcycled Model C on a 3-level Berger–Oliger hierarchy with
    regrids, every patch of a level advanced by one launch per stage, against one
    launch per patch on the same schedule.  Throughput counts cell updates:
    Σ_levels cells_l × substeps_l per coarse step (substeps 4^l: the stable dt
    shrinks with h²)."""
    import legendre_amd as lg
    from legendre_amd.workloads import AmrConfig

    A = AmrConfig(coarse_steps=args.amr_steps or AmrConfig.coarse_steps)
    ctx = lg.Context(0)
    ctx.set_tuning("modelc_kernel", {"auto": 0, "percell": 1, "stream": 2, "pair": 3}[args.kernel])
    ctx.set_tuning("amr_split_fill", getattr(args, "amr_split_fill", 0))
    ctx.set_tuning("modelc_strips", getattr(args, "strips", 0) or 0)  # level 0's strips (0: default)
    ctx.set_tuning("modelc_prefetch", getattr(args, "prefetch", 0) or 0)
    params = lg.ModelCParams(A.eps4, A.lambda_, A.d_thermal)
    pol = lg.RegridPolicy(A.every, A.buffer, A.efficiency, A.min_width, A.threshold, 0)
    grid = lg.CartesianGrid(ctx, [A.cells] * 2, [A.block] * 2, [0.0, 0.0], [A.h, A.h])
    ic = lg.Simulation.model_c(grid, params)
    ic.initialize_grains(A.grains(), A.delta)

    ctx_per_block = [False]  # set while the per-patch baseline runs

    def build():
        gg = lg.AmrGrid.from_patches(ctx, [A.cells] * 2, [A.block] * 2, [0.0, 0.0], [A.h, A.h],
                                     list(A.ratios), [])
        sim = lg.amr_simulation(gg, "modelc", params=params, subcycle=True, policy=pol)
        if args.graph and not ctx_per_block[0]:
            sim.set_graph_steps(args.graph)
        sim.field(0).copy_from(ic.field(0))
        sim.field(1).copy_from(ic.field(1))
        ctx.sync()
        return sim

    # Untimed warmup on a copy of the run, advanced past its second regrid (3 levels:
    # every kernel and pool size of the timed run used once); with --autotune the
    # launch shapes are tuned on that hierarchy.  The timed run starts from the
    # initial state as usual.
    tuning = None
    warmup_steps = A.every + 1
    probe = build()
    probe.step(A.dt, warmup_steps)
    if args.autotune:
        tuning = probe.autotune_amr(A.dt, reps=5)
        log(f"[amr] autotuned launch shapes: {tuning}")
    probe.close()

    def level_cells(sim):
        g = sim.amr
        cells = [0] * g.num_levels
        for b in range(g.num_blocks):
            lvl, lo, hi = g.patch(b)
            cells[lvl] += (hi[0] - lo[0]) * (hi[1] - lo[1])
        return cells

    def run(sim, chunks, tag):
        times, updates, shapes, first = [], [], [], []
        for c in range(chunks):
            t = time.perf_counter()
            sim.step(A.dt, 1)  # the regrid runs at the top of this step
            ctx.sync()
            t1 = time.perf_counter()
            sim.step(A.dt, A.every - 1)
            ctx.sync()
            times.append(time.perf_counter() - t)
            first.append(t1 - t)
            cells = level_cells(sim)
            updates.append(A.every * sum(n * 4 ** l for l, n in enumerate(cells)))
            shapes.append((sim.amr.num_blocks, cells))
            log(f"[amr] {tag}: coarse steps {(c + 1) * A.every}, {times[-1] * 1e3:.1f} ms "
                f"(regrid + first step {first[-1] * 1e3:.1f} ms), patches {sim.amr.num_blocks}, "
                f"cells per level {cells}")
        return times, updates, shapes, first

    nchunks = max(1, A.coarse_steps // A.every)
    sim = build()
    times, updates, shapes, first = run(sim, nchunks, "batched")
    nregrid, phases = sim.regrid_stats()
    log(f"[amr] {nregrid} regrids, seconds by phase: tag field to host {phases[0]:.3f}, "
        f"proposal {phases[1]:.3f}, new hierarchy {phases[2]:.3f}, migration {phases[3]:.3f}")
    sim.close()
    pb_chunks = min(nchunks, max(0, args.per_block_chunks))
    pb_times, pb_first, same = [], [], None
    if pb_chunks:
        ctx.set_tuning("amr_per_block", 1)
        ctx_per_block[0] = True  # the baseline launches eagerly, one launch per patch
        sim = build()
        pb_times, pb_updates, pb_shapes, pb_first = run(sim, pb_chunks, "per-patch launches")
        sim.close()
        ctx.set_tuning("amr_per_block", 0)
        same = all(pb_shapes[i] == shapes[i] for i in range(pb_chunks))
    total = sum(times)
    value = sum(updates) / total / 1e9
    speedup = sum(pb_times) / sum(times[:pb_chunks]) if pb_chunks else None
    speedup_steps = ((sum(pb_times) - sum(pb_first)) /
                     (sum(times[:pb_chunks]) - sum(first[:pb_chunks]))) if pb_chunks else None
    ic.close()
    ctx.close()
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": nchunks * A.every,
        "warmup": warmup_steps,
        "ms_per_step": round(total / (nchunks * A.every) * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"modelc_amr_{A.cells}sq_3levels_subcycled_{nchunks * A.every}coarse_steps",
            "base_cells": A.cells * A.cells,
            "ratios": list(A.ratios),
            "regrid_every": A.every,
            "graph_steps": args.graph,
            "seeds": A.nseeds,
            "noise_b": 0.0,
        },
        "roofline": roofline(32.0 * sum(updates), total, "every level stage, fill, restriction and regrid",
                             "32 B per cell update"),
        "amr": {
            "cell_updates": int(sum(updates)),
            "seconds": round(total, 3),
            "final_patches": shapes[-1][0],
            "final_cells_per_level": shapes[-1][1],
            "per_patch_chunks_compared": pb_chunks,
            "batched_ms_first_chunks": round(sum(times[:pb_chunks]) * 1e3, 3),
            "per_patch_ms_first_chunks": round(sum(pb_times) * 1e3, 3),
            "speedup_vs_per_patch_launches": round(speedup, 2) if speedup else None,
            "regrid_and_first_step_seconds": round(sum(first), 3),
            "regrids": nregrid,
            "regrid_phase_seconds": {"tag_field_to_host": round(float(phases[0]), 4),
                                     "proposal": round(float(phases[1]), 4),
                                     "new_hierarchy": round(float(phases[2]), 4),
                                     "migration": round(float(phases[3]), 4)},
            "speedup_excluding_regrid_steps": round(speedup_steps, 2) if speedup_steps else None,
            "same_hierarchy": same,
            "autotuned": tuning,
        },
    }
    return line


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["modelc", "adjoint", "ensemble", "amr", "jump", "observe", "heat3d"], default="modelc",
                    help="modelc: the headline plain step; adjoint: the calibration gradient; "
                         "ensemble: 512 batched realizations; amr: 3-level subcycled hierarchy")
    ap.add_argument("--jump-steps", type=int, default=0, help="jump workload steps (0 = 100)")
    ap.add_argument("--amr-steps", type=int, default=0, help="AMR coarse steps (0 = 500)")
    ap.add_argument("--amr-split-fill", type=int, default=0, choices=[0, 1],
                    help="1: the finest level's ghost fills in launches of their own")
    ap.add_argument("--per-block-chunks", type=int, default=2,
                    help="regrid intervals timed with one launch per patch (amr; 0 skips the baseline)")
    ap.add_argument("--ens-steps", type=int, default=0, help="ensemble steps (0 = 2000)")
    ap.add_argument("--per-block-steps", type=int, default=3,
                    help="steps timed with one launch per block (ensemble)")
    ap.add_argument("--cpu-members", type=int, default=16,
                    help="realizations re-run on the oracle (ensemble)")
    ap.add_argument("--adj-steps", type=int, default=0, help="adjoint workload steps (0 = 2000)")
    ap.add_argument("--snapshots", type=int, default=0, help="adjoint checkpoint slots (0 = 45)")
    ap.add_argument("--graph", type=int, default=0,
                    help="run the simulation's steps as replayed HIP graphs of this many steps "
                         "(modelc and amr workloads; 0 = eager launches)")
    ap.add_argument("--adj-kernel", choices=["auto", "passes", "stream"], default="auto",
                    help="reverse step: streaming kernel where the block shape allows (auto), "
                         "the three passes, or streaming required")
    ap.add_argument("--adj-rows", type=int, default=0,
                    help="rows per strip of the streaming reverse step (0 = chosen for the device)")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--noise", type=float, default=0.0)
    ap.add_argument("--mode", choices=["fused", "per-intent"], default="fused")
    ap.add_argument("--cells", type=int, default=10080)
    ap.add_argument("--block", type=int, default=126)
    ap.add_argument("--kernel", choices=["auto", "percell", "stream", "pair"], default="auto")
    ap.add_argument("--strips", type=int, default=0, help="row strips per block (0 = default)")
    ap.add_argument("--prefetch", type=int, default=0, choices=[0, 1, 2],
                    help="rows the streaming kernel loads ahead (0 = default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the reduced batched-AMR, ensemble and adjoint legs of the default run")
    ap.add_argument("--integrator", choices=["euler", "rk4"], default="euler",
                    help="rk4: RungeKutta4 (deterministic runs), four stages of whole-slab sweeps")
    ap.add_argument("--oriented", action="store_true",
                    help="multi-grain Model C (per-grain θ₀, the per-cell oriented kernel)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every launch shape of the Model C step first and use the fastest")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this backend "
              "runs on a single MI355X (multi-GPU sharding is a design note, not a build)",
              file=sys.stderr)
        return 2

    if args.workload in ("adjoint", "ensemble", "amr"):
        fn = {"adjoint": run_adjoint, "ensemble": run_ensemble, "amr": run_amr}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return 0
    if args.workload == "jump":
        return run_jump(args)
    if args.workload == "observe":
        return run_observe(args)
    if args.workload == "heat3d":
        return run_heat3d(args)

    from legendre_amd.workloads import ModelCConfig

    C = ModelCConfig(cells=args.cells, block=args.block, noise=args.noise)
    total_ms, per, wall, ceiling = run_gpu(args, C)
    ms_per_step = total_ms / args.steps
    value = C.interior_cells * args.steps / (total_ms / 1e3) / 1e9
    avg_launch_ms = sum(per) / len(per)
    # Field passes per step of the fused kernels (8 B per cell each): the Euler step
    # reads φ, u and writes them (4); the multi-grain cell also reads cos θ₀ and
    # sin θ₀ (6); the four RK4 stages move 16 passes per field (32), 40 oriented.
    # The per-intent mode runs many kernels per step: no single-kernel figure.
    passes = None
    if args.mode == "fused":
        if args.integrator == "rk4":
            passes = 40 if args.oriented else 32
        else:
            passes = 6 if args.oriented else 4
    bytes_per_launch = passes * 8 * C.interior_cells if passes else None
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9 if passes else None
    log(f"[bench] {args.steps} steps: {total_ms:.3f} ms on-stream, {wall * 1e3:.3f} ms wall; "
        f"per launch min {min(per):.4f} max {max(per):.4f} ms")
    traffic = read_traffic(os.path.join(ROOT, "profiles", "modelc_stream_kernel_10080sq.md"))
    if (args.cells != 10080 or args.block != 126 or args.noise != 0.0 or args.kernel == "percell"
            or args.mode != "fused" or args.oriented or args.integrator != "euler"):
        traffic = None  # the committed profile covers the default workload only
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(C, args.cpu_budget)
        log(f"[bench] cpu baseline: {cpu}")
    workload = f"model_c_{C.cells}sq_{C.block}blocks_b{C.noise:g}_{args.mode}"
    if args.oriented:
        workload += "_oriented"
    if args.integrator != "euler":
        workload += f"_{args.integrator}"
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
            "integrator": "euler_maruyama" if args.integrator == "euler" else args.integrator,
            "parallelism": "single_gpu",
        },
        "roofline": None if passes is None else {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": ("modelc_step_kernel" if args.kernel == "percell" or args.oriented
                       else "modelc_stream_kernel")
                      + (" (four RK4 stage launches per step)" if args.integrator == "rk4" else ""),
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_launch_ms, 5),
            "ceiling": ceiling,
            "frac_of_ceiling": round(achieved / ceiling["gbs"], 4),
        },
        "cpu_baseline": cpu,
    }
    default_variant = (args.cells == 10080 and args.block == 126 and args.noise == 0.0 and args.mode == "fused"
                       and args.integrator == "euler" and not args.oriented and args.kernel == "auto"
                       and not args.strips and not args.prefetch and not args.graph and not args.autotune)
    if default_variant and not args.no_extras:
        line["batched_and_adjoint"] = run_extras(args, value)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
