"""The second oracle: the GPU backend's own run-to-run determinism. Every path run
twice from the same inputs gives bit-identical results — the fused Model C step with
noise (and under graph replay), the ensemble with noise, the subcycled adaptive run
with noise and regrids, the calibration gradient, the jump driver, RK4 on
orientations, and the Parquet files the observer writes (byte for byte)."""
from __future__ import annotations

import filecmp
import os
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def twice(fn):
    a, b = fn(), fn()
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))
    return a


def test_fused_model_c_with_noise_and_graphs(ctx):
    say("determinism: fused Model C with noise, eager and graph replay")
    grid = lg.CartesianGrid(ctx, [252, 126], [126, 126], [0.0, 0.0], [0.4, 0.4])

    def run(graph=0):
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.01), wiener=True, seed=42)
        sim.initialize_grains(np.array([[20.0, 25.0, 2.0], [70.0, 20.0, 2.0]]), 0.55)
        sim.set_graph_steps(graph)
        sim.step(0.008, 37)
        out = (sim.field(0).download(), sim.field(1).download())
        sim.close()
        return out

    eager = twice(run)
    replay = twice(lambda: run(5))
    assert np.array_equal(eager[0], replay[0]) and np.array_equal(eager[1], replay[1])
    grid.close()


def test_ensemble_with_noise(ctx):
    say("determinism: ensemble with noise")
    sides = [(64, 64), (128, 64), (64, 192)]

    def run():
        reals = [lg.Realization(s, 0.04, 3.19, 2.0, 0.015, seed=7 + i) for i, s in enumerate(sides)]
        ens = lg.Ensemble(ctx, reals, block_cells=(64, 64), spacing=(0.4, 0.4))
        for i, s in enumerate(sides):
            ens.initialize_grains(i, np.array([[s[0] * 0.2, s[1] * 0.2, 2.0]]), 0.55)
        ens.step(0.008, 40)
        out = [ens.realization_field(i, f) for i in range(len(sides)) for f in (0, 1)]
        out.append(ens.solid_fraction())
        ens.close()
        return out

    twice(run)


def test_adaptive_run_with_noise_and_regrids(ctx, orc):
    say("determinism: subcycled AMR with noise and regrids")
    base = orc.grid([96, 96], [24, 24], [0, 0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0], [20.0, 15.0, 3.0]]), 0.7)

    def run():
        gg = lg.AmrGrid.from_patches(ctx, [96, 96], [24, 24], [0, 0], [0.4, 0.4], [2, 2], [])
        sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 2.0, 0.01), wiener=True,
                                seed=3, subcycle=True, policy=lg.RegridPolicy(3, 2, 0.75, 4, 0.12, 0))
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        sim.step(0.2 * 0.4 ** 2 / 2.0, 10)
        a = sim.amr
        m = H.amr_interior_mask(a)
        patches = np.array([[lvl, *lo, *hi] for lvl, lo, hi in (a.patch(b) for b in range(a.num_blocks))],
                           dtype=np.float64)
        out = (patches.reshape(-1), sim.field(0).download()[m], sim.field(1).download()[m])
        sim.close()
        return out

    twice(run)


def test_calibration_gradient(ctx, orc):
    say("determinism: calibration gradient")
    cells, block = [128, 128], [64, 64]
    og = orc.grid(cells, block, [0.0, 0.0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(og, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(og, phi, u, np.array([[25.6, 25.6, 3.0]]), 0.55)
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])

    def buf(a):
        b = grid.alloc_field()
        b.upload(a)
        return b

    target = phi.copy()
    target[::7] = 0.3

    def run():
        cal = lg.ModelCCalibration(grid, 40, 0.008, snapshots=6)
        gp, gu = grid.alloc_field(), grid.alloc_field()
        loss = cal.forward(lg.ModelCParams(0.05, 3.19, 2.0), buf(phi), buf(u), buf(target))
        grad = cal.backward(gp, gu)
        out = (np.array([loss]), np.asarray(grad), gp.download(), gu.download())
        cal.close()
        return out

    twice(run)
    grid.close()


def test_jump_driver_and_oriented_rk4(ctx):
    say("determinism: jump driver, oriented RK4")
    grid = lg.CartesianGrid(ctx, [4096], [512], [0.0], [1.0])
    n = grid.field_len(0)
    lam = np.linspace(0.0, 8.0, n)

    def jump():
        xs = [lg.DeviceBuffer(ctx, n) for _ in range(3)]
        bufs = []
        for v in (0.1, 0.5, 1.0, 2.0):
            b = lg.DeviceBuffer(ctx, n)
            b.upload(np.full(n, v))
            bufs.append(b)
        rate = lg.DeviceBuffer(ctx, n)
        rate.upload(lam)
        drivers = [lg.Driver(lg.LGD_DRIVER_TIME, 0, [bufs[0], None, None]),
                   lg.Driver(lg.LGD_DRIVER_WIENER, 0, [bufs[1], None, None]),
                   lg.Driver(lg.LGD_DRIVER_JUMP, 0, [None, bufs[2], bufs[3]], rate)]
        t = 0.0
        for _ in range(50):
            grid.apply_drivers(xs, [0, 0, 0], drivers, 0.01, 9, lg.time_bits(t))
            t += 0.01
        return [x.download() for x in xs]

    twice(jump)
    g2 = lg.CartesianGrid(ctx, [96, 64], [32, 32], [0.0, 0.0], [0.4, 0.4])

    def rk4():
        sim = lg.Simulation.model_c(g2, lg.ModelCParams(0.05, 3.19, 2.0))
        sim.set_integrator("rk4")
        sim.initialize_grains_oriented(np.array([[10.0, 10.0, 2.0, 0.4], [25.0, 15.0, 2.0, -0.3]]), 0.6)
        sim.step(0.008, 12)
        out = (sim.field(0).download(), sim.field(1).download())
        sim.close()
        return out

    twice(rk4)
    g2.close()
    grid.close()


def test_observer_files_are_byte_identical(ctx, tmp_path):
    say("determinism: observer files")
    grid = lg.CartesianGrid(ctx, [128, 96], [32, 32], [0.0, 0.0], [0.4, 0.4])
    dirs = []
    for k in range(2):
        d = tmp_path / f"run{k}"
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.01), wiener=True, seed=5)
        sim.initialize_grains(np.array([[20.0, 20.0, 3.0]]), 0.55)
        obs = lg.ParquetObserver(ctx, d, [("phi", 0), ("u", 1)], every=4, blocking=True)
        sim.attach_observer(obs)
        sim.step(0.008, 9)
        obs.flush()
        sim.close()
        obs.close()
        dirs.append(d)
    names = sorted(os.listdir(dirs[0]))
    assert names == sorted(os.listdir(dirs[1])) and len(names) == 4
    match, mismatch, errors = filecmp.cmpfiles(dirs[0], dirs[1], names, shallow=False)
    assert not mismatch and not errors, (mismatch, errors)
    grid.close()
