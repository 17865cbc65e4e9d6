"""GPU tests of the observation seam (src/core/observer.rs, src/io/parquet.rs): the
device gather + pinned snapshot ring + Parquet writer, read back with pyarrow.

- The reference's round-trip test (tests/parquet.rs:88-140) through the C ABI.
- A Model C simulation with an attached observer: snapshots at step 1 and every
  `every` steps hold the simulation's interiors bit for bit (eager and graph-
  replayed runs), the static file holds the oracle's cell centres.
- Multi-grain Model C: the static file carries θ₀.
- An adaptive run: one static file per grid epoch, every snapshot joined to its
  epoch's static file row for row, cell centres from the oracle's AmrGrid.
- Back-pressure: with the writer held, snapshots are dropped and counted, the solver
  never waits, and every queued snapshot reaches disk on release.
- Blocking mode (the synchronous ParquetObserver) writes every snapshot step."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")
pq = pytest.importorskip("pyarrow.parquet")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rows(field, cells, block, ghost):
    """A slab field in the writer's row order: blocks in order, interior cells
    dimension 0 fastest."""
    fv = H.field_view(field, cells, block, ghost)
    return fv[(slice(None),) + H.interior_slices(len(cells), ghost)].reshape(-1)


def table(path):
    return pq.ParquetFile(path).read()


def test_reference_round_trip(ctx, orc, tmp_path):
    """tests/parquet.rs:88-140: 4×4 grid, 2×2 blocks, phi dynamic (ghost 1), theta0
    static (ghost 0), every = 2; steps 1, 3, 4 observed."""
    say("parquet round trip (reference test)")
    cells, block = [4, 4], [2, 2]
    og = orc.grid(cells, block, [0.0, 0.0], [0.5, 0.5])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.5, 0.5])
    phi = np.zeros(orc.field_len(og, 1))
    theta = np.zeros(orc.field_len(og, 0))
    xs, ys = [], []
    pv, tv = H.field_view(phi, cells, block, 1), H.field_view(theta, cells, block, 0)
    for b in range(4):
        for j in range(2):
            for i in range(2):
                x, y = orc.cell_center(og, b, (i, j))[:2]
                xs.append(x)
                ys.append(y)
                pv[b, j + 1, i + 1] = np.float64(100.0) * x + y
                tv[b, j, i] = -(np.float64(1000.0) * y + x)
    bp, bt = lg.DeviceBuffer(ctx, phi.size), lg.DeviceBuffer(ctx, theta.size)
    bp.upload(phi)
    bt.upload(theta)
    obs = lg.ParquetObserver(ctx, tmp_path, [("phi", 0)], every=2, statics=[("theta0", 0)],
                             blocking=True)
    for step, t in ((1, 0.25), (3, 0.75), (4, 1.0)):
        obs.observe(grid, step, t, 0, [bp], [1], [bt], [0])
    obs.flush()
    st = obs.stats()
    assert st["written"] == 3 and st["dropped"] == 0 and st["failed"] == 0, st
    assert sorted(os.listdir(tmp_path)) == ["snap_0000001.parquet", "snap_0000004.parquet",
                                            "static_0000000.parquet"]
    s = table(tmp_path / "static_0000000.parquet")
    assert s.column_names == ["x", "y", "level", "patch", "theta0"]
    assert np.array_equal(s["x"].to_numpy(), xs) and np.array_equal(s["y"].to_numpy(), ys)
    assert np.array_equal(s["theta0"].to_numpy(), rows(theta, cells, block, 0))
    assert np.array_equal(s["level"].to_numpy(), np.zeros(16))
    assert np.array_equal(s["patch"].to_numpy(), np.repeat(np.arange(4.0), 4))
    for step, t in ((1, 0.25), (4, 1.0)):
        sn = table(tmp_path / f"snap_{step:07d}.parquet")
        assert sn.column_names == ["step", "t", "epoch", "phi"]
        assert np.array_equal(sn["phi"].to_numpy(), rows(phi, cells, block, 1))
        assert list(sn["step"].to_numpy()) == [step] * 16 and list(sn["epoch"].to_numpy()) == [0] * 16
        assert list(sn["t"].to_numpy()) == [t] * 16
    with pytest.raises(lg.LegendreError):
        lg.ParquetObserver(ctx, tmp_path / "x", [], every=1)
    obs.close()
    grid.close()


@pytest.mark.parametrize("graph", [0, 3])
def test_simulation_snapshots_hold_the_state(ctx, orc, tmp_path, graph):
    cells, block = [96, 64], [32, 32]
    say(f"model C observer graph={graph}")
    og = orc.grid(cells, block, [0.0, 0.0], [0.4, 0.4])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])
    grains = np.array([[10.0, 12.0, 3.0], [28.0, 14.0, 2.0]])
    p = lg.ModelCParams(0.05, 3.19, 2.0, 0.01)
    sim = lg.Simulation.model_c(grid, p, wiener=True, seed=7)
    sim.initialize_grains(grains, 0.55)
    sim.set_graph_steps(graph)
    obs = lg.ParquetObserver(ctx, tmp_path, [("phi", 0), ("u", 1)], every=5, ring=3, blocking=True)
    sim.attach_observer(obs)
    ref = lg.Simulation.model_c(grid, p, wiener=True, seed=7)
    ref.initialize_grains(grains, 0.55)
    expect = {}
    done = 0
    for stop in (1, 5, 10):
        ref.step(0.008, stop - done)
        done = stop
        expect[stop] = (ref.field(0).download(), ref.field(1).download(), ref.time)
    sim.step(0.008, 12)
    obs.flush()
    assert sorted(os.listdir(tmp_path)) == ["snap_0000001.parquet", "snap_0000005.parquet",
                                            "snap_0000010.parquet", "static_0000000.parquet"]
    for step, (phi, u, t) in expect.items():
        sn = table(tmp_path / f"snap_{step:07d}.parquet")
        assert np.array_equal(sn["phi"].to_numpy(), rows(phi, cells, block, 1)), step
        assert np.array_equal(sn["u"].to_numpy(), rows(u, cells, block, 1)), step
        assert sn["t"].to_numpy()[0] == t
    s = table(tmp_path / "static_0000000.parquet")
    n = 0
    for b in range(orc.num_blocks(og)):
        for j in range(block[1]):
            for i in range(block[0]):
                x, y = orc.cell_center(og, b, (i, j))[:2]
                assert s["x"][n].as_py() == x and s["y"][n].as_py() == y
                n += 1
    assert obs.stats()["dropped"] == 0
    sim.close()
    ref.close()
    obs.close()
    grid.close()


def test_oriented_static_file_carries_theta0(ctx, orc, tmp_path):
    cells, block = [48, 40], [24, 20]
    say("oriented static theta0")
    og = orc.grid(cells, block, [0.0, 0.0], [0.4, 0.4])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])
    grains = np.array([[5.0, 4.0, 2.0, 0.3], [14.0, 11.0, 1.5, -0.7]])
    sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0))
    sim.initialize_grains_oriented(grains, 0.6)
    obs = lg.ParquetObserver(ctx, tmp_path, [("phi", 0)], every=2, statics=[("theta0", 0)],
                             blocking=True)
    sim.attach_observer(obs)
    sim.step(0.008, 2)
    obs.flush()
    theta = np.zeros(orc.field_len(og, 0))
    orc.initialize_orientations(og, theta, grains)
    s = table(tmp_path / "static_0000000.parquet")
    assert np.array_equal(s["theta0"].to_numpy(), rows(theta, cells, block, 0))
    assert sorted(os.listdir(tmp_path)) == ["snap_0000001.parquet", "snap_0000002.parquet",
                                            "static_0000000.parquet"]
    sim.close()
    obs.close()
    grid.close()


def test_adaptive_run_writes_one_static_file_per_epoch(ctx, orc, tmp_path):
    say("amr observer epochs")
    base = orc.grid([96, 96], [24, 24], [0, 0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0], [20.0, 15.0, 3.0]]), 0.7)
    pol = lg.RegridPolicy(3, 2, 0.75, 4, 0.12, 0)
    gg = lg.AmrGrid.from_patches(ctx, [96, 96], [24, 24], [0, 0], [0.4, 0.4], [2, 2], [])
    sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 2.0), subcycle=True,
                            policy=pol)
    sim.field(0).upload(phi)
    sim.field(1).upload(u)
    obs = lg.ParquetObserver(ctx, tmp_path, [("phi", 0), ("u", 1)], every=2, ring=4, blocking=True)
    sim.attach_observer(obs)
    snaps = {}
    for s in range(1, 14):
        sim.step(0.2 * 0.4 ** 2 / 2.0)
        if s == 1 or s % 2 == 0:
            a = sim.amr
            m = H.amr_interior_mask(a)
            patches = [a.patch(b) for b in range(a.num_blocks)]
            snaps[s] = (patches, sim.field(0).download()[m], sim.field(1).download()[m])
    obs.flush()
    st = obs.stats()
    assert st["dropped"] == 0 and st["failed"] == 0, st
    statics = sorted(f for f in os.listdir(tmp_path) if f.startswith("static_"))
    assert len(statics) >= 2  # the hierarchy changed at least once
    for s, (patches, pv, uv) in snaps.items():
        sn = table(tmp_path / f"snap_{s:07d}.parquet")
        epoch = int(sn["epoch"][0].as_py())
        stf = table(tmp_path / f"static_{epoch:07d}.parquet")
        assert stf.num_rows == sn.num_rows == pv.size
        # Interiors in patch order, each patch row-major with x fastest: the mask order.
        assert np.array_equal(sn["phi"].to_numpy(), pv) and np.array_equal(sn["u"].to_numpy(), uv)
        levels = [[] for _ in range(2)]
        for (lvl, lo, hi) in patches:
            if lvl > 0:
                levels[lvl - 1].append((lo, hi))
        oa = orc.AmrGrid.from_patches(base, [2, 2], [lv for lv in levels if lv])
        n = 0
        for b, (lvl, lo, hi) in enumerate(patches):
            for j in range(hi[1] - lo[1]):
                for i in range(hi[0] - lo[0]):
                    if (i + j) % 7 == 0:  # a sample of every patch
                        x, y = oa.cell_center(b, i, j)
                        assert stf["x"][n].as_py() == x and stf["y"][n].as_py() == y
                        assert stf["level"][n].as_py() == lvl and stf["patch"][n].as_py() == b
                    n += 1
    sim.close()
    obs.close()


def test_a_held_writer_drops_snapshots_without_stalling_the_solver(ctx, orc, tmp_path):
    say("observer back-pressure")
    cells, block = [64, 64], [32, 32]
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])
    sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0))
    sim.initialize_grains(np.array([[12.0, 12.0, 3.0]]), 0.55)
    obs = lg.ParquetObserver(ctx, tmp_path, [("phi", 0)], every=1, ring=3)
    sim.attach_observer(obs)
    obs.pause(True)
    sim.step(0.008, 6)  # steps 1-3 take the three ring buffers (step 1 also the static one)
    ctx.sync()
    st = obs.stats()
    assert st["written"] == 0 and st["dropped"] == 3, st
    obs.flush()  # releases the writer
    st = obs.stats()
    assert st["written"] == 4 and st["dropped"] == 3, st
    assert sorted(os.listdir(tmp_path)) == ["snap_0000001.parquet", "snap_0000002.parquet",
                                            "snap_0000003.parquet", "static_0000000.parquet"]
    sim.step(0.008, 1)  # buffers are back: step 7 snapshots again
    obs.flush()
    assert obs.stats()["written"] == 5 and (tmp_path / "snap_0000007.parquet").exists()
    obs.close()  # detaches itself: the simulation steps on without it
    sim.step(0.008, 2)
    assert not (tmp_path / "snap_0000008.parquet").exists()
    sim.close()
    obs.close()
    grid.close()


def test_three_dimensional_heat_has_z_coordinates(ctx, orc, tmp_path):
    say("heat 3-D observer")
    cells, block = [16, 12, 8], [8, 6, 4]
    og = orc.grid(cells, block, [0.5, -1.0, 2.0], [0.1, 0.2, 0.3])
    grid = lg.CartesianGrid(ctx, cells, block, [0.5, -1.0, 2.0], [0.1, 0.2, 0.3])
    sim = lg.Simulation.heat(grid, 1.0)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(orc.field_len(og, 1))
    sim.field(0).upload(u0)
    obs = lg.ParquetObserver(ctx, tmp_path, [("u", 0)], every=1, blocking=True)
    sim.attach_observer(obs)
    sim.step(1e-4, 1)
    obs.flush()
    s = table(tmp_path / "static_0000000.parquet")
    assert s.column_names == ["x", "y", "z", "level", "patch"]
    n = 0
    for b in range(orc.num_blocks(og)):
        for k in range(block[2]):
            for j in range(block[1]):
                for i in range(block[0]):
                    if n % 5 == 0:
                        want = orc.cell_center(og, b, (i, j, k))
                        assert (s["x"][n].as_py(), s["y"][n].as_py(), s["z"][n].as_py()) == tuple(want[:3])
                    n += 1
    sn = table(tmp_path / "snap_0000001.parquet")
    assert np.array_equal(sn["u"].to_numpy(), rows(sim.field(0).download(), cells, block, 1))
    sim.close()
    obs.close()
    grid.close()
