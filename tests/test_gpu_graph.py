"""Whole-step HIP graphs (lgd_sim_set_graph_steps): a simulation replaying captured
steps computes bitwise what eager launches compute — noise included (its salt comes
from a device copy of t advanced inside the graph), across ping-pong parities, odd
chunk lengths, AMR subcycling, regrids between graphs, and dt and tuning changes.
Each case is checked after every replay, with eager work (downloads, other
simulations) between the replays."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def fields(sim):
    return [sim.field(f).download() for f in range(sim.nfields)]


def assert_same(a, b, where):
    fa, fb = fields(a), fields(b)
    for f, (x, y) in enumerate(zip(fa, fb)):
        assert np.array_equal(x, y), f"{where}: field {f} differs"
    assert a.time == b.time, where
    assert a.step_index == b.step_index, where


def grains_state(orc, cells, block, grains):
    g = orc.grid(cells, block, [0, 0], [0.4, 0.4])
    n = orc.field_len(g, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(g, phi, u, np.array(grains), 0.55)
    return phi, u


def run_pair(make, chunks, graph_steps, dts=None):
    eager, graph = make(), make()
    graph.set_graph_steps(graph_steps)
    for i, n in enumerate(chunks):
        dt = dts[i] if dts else None
        eager.step(dt, n)
        graph.step(dt, n)
        assert_same(eager, graph, f"after chunk {i} ({n} steps)")
    eager.close()
    graph.close()


@pytest.mark.parametrize("graph_steps", [1, 3, 4])
@pytest.mark.parametrize("mode", ["fused", "per_intent"])
def test_modelc_with_noise_graph_equals_eager(ctx, orc, mode, graph_steps):
    say(f"graph modelc mode={mode} graph_steps={graph_steps}")
    cells, block = [64, 48], [32, 24]
    phi, u = grains_state(orc, cells, block, [[9.0, 8.0, 2.4], [17.0, 12.0, 1.6]])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])
    m = lg.LGD_MODE_FUSED if mode == "fused" else lg.LGD_MODE_PER_INTENT

    def make():
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.02), wiener=True,
                                    seed=7, mode=m)
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        return sim

    dt = 0.008
    run_pair(make, [5, 1, 7, 4], graph_steps, dts=[dt] * 4)
    grid.close()


def test_heat_graph_equals_eager_with_periodic_faces(ctx):
    say("graph heat 3-D periodic")
    grid = lg.CartesianGrid(ctx, [24, 16, 8], [8, 8, 4], [0.0, 0.0, 0.0], [0.1, 0.1, 0.1],
                            [True, False, True])
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(grid.field_len(1))

    def make():
        sim = lg.Simulation.heat(grid, 1.0)
        sim.field(0).upload(u0)
        return sim

    run_pair(make, [3, 3, 2], 3, dts=[1e-3] * 3)
    grid.close()


def test_graph_follows_dt_and_tuning_changes(ctx, orc):
    """A change of dt or of a launch knob re-captures; results stay eager-equal."""
    say("graph dt and tuning changes")
    cells, block = [64, 48], [32, 24]
    phi, u = grains_state(orc, cells, block, [[9.0, 8.0, 2.4]])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])

    def make():
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.02), wiener=True, seed=3)
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        return sim

    eager, graph = make(), make()
    graph.set_graph_steps(2)
    try:
        for i, (dt, strips) in enumerate([(0.008, 0), (0.004, 0), (0.004, 3), (0.008, 1)]):
            ctx.set_tuning("modelc_strips", strips)
            eager.step(dt, 4)
            graph.step(dt, 4)
            assert_same(eager, graph, f"phase {i}")
    finally:
        ctx.set_tuning("modelc_strips", 0)
    # turning graphs off mid-run continues the same trajectory
    graph.set_graph_steps(0)
    eager.step(0.008, 3)
    graph.step(0.008, 3)
    assert_same(eager, graph, "graphs off")
    eager.close()
    graph.close()
    grid.close()


@pytest.mark.parametrize("subcycle", [False, True])
@pytest.mark.parametrize("graph_steps", [1, 3])
def test_adaptive_amr_modelc_with_noise_graph_equals_eager(ctx, orc, subcycle, graph_steps):
    """Regrids every 4 steps split the graphs; each new hierarchy re-captures."""
    say(f"graph amr subcycle={subcycle} graph_steps={graph_steps}")
    base = orc.grid([48, 48], [24, 24], [0, 0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0], [12.0, 9.0, 2.0]]), 0.7)
    pol = lg.RegridPolicy(4, 2, 0.8, 4, 0.15, 0)

    def make():
        gg = lg.AmrGrid.from_patches(ctx, [48, 48], [24, 24], [0, 0], [0.4, 0.4], [2, 2], [])
        sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 2.0, 0.01),
                                wiener=True, seed=5, subcycle=subcycle, policy=pol)
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        return sim

    dt = 0.2 * (0.4 if subcycle else 0.1) ** 2 / 2.0
    eager, graph = make(), make()
    graph.set_graph_steps(graph_steps)
    for i, n in enumerate([3, 5, 1, 7, 4]):
        eager.step(dt, n)
        graph.step(dt, n)
        assert eager.amr.num_blocks == graph.amr.num_blocks, f"chunk {i}"
        assert_same(eager, graph, f"after chunk {i}")
    assert graph.amr.num_levels >= 2
    eager.close()
    graph.close()


def test_static_amr_heat_graph_equals_eager(ctx):
    say("graph static amr heat")
    levels = [[((4, 4), (20, 24)), ((20, 8), (28, 20))], [((12, 12), (36, 40)), ((40, 20), (52, 36))]]
    rng = np.random.default_rng(9)

    def make(u0=[]):
        gg = lg.AmrGrid.from_patches(ctx, [16, 16], [8, 8], [0, 0], [0.5, 0.5], [2, 2], levels)
        sim = lg.amr_simulation(gg, "heat", kappa=0.7, subcycle=True)
        if not u0:
            u0.append(rng.standard_normal(sim.amr.field_len))
        sim.field(0).upload(u0[0])
        return sim

    run_pair(make, [2, 2, 5], 2, dts=[0.2 * 0.25 * 0.5 / 0.7] * 3)
