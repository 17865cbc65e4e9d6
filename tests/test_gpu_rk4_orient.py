"""GPU parity of RungeKutta4 (src/integrators/rk4.rs) and of multi-grain Model C
(ModelC::with_orientations: finite_volume.rs:188-235, phasefield.rs:227-314,
378-392) against the oracle, bitwise: heat and Model C under RK4, the oriented
vector field on its own, oriented forward-Euler steps (fused and per-intent) and
oriented RK4 steps; plus the driver-contract error and graph replay of RK4."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def interior_mask(cells, block, ghost=1):
    nb = 1
    for c, b in zip(cells, block):
        nb *= c // b
    shape = [b + 2 * ghost for b in reversed(block)]
    m = np.zeros([nb] + shape, dtype=bool)
    sl = tuple([slice(None)] + [slice(ghost, -ghost)] * len(block))
    m[sl] = True
    return m.reshape(-1)


GRAINS = np.array([[6.0, 5.5, 2.4, 0.35], [16.0, 12.0, 1.8, -0.6], [9.0, 14.0, 1.5, 1.1]])
CELLS, BLOCK = [64, 48], [32, 24]


def oriented_state(orc, cells=CELLS, block=BLOCK, grains=GRAINS):
    og = orc.grid(cells, block, [0, 0], [0.4, 0.4])
    n = orc.field_len(og, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(og, phi, u, grains[:, :3], 0.6)
    theta = np.zeros(orc.field_len(og, 0))
    orc.initialize_orientations(og, theta, grains)
    return og, phi, u, theta


@pytest.mark.parametrize("case", ["3d_periodic", "2d_rule_faces"])
def test_heat_rk4_matches_oracle_bitwise(ctx, orc, case):
    say(f"heat rk4 {case}")
    if case == "3d_periodic":
        cells, block, periodic, bcs = [16, 16, 8], [8, 8, 4], [True, False, True], None
    else:
        cells, block, periodic = [32, 24], [16, 12], [False, False]
        bcs = [(lg.LGD_BC_DIRICHLET, 0.3), (lg.LGD_BC_FLUX, 0.1), (lg.LGD_BC_EXTRAPOLATE, 0.0),
               (lg.LGD_BC_MIRROR, 0.0)]
    h = [0.1] * len(cells)
    og = orc.grid(cells, block, [0] * len(cells), h, periodic)
    rng = np.random.default_rng(11)
    u0 = rng.standard_normal(orc.field_len(og, 1))
    want = u0.copy()
    dt = 0.2 * 0.01 / (2 * len(cells))
    orc.heat_rk4_run(og, 1.0, want, 7, dt, bcs=bcs)
    grid = lg.CartesianGrid(ctx, cells, block, [0.0] * len(cells), h, periodic)
    sim = lg.Simulation.heat(grid, 1.0, bcs=None if bcs is None else [lg.FaceBc(k, v) for k, v in bcs],
                             mode=lg.LGD_MODE_PER_INTENT)
    sim.set_integrator("rk4")
    sim.field(0).upload(u0)
    sim.step(dt, 7)
    m = interior_mask(cells, block)
    got = sim.field(0).download()
    assert np.array_equal(got[m], want[m])
    sim.close()
    grid.close()


def test_modelc_rk4_matches_oracle_bitwise(ctx, orc):
    say("modelc rk4")
    og, phi, u, _ = oriented_state(orc)
    p = orc.params(0.05, 3.19, 2.0)
    wp, wu = phi.copy(), u.copy()
    orc.modelc_rk4_run(og, p, wp, wu, 12, 0.008)
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    for mode in (lg.LGD_MODE_FUSED, lg.LGD_MODE_PER_INTENT):
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0), mode=mode)
        sim.set_integrator("rk4")
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        sim.step(0.008, 12)
        m = interior_mask(CELLS, BLOCK)
        assert np.array_equal(sim.field(0).download()[m], wp[m])
        assert np.array_equal(sim.field(1).download()[m], wu[m])
        sim.close()
    grid.close()


def test_oriented_vector_field_matches_oracle_bitwise(ctx, orc):
    say("oriented rhs")
    og, phi, u, theta = oriented_state(orc)
    orc.fill_ghosts(og, phi, 1)
    orc.fill_ghosts(og, u, 1)
    p = orc.params(0.05, 3.19, 2.0)
    kp, ku = np.zeros(phi.size), np.zeros(phi.size)
    orc.modelc_rhs_oriented(og, p, phi, u, theta, kp, ku)
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    bufs = [lg.DeviceBuffer(ctx, x.size) for x in (phi, u, theta, theta, kp, ku)]
    bufs[0].upload(phi)
    bufs[1].upload(u)
    bufs[2].upload(np.cos(theta))
    bufs[3].upload(np.sin(theta))
    grid.modelc_rhs_oriented(lg.ModelCParams(0.05, 3.19, 2.0), *bufs)
    m = interior_mask(CELLS, BLOCK)
    assert np.array_equal(bufs[4].download()[m], kp[m])
    assert np.array_equal(bufs[5].download()[m], ku[m])
    grid.close()


@pytest.mark.parametrize("integrator", ["euler", "rk4"])
@pytest.mark.parametrize("mode", ["fused", "per_intent"])
def test_oriented_steps_match_oracle_bitwise(ctx, orc, mode, integrator):
    say(f"oriented steps mode={mode} integrator={integrator}")
    og, phi, u, theta = oriented_state(orc)
    p = orc.params(0.05, 3.19, 2.0)
    wp, wu = phi.copy(), u.copy()
    if integrator == "rk4":
        orc.modelc_rk4_run(og, p, wp, wu, 10, 0.008, theta=theta)
    else:
        orc.modelc_oriented_run(og, p, wp, wu, theta, 25, 0.008)
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    m_ = lg.LGD_MODE_FUSED if mode == "fused" else lg.LGD_MODE_PER_INTENT
    sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0), mode=m_)
    sim.set_integrator(integrator)
    sim.initialize_grains_oriented(GRAINS, 0.6)
    m = interior_mask(CELLS, BLOCK)
    assert np.array_equal(sim.field(0).download()[m], phi[m])
    sim.step(0.008, 10 if integrator == "rk4" else 25)
    assert np.array_equal(sim.field(0).download()[m], wp[m])
    assert np.array_equal(sim.field(1).download()[m], wu[m])
    sim.close()
    grid.close()


def test_oriented_fused_and_per_intent_agree_with_noise(ctx, orc):
    """With a Wiener driver the two execution modes draw the same deviates and add
    them in the same order: bitwise equal."""
    say("oriented noise fused vs per-intent")
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    out = []
    for mode in (lg.LGD_MODE_FUSED, lg.LGD_MODE_PER_INTENT):
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.02), wiener=True, seed=5,
                                    mode=mode)
        sim.initialize_grains_oriented(GRAINS, 0.6)
        sim.step(0.008, 15)
        m = interior_mask(CELLS, BLOCK)
        out.append((sim.field(0).download()[m], sim.field(1).download()[m]))
        sim.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    grid.close()


def test_rk4_rejects_a_stochastic_model_and_replays_as_a_graph(ctx, orc):
    say("rk4 contract and graph")
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    noisy = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.02), wiener=True)
    with pytest.raises(lg.LegendreError) as e:
        noisy.set_integrator("rk4")
    assert e.value.status == 5
    noisy.close()
    sims = []
    for graph in (0, 3):
        sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0))
        sim.set_integrator("rk4")
        sim.initialize_grains_oriented(GRAINS, 0.6)
        sim.set_graph_steps(graph)
        sims.append(sim)
    for n in (4, 3, 5):
        for sim in sims:
            sim.step(0.008, n)
        a, b = (s.field(0).download() for s in sims)
        assert np.array_equal(a, b)
    for sim in sims:
        sim.close()
    grid.close()


@pytest.mark.parametrize("integrator", ["euler", "rk4"])
def test_zero_orientation_is_bitwise_the_axis_aligned_path(ctx, orc, integrator):
    """tests/model_c.rs (zero_orientation_is_bitwise_identical_to_legacy_path): θ₀ = 0
    everywhere through the oriented kernels equals the axis-aligned model bit for bit."""
    say(f"zero orientation == axis-aligned, {integrator}")
    grid = lg.CartesianGrid(ctx, CELLS, BLOCK, [0.0, 0.0], [0.4, 0.4])
    out = []
    for oriented in (False, True):
        for mode in (lg.LGD_MODE_FUSED, lg.LGD_MODE_PER_INTENT):
            sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0), mode=mode)
            sim.set_integrator(integrator)
            sim.initialize_grains(GRAINS[:, :3], 0.6)
            if oriented:
                sim.set_orientation(np.zeros(orc.num_blocks(orc.grid(CELLS, BLOCK, [0, 0], [0.4, 0.4]))
                                             * BLOCK[0] * BLOCK[1]))
            sim.step(0.008, 12)
            m = interior_mask(CELLS, BLOCK)
            out.append((sim.field(0).download()[m], sim.field(1).download()[m]))
            sim.close()
    for phi, u in out[1:]:
        assert np.array_equal(phi, out[0][0]) and np.array_equal(u, out[0][1])
    grid.close()


RK4_STREAM_CASES = [
    # cells, block, periodic, (strips, rows ahead)
    ([64, 48], [32, 24], [False, False], (4, 2)),
    ([90, 60], [30, 20], [True, False], (3, 1)),
    ([252, 126], [126, 63], [False, True], (4, 2)),
    ([12, 8], [1, 2], [True, True], (1, 2)),
    ([256, 40], [128, 20], [False, False], (2, 2)),  # 128 wide: one thread per cell
]


@pytest.mark.parametrize("case", range(len(RK4_STREAM_CASES)))
def test_modelc_rk4_stream_stages_match_per_cell_and_oracle(ctx, orc, case):
    """The fused RK4 stages on the streaming kernel (face sharing, two columns per
    lane) against the one-thread-per-cell stages, whole slabs bit for bit, and
    against the oracle's RungeKutta4 on the interiors."""
    cells, block, periodic, (strips, prefetch) = RK4_STREAM_CASES[case]
    say(f"rk4 streaming stages cells={cells} block={block} periodic={periodic}")
    h = 0.4
    grains = np.array([[0.3 * cells[0] * h, 0.45 * cells[1] * h, 2.0],
                       [0.75 * cells[0] * h, 0.2 * cells[1] * h, 1.2]])
    og = orc.grid(cells, block, [0, 0], [h, h], periodic=periodic)
    n = orc.field_len(og, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(og, phi, u, grains, 0.6)
    wp, wu = phi.copy(), u.copy()
    orc.modelc_rk4_run(og, orc.params(0.05, 3.19, 2.0), wp, wu, 8, 0.008)
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [h, h], periodic=periodic)
    outs = []
    try:
        for kernel, s, pf in ((2, strips, prefetch), (1, 0, 0)):
            ctx.set_tuning("modelc_kernel", kernel)
            ctx.set_tuning("modelc_strips", s)
            ctx.set_tuning("modelc_prefetch", pf)
            sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0), mode=lg.LGD_MODE_FUSED)
            sim.set_integrator("rk4")
            sim.field(0).upload(phi)
            sim.field(1).upload(u)
            sim.step(0.008, 8)
            outs.append((sim.field(0).download(), sim.field(1).download()))
            sim.close()
    finally:
        for key in ("modelc_kernel", "modelc_strips", "modelc_prefetch"):
            ctx.set_tuning(key, 0)
    grid.close()
    m = interior_mask(cells, block)
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0][m], wp[m]) and np.array_equal(outs[0][1][m], wu[m])
