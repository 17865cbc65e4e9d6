"""GPU parity of the three-level adaptive path and of the reference's AMR tests that
pin the transfers (tests/amr.rs), against the oracle on the same inputs: the
hierarchy after every step, and the fields (bitwise without noise, within 1e-10
with it)."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def patches(g):
    return [g.patch(b) for b in range(g.num_blocks)]


def test_three_level_regrid_builds_level_one_from_nesting_flags(ctx, orc):
    """The level-1 boxes come from the nesting flags alone (adapt.rs:194-211)."""
    say("amr nesting-flag regrid (stripes)")
    S = H.STRIPES
    og = orc.AmrGrid.from_patches(orc.grid(S["cells"], S["block"], [0, 0], [S["h"], S["h"]]),
                                  S["ratios"], S["levels"])
    f = H.stripes_field(og, og.field_len(1))
    sim_o = orc.AmrSim(og, "heat", kappa=0.05, policy=orc.RegridPolicy(*S["policy"]))
    sim_o.field(0)[:] = f
    gg = lg.AmrGrid.from_patches(ctx, S["cells"], S["block"], [0, 0], [S["h"], S["h"]], S["ratios"],
                                 S["levels"])
    sim_g = lg.amr_simulation(gg, "heat", kappa=0.05, policy=lg.RegridPolicy(*S["policy"]))
    sim_g.field(0).upload(f)
    for s in range(6):
        sim_o.step(0.01)
        sim_g.step(0.01)
        assert patches(sim_g.amr) == patches(sim_o.grid), f"step {s}"
        if s == 0:
            assert sorted(p for p in patches(sim_g.amr) if p[0] > 0) == S["expect"]
    m = H.amr_interior_mask(sim_g.amr)
    assert np.array_equal(sim_g.field(0).download()[m], sim_o.field(0)[m])
    sim_g.close()


@pytest.mark.parametrize("noise", [0.0, 0.01])
def test_three_level_adaptive_subcycled_model_c_matches_oracle(ctx, orc, noise):
    """The batched-AMR configuration at reduced size: 128² base in 32² blocks, ratios
    [2, 2], subcycled Model C, gradient tagging at 0.1 (buffer 2, efficiency 0.8,
    minimum width 8), regrid every 4 coarse steps, 4 seeds; 40 coarse steps with the
    hierarchy compared after every one."""
    say(f"amr 3-level adaptive subcycled model C noise={noise}")
    h, n, blk = 0.4, 128, 32
    base = orc.grid([n, n], [blk, blk], [0, 0], [h, h])
    pol = (4, 2, 0.8, 8, 0.1, 0)
    sim_o = orc.AmrSim(orc.AmrGrid.from_patches(base, [2, 2], []), "modelc",
                       params=orc.params(0.05, 3.19, 2.0, noise), seed=3, wiener=noise > 0,
                       integ="subcycle", policy=orc.RegridPolicy(*pol))
    seeds = orc.seed_positions(42, 4, n * h, 2.0)
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, seeds, 0.55)
    sim_o.field(0)[:] = phi
    sim_o.field(1)[:] = u
    gg = lg.AmrGrid.from_patches(ctx, [n, n], [blk, blk], [0, 0], [h, h], [2, 2], [])
    sim_g = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.05, 3.19, 2.0, noise),
                              wiener=noise > 0, seed=3, subcycle=True, policy=lg.RegridPolicy(*pol))
    sim_g.field(0).upload(phi)
    sim_g.field(1).upload(u)
    dt = 0.2 * h * h / 2.0
    deepest = 0
    for s in range(40):
        sim_o.step(dt)
        sim_g.step(dt)
        assert sim_g.time == sim_o.time
        assert patches(sim_g.amr) == patches(sim_o.grid), f"step {s}: hierarchies differ"
        deepest = max(deepest, sim_g.amr.num_levels)
    assert deepest == 3
    assert H.proper_nesting_violations(sim_g.amr) == []
    m = H.amr_interior_mask(sim_g.amr)
    gp, gu = sim_g.field(0).download()[m], sim_g.field(1).download()[m]
    if noise == 0.0:
        assert np.array_equal(gp, sim_o.field(0)[m]) and np.array_equal(gu, sim_o.field(1)[m])
    else:
        assert np.max(np.abs(gp - sim_o.field(0)[m])) <= 1e-10
        assert np.max(np.abs(gu - sim_o.field(1)[m])) <= 1e-10
    sim_g.close()


# Static three-level hierarchies with a refinement ratio other than 2: (ratios,
# level-1 boxes, level-2 boxes), on a 48² base in 24² blocks.  [3, 2] has nine
# level-1 substeps per coarse step and a ratio-2 finest level (its restriction shares
# its last stage launch); [2, 3] has a ratio-3 finest level (restriction in launches
# of its own).  Each has a level-2 patch that spans two level-1 patches.
MIXED_RATIOS = {
    "3x2": ([3, 2], [((12, 12), (60, 48)), ((60, 12), (96, 42))],
            [((30, 30), (90, 80)), ((100, 30), (170, 76))]),
    "2x3": ([2, 3], [((8, 8), (40, 32)), ((40, 8), (56, 24))],
            [((30, 30), (90, 72)), ((96, 30), (150, 60))]),
}


@pytest.mark.parametrize("noise", [0.0, 0.02])
@pytest.mark.parametrize("name", list(MIXED_RATIOS))
def test_mixed_ratio_static_hierarchy_model_c_matches_oracle(ctx, orc, name, noise):
    """Subcycled Model C on a static three-level hierarchy with ratios [3, 2] and
    [2, 3], whose levels take their first fill from the parent's stage launch at every
    ratio: fields after 6 coarse steps bitwise against the oracle without noise,
    within 1e-10 with it."""
    say(f"amr mixed ratios {name} noise={noise}")
    ratios, l1, l2 = MIXED_RATIOS[name]
    h, n, blk = 0.4, 48, 24
    levels = [l1, l2]
    og = orc.AmrGrid.from_patches(orc.grid([n, n], [blk, blk], [0, 0], [h, h]), ratios, levels)
    gg = lg.AmrGrid.from_patches(ctx, [n, n], [blk, blk], [0, 0], [h, h], ratios, levels)
    sim_o = orc.AmrSim(og, "modelc", params=orc.params(0.05, 3.19, 2.0, noise), seed=5,
                       wiener=noise > 0, integ="subcycle")
    gr = sim_o.grid
    phi, u = sim_o.field(0), sim_o.field(1)
    for b in range(gr.num_blocks):
        vp, vu = gr.view(phi, b), gr.view(u, b)
        ny, nx = vp.shape[0] - 2, vp.shape[1] - 2
        for j in range(ny):
            for i in range(nx):
                x, y = gr.cell_center(b, i, j)
                d = math.hypot(x - 6.5, y - 5.0) - 3.0
                vp[j + 1, i + 1] = -math.tanh(d / math.sqrt(2.0))
                vu[j + 1, i + 1] = -0.55
    sim_g = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.05, 3.19, 2.0, noise),
                              wiener=noise > 0, seed=5, subcycle=True)
    sim_g.field(0).upload(phi)
    sim_g.field(1).upload(u)
    dt = 0.2 * h * h / 2.0
    sim_o.step(dt, 6)
    sim_g.step(dt, 6)
    assert sim_g.time == sim_o.time
    m = H.amr_interior_mask(sim_g.amr)
    gp, gu = sim_g.field(0).download()[m], sim_g.field(1).download()[m]
    if noise == 0.0:
        assert np.array_equal(gp, sim_o.field(0)[m]) and np.array_equal(gu, sim_o.field(1)[m])
    else:
        assert np.max(np.abs(gp - sim_o.field(0)[m])) <= 1e-10
        assert np.max(np.abs(gu - sim_o.field(1)[m])) <= 1e-10
    sim_g.close()


def lin(c):
    return H._fma(3.0, c[0], 2.0) - c[1]


def box_tagger():
    calls = []

    def tag(grid, level):
        if level != 0:
            return []
        shift = 2 if calls else 0
        calls.append(level)
        return [(x, y) for y in range(2 + shift, 6 + shift) for x in range(2 + shift, 6 + shift)]

    return tag


def test_migration_preserves_linear_fields_exactly(ctx, orc):
    """tests/amr.rs:846-900 through the C ABI's TagCells callback: a forced box moved
    mid-run under a frozen model (heat with κ = 0) keeps the linear field, through
    the same-level copy and the prolongation; the bits equal the oracle's."""
    say("amr migration of a linear field (box tagger)")
    pol = (1, 0, 0.9, 2, 0.0, 0)
    base = orc.grid([16, 16], [8, 8], [0, 0], [0.5, 0.5])
    sim_o = orc.AmrSim(orc.AmrGrid.from_patches(base, [2], []), "heat", kappa=0.0,
                       policy=orc.RegridPolicy(*pol))
    sim_o.set_tagger(box_tagger())
    sim_o.grid.fill_by_center(sim_o.field(0), lin)
    gg = lg.AmrGrid.from_patches(ctx, [16, 16], [8, 8], [0, 0], [0.5, 0.5], [2], [])
    sim_g = lg.amr_simulation(gg, "heat", kappa=0.0, policy=lg.RegridPolicy(*pol))
    sim_g.set_tagger(box_tagger())
    sim_g.field(0).upload(sim_o.field(0))
    for s, box in enumerate([((4, 4), (12, 12)), ((8, 8), (16, 16))]):
        sim_o.step(1e-3)
        sim_g.step(1e-3)
        assert [p for p in patches(sim_g.amr) if p[0] == 1] == [(1,) + box], f"step {s}"
        assert patches(sim_g.amr) == patches(sim_o.grid)
    m = H.amr_interior_mask(sim_g.amr)
    got = sim_g.field(0).download()
    assert np.array_equal(got[m], sim_o.field(0)[m])
    gr = sim_o.grid
    for b in gr.blocks_at(1):
        v = gr.view(got, b)
        for j in range(8):
            for i in range(8):
                want = lin(gr.cell_center(b, i, j))
                assert abs(v[j + 1, i + 1] - want) <= 1e-12 * max(1.0, abs(want))
    sim_g.close()


def test_a_failing_tagger_fails_the_step(ctx):
    say("amr failing tagger")
    gg = lg.AmrGrid.from_patches(ctx, [16, 16], [8, 8], [0, 0], [0.5, 0.5], [2], [])
    sim = lg.amr_simulation(gg, "heat", kappa=0.0, policy=lg.RegridPolicy(1, 0, 0.9, 2, 0.0, 0))

    def bad(grid, level):
        raise RuntimeError("no tags")

    sim.set_tagger(bad)
    with pytest.raises(lg.LegendreError) as e:
        sim.step(1e-3)
    assert e.value.status == 1
    sim.set_tagger(None)  # the GradientTagger again
    sim.step(1e-3)
    sim.close()


@pytest.mark.parametrize("model", ["heat", "modelc"])
def test_simulation_on_amr_matches_cartesian_bitwise(ctx, orc, model):
    """tests/amr.rs:192: an AmrGrid without patches runs the Cartesian simulation bit
    for bit (Model C with noise: the keys are the same block ids and cells)."""
    say(f"amr without patches == cartesian ({model})")
    cells, block, h = [96, 64], [32, 32], 0.4
    gg = lg.AmrGrid.from_patches(ctx, cells, block, [0, 0], [h, h], [], [])
    cg = lg.CartesianGrid(ctx, cells, block, [0, 0], [h, h])
    og = orc.grid(cells, block, [0, 0], [h, h])
    if model == "heat":
        sa = lg.amr_simulation(gg, "heat", kappa=0.7)
        sc = lg.Simulation.heat(cg, 0.7)
        u = np.zeros(orc.field_len(og, 1))
        uv = H.field_view(u, cells, block, 1)
        for b in range(uv.shape[0]):
            for j in range(block[1]):
                for i in range(block[0]):
                    x, y = orc.cell_center(og, b, (i, j))
                    uv[b, j + 1, i + 1] = math.cos(0.7 * x) + math.sin(1.3 * y)
        fields = [u]
        dt = 0.2 * h * h / 0.7
    else:
        p = lg.ModelCParams(0.05, 3.19, 2.0, 0.02)
        sa = lg.amr_simulation(gg, "modelc", params=p, wiener=True, seed=9)
        sc = lg.Simulation.model_c(cg, p, wiener=True, seed=9)
        phi = np.zeros(orc.field_len(og, 1))
        u = np.zeros_like(phi)
        orc.initialize_grains(og, phi, u, np.array([[14.0, 9.0, 2.0], [26.0, 14.0, 1.2]]), 0.55)
        fields = [phi, u]
        dt = 0.008
    for i, f in enumerate(fields):
        sa.field(i).upload(f)
        sc.field(i).upload(f)
    sa.step(dt, 30)
    sc.step(dt, 30)
    m = np.zeros(fields[0].size, bool)
    H.field_view(m, cells, block, 1)[(slice(None),) + H.interior_slices(2, 1)] = True
    for i in range(len(fields)):
        assert np.array_equal(sa.field(i).download()[m], sc.field(i).download()[m])
    sa.close()
    sc.close()


def test_interface_run_stays_physical(ctx, orc):
    """tests/amr.rs:491-537 on the device, bitwise against the oracle."""
    say("amr interface run")
    kappa, hf = 0.7, 0.25
    dt = 0.2 * hf * hf / kappa
    levels = [[((8, 8), (24, 24))]]
    base = orc.grid([16, 16], [8, 8], [0, 0], [0.5, 0.5])
    sim_o = orc.AmrSim(orc.AmrGrid.from_patches(base, [2], levels), "heat", kappa=kappa)
    init = lambda c: math.exp(-H._fma(c[0] - 4.0, c[0] - 4.0, (c[1] - 4.0) * (c[1] - 4.0)) / 1.5)
    sim_o.grid.fill_by_center(sim_o.field(0), init)
    gg = lg.AmrGrid.from_patches(ctx, [16, 16], [8, 8], [0, 0], [0.5, 0.5], [2], levels)
    sim_g = lg.amr_simulation(gg, "heat", kappa=kappa)
    sim_g.field(0).upload(sim_o.field(0))
    sim_o.step(dt, 200)
    sim_g.step(dt, 200)
    m = H.amr_interior_mask(sim_g.amr)
    got = sim_g.field(0).download()[m]
    assert np.array_equal(got, sim_o.field(0)[m])
    assert np.all(np.isfinite(got)) and got.max() < 0.9 and got.max() > 0.0 and got.min() > -1e-9
    sim_g.close()


def slab_pattern(nslab: int, b: int) -> np.ndarray:
    """tests/amr.rs:35-45: slab entry i of block b = ((i·37 + b·101) mod 89)·0.037 − 1.2
    (mul_add)."""
    return np.array([H._fma(float((i * 37 + b * 101) % 89), 0.037, -1.2) for i in range(nslab)])


@pytest.mark.parametrize("model", ["heat", "modelc"])
def test_a_patch_steps_like_its_uniform_twin_bitwise(ctx, orc, model):
    """tests/amr.rs:67-125: the Laplacian (heat) and the Karma–Rappel operator (Model
    C) on a 6×6 level-1 patch equal the uniform single-block twin bit for bit.  The
    patch is stepped by the AMR level kernel after its ghost fill; the twin holds the
    same slab, ghosts included, and takes the per-intent Cartesian step."""
    say(f"amr patch == uniform twin ({model})")
    levels = [[((4, 4), (10, 10))]]
    og = orc.AmrGrid.from_patches(orc.grid([8, 8], [8, 8], [0, 0], [0.5, 0.5]), [2], levels)
    n = og.field_len(1)
    fields = []
    for k in range(2 if model == "modelc" else 1):
        f = np.zeros(n)
        for b in range(og.num_blocks):
            v = og.view(f, b).reshape(-1)
            v[:] = slab_pattern(v.size, b + 3 * k)
        if model == "modelc" and k == 0:
            f = np.tanh(f)  # phi within (-1, 1)
        fields.append(f)
    # The ghost values the AMR step's fill writes (the device fill is bitwise the
    # oracle's restriction + fill).
    filled = [f.copy() for f in fields]
    for f in filled:
        og.restrict(f)
        og.fill_ghosts(f)
    gg = lg.AmrGrid.from_patches(ctx, [8, 8], [8, 8], [0, 0], [0.5, 0.5], [2], levels)
    tg = lg.CartesianGrid(ctx, [6, 6], [6, 6], [0, 0], [0.25, 0.25])
    dt = 1e-3
    if model == "heat":
        sa = lg.amr_simulation(gg, "heat", kappa=0.7)
    else:
        p = lg.ModelCParams(0.06, 3.19, 2.0, 0.0)
        sa = lg.amr_simulation(gg, "modelc", params=p)
    for i, f in enumerate(fields):
        sa.field(i).upload(f)
    sa.step(dt)
    # Twin: the patch slab (block 1) with its filled ghost ring.
    tw = []
    for f in filled:
        d = tg.alloc_field(1)
        d.upload(og.view(f, 1).reshape(-1).copy())
        tw.append(d)
    ks = [tg.alloc_field(1) for _ in tw]
    if model == "heat":
        tg.heat_rhs(0.7, tw[0], ks[0])
    else:
        tg.modelc_rhs(p, tw[0], tw[1], ks[0], ks[1])
    for x, k in zip(tw, ks):
        tg.axpy(x, k, dt)
    sl = (slice(1, -1), slice(1, -1))
    for i, x in enumerate(tw):
        got = og.view(sa.field(i).download(), 1)[sl]
        want = x.download().reshape(8, 8)[sl]
        assert np.array_equal(got, want), f"field {i}"
    sa.close()


@pytest.mark.parametrize("model", ["modelc", "heat"])
def test_fused_next_fill_matches_split_fill_bitwise(ctx, orc, model):
    """The finest level's stage launch filling the next substep's ghosts, and a level's
    stage launch filling its finer level's first substep's ghosts (the defaults),
    compute the bits of a separate fill launch per substep: three-level adaptive
    subcycled runs, Model C with noise and heat, hierarchy and fields after every
    step."""
    say(f"amr fused next fill vs split fill ({model})")
    h, n, blk = 0.4, 96, 24
    base = orc.grid([n, n], [blk, blk], [0, 0], [h, h])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[10.0, 9.0, 3.0], [22.0, 25.0, 2.0]]), 0.55)
    pol = lg.RegridPolicy(4, 2, 0.8, 4, 0.1, 0)
    runs = []
    try:
        for split in (0, 1):
            ctx.set_tuning("amr_split_fill", split)
            gg = lg.AmrGrid.from_patches(ctx, [n, n], [blk, blk], [0, 0], [h, h], [2, 2], [])
            if model == "modelc":
                sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.05, 3.19, 2.0, 0.01),
                                        wiener=True, seed=4, subcycle=True, policy=pol)
                sim.field(1).upload(u)
                dt = 0.2 * h * h / 2.0
            else:
                sim = lg.amr_simulation(gg, "heat", kappa=0.7, subcycle=True, policy=pol)
                dt = 0.25 * h * h / 0.7
            sim.field(0).upload(phi)
            trace = []
            for s in range(16):
                sim.step(dt)
                a = sim.amr
                m = H.amr_interior_mask(a)
                trace.append((patches(a), [sim.field(f).download()[m] for f in range(sim.nfields)]))
            runs.append(trace)
            sim.close()
    finally:
        ctx.set_tuning("amr_split_fill", 0)
    assert any(p[0] == 2 for tr in runs[0] for p in tr[0])  # level 2 existed
    for s, (a, b) in enumerate(zip(*runs)):
        assert a[0] == b[0], f"step {s}: hierarchies differ"
        for fa, fb in zip(a[1], b[1]):
            assert np.array_equal(fa, fb), f"step {s}: fields differ"


def test_amr_launch_shapes_change_no_bit(ctx, orc):
    """Every level-stage form (streaming at several strip heights, per-cell at several
    tile shapes, per-patch launches), fill / restriction width, level-0 kernel (one
    block or two per wavefront, or per-cell with its own fill) and the autotuner's
    choice give the bits of the
    default shapes: three-level adaptive subcycled Model C with noise, hierarchy and
    fields after every step; the autotuner leaves the simulation it times bitwise
    unchanged."""
    say("amr launch shapes")
    h, n, blk = 0.4, 96, 24
    base = orc.grid([n, n], [blk, blk], [0, 0], [h, h])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[10.0, 9.0, 3.0], [22.0, 25.0, 2.0]]), 0.55)
    pol = lg.RegridPolicy(4, 2, 0.8, 4, 0.1, 0)
    dt = 0.2 * h * h / 2.0

    def run(knobs, tune=False):
        for k, v in knobs.items():
            ctx.set_tuning(k, v)
        gg = lg.AmrGrid.from_patches(ctx, [n, n], [blk, blk], [0, 0], [h, h], [2, 2], [])
        sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.05, 3.19, 2.0, 0.01),
                                wiener=True, seed=4, subcycle=True, policy=pol)
        sim.field(0).upload(phi)
        sim.field(1).upload(u)
        trace = []
        for s in range(12):
            sim.step(dt)
            if tune and s == 5:
                r = sim.autotune_amr(dt, reps=2)
                assert r["candidates"] >= 12 and r["ms"] > 0.0
            a = sim.amr
            m = H.amr_interior_mask(a)
            trace.append((patches(a), sim.field(0).download()[m], sim.field(1).download()[m]))
        sim.close()
        return trace

    shapes = [{}, {"amr_kernel": 1, "amr_tile": 1, "amr_wg": 64},
              {"amr_kernel": 1, "amr_tile": 2, "amr_wg": 1024},
              {"amr_kernel": 1, "amr_tile": 6, "modelc_kernel": 3},
              {"amr_rows": 2, "amr_wg": 128}, {"amr_rows": 5, "modelc_kernel": 2},
              {"amr_rows": 16, "modelc_kernel": 2, "modelc_strips": 1, "modelc_prefetch": 1},
              {"amr_per_block": 1}, {"amr_per_block": 1, "amr_kernel": 1},
              {"amr_kernel": 1, "modelc_kernel": 1}, {"modelc_kernel": 1}]
    try:
        ref = run({})
        assert any(p[0] == 2 for tr in ref for p in tr[0])
        for knobs in shapes[1:]:
            got = run(knobs)
            for s, (a, b) in enumerate(zip(ref, got)):
                assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), (knobs, s)
            for k in knobs:
                ctx.set_tuning(k, 0)
        got = run({}, tune=True)
        for s, (a, b) in enumerate(zip(ref, got)):
            assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), ("tuned", s)
    finally:
        for k in ("amr_tile", "amr_wg", "amr_kernel", "amr_rows", "amr_per_block", "modelc_kernel",
                  "modelc_strips", "modelc_prefetch"):
            ctx.set_tuning(k, 0)
