"""GPU parity of the AMR path (2-D): device transfers, level kernels, global-dt and
subcycled stepping, and adaptive regrid, against the oracle's restatement of
src/geometry/amr and src/integrators/subcycling.rs on the same inputs."""
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


HIERARCHIES = {
    # name: (cells, block, h, ratios, levels)
    "two_patches": ([8, 8], [4, 4], 0.5, [2], [[((4, 4), (8, 12)), ((8, 4), (12, 12))]]),
    "three_levels": ([16, 16], [8, 8], 0.5, [2, 2],
                     [[((4, 4), (20, 24)), ((20, 8), (28, 20))], [((12, 12), (36, 40)), ((40, 20), (52, 36))]]),
    "edge_patch": ([12, 12], [6, 6], 0.4, [2], [[((0, 0), (10, 8)), ((14, 16), (24, 24))]]),
}


def pair(ctx, orc, name):
    cells, block, h, ratios, levels = HIERARCHIES[name]
    og = orc.AmrGrid.from_patches(orc.grid(cells, block, [0, 0], [h, h]), ratios, levels)
    gg = lg.AmrGrid.from_patches(ctx, cells, block, [0, 0], [h, h], ratios, levels)
    assert og.field_len(1) == gg.field_len
    for b in range(og.num_blocks):
        assert og.patch(b) == gg.patch(b)
        assert og.offset(b, 1) == gg.offset(b)
    return og, gg


def smooth(c):
    return math.cos(0.7 * c[0]) + math.sin(1.3 * c[1]) + 0.1 * c[0] * c[1]


@pytest.mark.parametrize("name", list(HIERARCHIES))
def test_fill_and_restrict_match_oracle(ctx, orc, name):
    say(f"amr fill/restrict {name}")
    og, gg = pair(ctx, orc, name)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(og.field_len(1))
    d = lg.DeviceBuffer(ctx, f.size)
    d.upload(f)
    for l in range(og.num_levels - 1, 0, -1):
        og.restrict_level(f, l)
        gg.restrict_level(d, l)
    assert np.array_equal(d.download(), f)
    for l in range(og.num_levels):
        og.fill_level(f, l)
        gg.fill_level(d, l)
    assert np.array_equal(d.download(), f)
    # time interpolation of the coarse source
    old = rng.standard_normal(f.size)
    dold = lg.DeviceBuffer(ctx, f.size)
    dold.upload(old)
    for l in range(1, og.num_levels):
        og.fill_level(f, l, old=old, alpha=0.25)
        gg.fill_level(d, l, old=dold, alpha=0.25)
    assert np.array_equal(d.download(), f)


def test_invalid_hierarchies_are_rejected(ctx):
    say("amr invalid hierarchy")
    for levels in ([[((4, 4), (8, 8)), ((6, 6), (10, 10))]], [[((3, 4), (7, 8))]], [[((4, 4), (4, 8))]]):
        with pytest.raises(lg.LegendreError) as e:
            lg.AmrGrid.from_patches(ctx, [8, 8], [4, 4], [0, 0], [0.5, 0.5], [2], levels)
        assert e.value.status == 1


@pytest.mark.parametrize("name", list(HIERARCHIES))
@pytest.mark.parametrize("subcycle", [False, True])
def test_heat_on_static_hierarchy_matches_oracle_bitwise(ctx, orc, name, subcycle):
    say(f"amr heat {name} subcycle={subcycle}")
    og, gg = pair(ctx, orc, name)
    kappa = 0.7
    sim_o = orc.AmrSim(og, "heat", kappa=kappa, integ="subcycle" if subcycle else "global")
    sim_o.grid.fill_by_center(sim_o.field(0), smooth)
    sim_g = lg.amr_simulation(gg, "heat", kappa=kappa, subcycle=subcycle)
    sim_g.field(0).upload(sim_o.field(0))
    h0 = HIERARCHIES[name][2]
    nlev = len(HIERARCHIES[name][3]) + 1
    dt = 0.25 * (h0 if subcycle else h0 / 2 ** (nlev - 1)) ** 2 / kappa
    sim_o.step(dt, 12)
    sim_g.step(dt, 12)
    assert sim_g.time == sim_o.time
    m = H.amr_interior_mask(sim_g.amr)  # ghost entries are scratch after a step
    assert np.array_equal(sim_g.field(0).download()[m], sim_o.field(0)[m])


# Static Model C hierarchies: (base cells, base block, level-1 boxes, seeds (x, y, r)).
# "wide" has level-1 patches 160 and 64 cells wide, which the streaming level stage
# cuts into three and two column segments.
MODELC_LAYOUTS = {
    "small": ([48, 48], [24, 24], [[((8, 8), (40, 32)), ((40, 8), (56, 24))]], [(6.0, 5.0, 3.0)]),
    "wide": ([96, 64], [48, 32], [[((8, 8), (168, 40)), ((8, 48), (72, 120))]],
             [(12.0, 5.0, 3.0), (6.0, 14.0, 2.0)]),
}


@pytest.mark.parametrize("subcycle", [False, True])
@pytest.mark.parametrize("noise", [0.0, 0.02])
@pytest.mark.parametrize("layout", list(MODELC_LAYOUTS))
@pytest.mark.parametrize("form", ["streaming", "per-cell"])
def test_modelc_on_static_hierarchy_matches_oracle(ctx, orc, subcycle, noise, layout, form):
    """Model C on a static two-level hierarchy against the oracle, through the
    streaming level stage and the per-cell one: bitwise without noise, within 1e-10
    with it (the noise draws agree to a few ulp)."""
    say(f"amr modelc subcycle={subcycle} noise={noise} layout={layout} form={form}")
    cells, block, levels, seeds = MODELC_LAYOUTS[layout]
    h = 0.4
    ctx.set_tuning("amr_kernel", 1 if form == "per-cell" else 0)
    og = orc.AmrGrid.from_patches(orc.grid(cells, block, [0, 0], [h, h]), [2], levels)
    gg = lg.AmrGrid.from_patches(ctx, cells, block, [0, 0], [h, h], [2], levels)
    p = orc.params(0.05, 3.19, 2.0, noise)
    sim_o = orc.AmrSim(og, "modelc", params=p, seed=11, wiener=noise > 0,
                       integ="subcycle" if subcycle else "global")
    gr = sim_o.grid
    phi, u = sim_o.field(0), sim_o.field(1)
    for b in range(gr.num_blocks):
        vp, vu = gr.view(phi, b), gr.view(u, b)
        ny, nx = vp.shape[0] - 2, vp.shape[1] - 2
        for j in range(ny):
            for i in range(nx):
                x, y = gr.cell_center(b, i, j)
                d = min(math.hypot(x - sx, y - sy) - r for sx, sy, r in seeds)
                vp[j + 1, i + 1] = -math.tanh(d / math.sqrt(2.0))
                vu[j + 1, i + 1] = -0.55
    sim_g = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.05, 3.19, 2.0, noise),
                              wiener=noise > 0, seed=11, subcycle=subcycle)
    sim_g.field(0).upload(phi)
    sim_g.field(1).upload(u)
    dt = 0.2 * (h if subcycle else h / 2) ** 2 / 2.0
    sim_o.step(dt, 20)
    try:
        sim_g.step(dt, 20)
    finally:
        ctx.set_tuning("amr_kernel", 0)
    gp, gu = sim_g.field(0).download(), sim_g.field(1).download()
    m = H.amr_interior_mask(sim_g.amr)  # ghost entries are scratch after a step
    if noise == 0.0:
        assert np.array_equal(gp[m], sim_o.field(0)[m]) and np.array_equal(gu[m], sim_o.field(1)[m])
    else:
        assert np.max(np.abs(gp[m] - sim_o.field(0)[m])) <= 1e-10
        assert np.max(np.abs(gu[m] - sim_o.field(1)[m])) <= 1e-10


def gaussian(c):
    dx, dy = c[0] - 4.0, c[1] - 4.0
    return math.exp(-H._fma(dx, dx, dy * dy) / 0.4)


@pytest.mark.parametrize("subcycle", [False, True])
def test_adaptive_heat_matches_oracle_and_golden_numbers(ctx, orc, subcycle):
    """The reference's adaptive test (tests/amr.rs:682-788) on the device: the
    same regrid decisions (hierarchy after every step) and the same bits; with a
    global dt (the reference's configuration) the device state reproduces the
    reference's golden numbers err_adaptive = 0.001680900 and refined_cells = 400."""
    say(f"amr adaptive heat subcycle={subcycle}")
    kappa = 0.7
    dt = 0.2 * 0.25 * 0.25 / kappa
    pol = (4, 2, 0.7, 2, 0.05, 0)
    base = orc.grid([16, 16], [8, 8], [0, 0], [0.5, 0.5])
    sim_o = orc.AmrSim(orc.AmrGrid.from_patches(base, [2], []), "heat", kappa=kappa,
                       integ="subcycle" if subcycle else "global", policy=orc.RegridPolicy(*pol))
    sim_o.grid.fill_by_center(sim_o.field(0), gaussian)
    gg = lg.AmrGrid.from_patches(ctx, [16, 16], [8, 8], [0, 0], [0.5, 0.5], [2], [])
    sim_g = lg.amr_simulation(gg, "heat", kappa=kappa, subcycle=subcycle, policy=lg.RegridPolicy(*pol))
    sim_g.field(0).upload(sim_o.field(0))
    refined = 0
    for s in range(60):
        sim_o.step(dt)
        sim_g.step(dt)
        og, ga = sim_o.grid, sim_g.amr
        assert og.num_blocks == ga.num_blocks, f"step {s}"
        for b in range(og.num_blocks):
            assert og.patch(b) == ga.patch(b), f"step {s} block {b}"
        refined = max(refined, sum((hi[0] - lo[0]) * (hi[1] - lo[1])
                                   for lev, lo, hi in (ga.patch(b) for b in range(ga.num_blocks)) if lev == 1))
    m = H.amr_interior_mask(sim_g.amr)
    assert np.array_equal(sim_g.field(0).download()[m], sim_o.field(0)[m])
    if subcycle:
        return
    # The golden numbers from the device state: restrict on the device, gather
    # level 0, compare with the fine uniform run averaged onto the coarse cells.
    ga = sim_g.amr
    f = sim_g.field(0)
    if ga.num_levels > 1:  # the bump has usually diffused below the threshold by now
        ga.restrict_level(f, 1)
    v = f.download()
    amr0 = np.zeros((16, 16))
    for b in range(ga.num_blocks):
        lev, lo, hi = ga.patch(b)
        if lev == 0:
            o = ga.offset(b)
            amr0[lo[1]:hi[1], lo[0]:hi[0]] = v[o:o + 100].reshape(10, 10)[1:-1, 1:-1]
    ug = orc.grid([32, 32], [16, 16], [0, 0], [0.25, 0.25])
    xs = (np.arange(32) + 0.5) * 0.25
    uf = H.global_to_field(np.array([[gaussian((x, y)) for x in xs] for y in xs]), [32, 32], [16, 16], 1)
    orc.heat_run(ug, kappa, uf, 60, dt)
    fine = H.field_to_global(uf, [32, 32], [16, 16], 1)
    ref = np.zeros((16, 16))
    for y in range(32):
        for x in range(32):
            ref[y // 2, x // 2] = H._fma(fine[y, x], 0.25, ref[y // 2, x // 2])
    err_adaptive = float(np.max(np.abs(amr0 - ref)))
    assert abs(err_adaptive - 0.001680900) <= 1e-6, err_adaptive
    assert refined == 400


def test_adaptive_subcycled_model_c_matches_oracle_bitwise(ctx, orc):
    """tests/amr.rs:1181-1228 on the device: adaptive, subcycled Model C."""
    say("amr adaptive subcycled model C")
    base = orc.grid([48, 48], [24, 24], [0, 0], [0.4, 0.4])
    p = orc.params(0.06, 3.19, 0.6267 * 3.19, 0.0)
    pol = (4, 2, 0.8, 4, 0.15, 0)
    sim_o = orc.AmrSim(orc.AmrGrid.from_patches(base, [2], []), "modelc", params=p, seed=0,
                       integ="subcycle", policy=orc.RegridPolicy(*pol))
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0]]), 0.7)
    sim_o.field(0)[:] = phi
    sim_o.field(1)[:] = u
    gg = lg.AmrGrid.from_patches(ctx, [48, 48], [24, 24], [0, 0], [0.4, 0.4], [2], [])
    sim_g = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 0.6267 * 3.19, 0.0),
                              subcycle=True, policy=lg.RegridPolicy(*pol))
    sim_g.field(0).upload(phi)
    sim_g.field(1).upload(u)
    dt = 0.2 * 0.4**2 / p.d_thermal
    sim_o.step(dt, 40)
    sim_g.step(dt, 40)
    assert sim_g.amr.num_levels == 2
    assert sim_g.amr.num_blocks == sim_o.grid.num_blocks
    m = H.amr_interior_mask(sim_g.amr)
    assert np.array_equal(sim_g.field(0).download()[m], sim_o.field(0)[m])
    assert np.array_equal(sim_g.field(1).download()[m], sim_o.field(1)[m])


def test_per_block_dispatch_matches_batched_bitwise(ctx, orc):
    """One launch per patch (the per-block dispatch baseline) computes the same bits
    as one launch per level, noise included."""
    say("amr per-block vs batched dispatch")
    base = orc.grid([48, 48], [24, 24], [0, 0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0], [12.0, 9.0, 2.0]]), 0.7)
    pol = lg.RegridPolicy(4, 2, 0.8, 4, 0.15, 0)
    outs = []
    try:
        for per_block in (0, 1):
            ctx.set_tuning("amr_per_block", per_block)
            gg = lg.AmrGrid.from_patches(ctx, [48, 48], [24, 24], [0, 0], [0.4, 0.4], [2, 2], [])
            sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 2.0, 0.01),
                                    wiener=True, seed=5, subcycle=True, policy=pol)
            sim.field(0).upload(phi)
            sim.field(1).upload(u)
            sim.step(0.2 * 0.4**2 / 2.0, 24)
            m = H.amr_interior_mask(sim.amr)
            outs.append((sim.amr.num_blocks, sim.field(0).download()[m], sim.field(1).download()[m]))
            sim.close()
    finally:
        ctx.set_tuning("amr_per_block", 0)
    assert outs[0][0] == outs[1][0] and outs[0][0] > 4
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])


@pytest.mark.parametrize("subcycle", [False, True])
def test_device_regrid_matches_host_regrid_bitwise(ctx, orc, subcycle):
    """Tagging, dilation and migration on the device (the default) against the host
    path (the "regrid_host" knob): the same hierarchy after every step and the same
    bits, with noise and three levels."""
    say(f"amr device vs host regrid subcycle={subcycle}")
    base = orc.grid([96, 96], [24, 24], [0, 0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[0.4, 0.4, 4.0], [20.0, 15.0, 3.0], [30.0, 32.0, 2.0]]), 0.7)
    pol = lg.RegridPolicy(3, 2, 0.75, 4, 0.12, 0)
    dt = 0.2 * (0.4 if subcycle else 0.1) ** 2 / 2.0
    runs = []
    try:
        for host in (0, 1):
            ctx.set_tuning("regrid_host", host)
            gg = lg.AmrGrid.from_patches(ctx, [96, 96], [24, 24], [0, 0], [0.4, 0.4], [2, 2], [])
            sim = lg.amr_simulation(gg, "modelc", params=lg.ModelCParams(0.06, 3.19, 2.0, 0.01),
                                    wiener=True, seed=9, subcycle=subcycle, policy=pol)
            sim.field(0).upload(phi)
            sim.field(1).upload(u)
            trace = []
            for s in range(13):
                sim.step(dt)
                a = sim.amr
                m = H.amr_interior_mask(a)
                trace.append(([a.patch(b) for b in range(a.num_blocks)], sim.field(0).download()[m],
                              sim.field(1).download()[m]))
            runs.append(trace)
            sim.close()
    finally:
        ctx.set_tuning("regrid_host", 0)
    dev, host = runs
    assert len(dev[-1][0]) > 16 and any(p[0] == 2 for p in dev[-1][0])
    for s, (a, b) in enumerate(zip(dev, host)):
        assert a[0] == b[0], f"step {s}: hierarchies differ"
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), f"step {s}: fields differ"
