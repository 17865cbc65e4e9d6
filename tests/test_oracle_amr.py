"""Pins the oracle's AMR restatement (2-D): hierarchy validation, restriction,
ghost fill with bilinear prolongation, Berger–Rigoutsos clustering, regrid +
migration and Berger–Oliger subcycling — against the reference's own tests
(tests/amr.rs, src/geometry/amr/{grid,cluster}.rs), including its golden numbers
(err_adaptive = 0.001680900, err_coarse = 0.003716000, refined_cells = 400).

Function → pin:
  from_patches        reference validation cases (ratio, empty, overlap, nesting, alignment)
  restrict/fill       exact reproduction of linear fields (amr.rs:227-288)
  fill (exchange)     full fine tiling ≡ uniform fine grid, bitwise (amr.rs:348-437)
  subcycling          fully tiled subcycled fine ≡ uniform fine at dt/4, bitwise (amr.rs:1015-1110)
  cluster             reference box tests (cluster.rs:320-523)
  tag/regrid/migrate  golden numbers (amr.rs:682-788); coverage of tagged cells (amr.rs:599-677)
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import helpers as H


def base_grid(orc, cells, block, h, origin=(0.0, 0.0)):
    return orc.grid(cells, block, list(origin), [h, h])


# ------------------------------------------------------------- hierarchy
def test_from_patches_validates_like_the_reference(orc):
    base = base_grid(orc, [8, 8], [4, 4], 0.5)
    ok = orc.AmrGrid.from_patches(base, [2], [[((4, 4), (8, 12)), ((8, 4), (12, 12))]])
    assert ok.num_blocks == 6 and ok.num_levels == 2
    bad = {
        "ratio": ([1], [[((4, 4), (8, 8))]]),
        "more levels than ratios": ([2], [[((4, 4), (8, 8))], [((8, 8), (12, 12))]]),
        "empty": ([2], [[((4, 4), (4, 8))]]),
        "overlap": ([2], [[((4, 4), (8, 8)), ((6, 6), (10, 10))]]),
        "misaligned": ([2], [[((3, 4), (7, 8))]]),
        "outside the domain": ([2], [[((12, 12), (18, 18))]]),
    }
    for name, (ratios, levels) in bad.items():
        with pytest.raises(ValueError):
            orc.AmrGrid.from_patches(base, ratios, levels)
    # Proper nesting: a level-2 patch must sit inside level 1 grown by one.
    with pytest.raises(ValueError):
        orc.AmrGrid.from_patches(base, [2, 2], [[((4, 4), (8, 8))], [((8, 8), (16, 16))]])
    nested = orc.AmrGrid.from_patches(base, [2, 2], [[((2, 2), (14, 14))], [((8, 8), (16, 16))]])
    assert nested.num_levels == 3


def test_find_patch_and_cell_keys(orc):
    base = base_grid(orc, [8, 8], [4, 4], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [[((4, 4), (8, 12)), ((8, 4), (12, 12))]])
    assert g.find_patch(1, (5, 5)) == 4 and g.find_patch(1, (9, 11)) == 5
    assert g.find_patch(1, (2, 2)) is None and g.find_patch(0, (7, 0)) == 1
    assert g.cell_center(4, 0, 0) == (4.5 * 0.25, 4.5 * 0.25)


# --------------------------------------------------------------- intergrid
def test_linear_field_ghosts_and_restriction_are_exact(orc):
    # tests/amr.rs:227-288
    lin = lambda c: H._fma(3.0, c[0], 2.0) - c[1]
    base = base_grid(orc, [8, 8], [4, 4], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [[((4, 4), (8, 12)), ((8, 4), (12, 12))]])
    f = np.zeros(g.field_len(1))
    g.fill_by_center(f, lin)
    v0 = g.view(f, 0)
    v0[1 + 2 : 1 + 4, 1 + 2 : 1 + 4] = 99.0  # corrupt coarse cells under the patches
    # (block 0 is the base block at (0,0); the patches cover coarse [2,6)x[2,6))
    for b in range(4):
        v = g.view(f, b)
        _, lo, hi = g.patch(b)
        for j in range(4):
            for i in range(4):
                if 2 <= lo[0] + i < 6 and 2 <= lo[1] + j < 6:
                    v[j + 1, i + 1] = 99.0
    g.restrict(f)
    g.fill_ghosts(f)
    for fb in g.blocks_at(1):
        _, lo, hi = g.patch(fb)
        nx, ny = hi[0] - lo[0], hi[1] - lo[1]
        v = g.view(f, fb)
        for j in range(-1, ny + 1):
            for i in range(-1, nx + 1):
                if 0 <= i < nx and 0 <= j < ny:
                    continue
                assert abs(v[j + 1, i + 1] - lin(g.cell_center(fb, i, j))) <= 1e-12
    for b in g.blocks_at(0):
        v = g.view(f, b)
        for j in range(4):
            for i in range(4):
                assert abs(v[j + 1, i + 1] - lin(g.cell_center(b, i, j))) <= 1e-12


def _gather_level(g, f, level, n):
    out = np.zeros((n, n))
    for b in g.blocks_at(level):
        _, lo, hi = g.patch(b)
        v = g.view(f, b)
        out[lo[1] : hi[1], lo[0] : hi[0]] = v[1:-1, 1:-1]
    return out


def _uniform_heat(orc, init, n, block, h, kappa, dt, steps):
    ug = orc.grid([n, n], block, [0, 0], [h, h])
    xs = (np.arange(n) + 0.5) * h
    u = np.array([[init((x, y)) for x in xs] for y in xs])
    f = H.global_to_field(u, [n, n], block, 1)
    orc.heat_run(ug, kappa, f, steps, dt)
    return H.field_to_global(f, [n, n], block, 1)


def test_full_fine_tiling_matches_uniform_grid_bitwise(orc):
    # tests/amr.rs:348-437
    kappa, hf = 0.7, 0.25
    dt = 0.2 * hf * hf / kappa
    init = lambda c: math.cos(0.7 * c[0]) + math.sin(1.3 * c[1])
    base = base_grid(orc, [8, 8], [8, 8], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [[((0, 0), (8, 16)), ((8, 0), (16, 16))]])
    sim = orc.AmrSim(g, "heat", kappa=kappa)
    gg = sim.grid
    gg.fill_by_center(sim.field(0), init)
    sim.step(dt, 100)
    amr = _gather_level(sim.grid, sim.field(0), 1, 16)
    uni = _uniform_heat(orc, init, 16, [8, 16], hf, kappa, dt, 100)
    assert np.array_equal(amr, uni)


def test_fully_tiled_subcycled_fine_matches_uniform_fine_bitwise(orc):
    # tests/amr.rs:1015-1110
    kappa = 0.7
    init = lambda c: math.cos(0.6 * c[0]) + math.sin(1.1 * c[1])
    base = base_grid(orc, [8, 8], [8, 8], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [[((0, 0), (8, 16)), ((8, 0), (16, 16))]])
    sim = orc.AmrSim(g, "heat", kappa=kappa, integ="subcycle")
    sim.grid.fill_by_center(sim.field(0), init)
    dt_coarse = 0.25 * 0.5**2 / kappa
    sim.step(dt_coarse, 25)
    amr = _gather_level(sim.grid, sim.field(0), 1, 16)
    uni = _uniform_heat(orc, init, 16, [8, 16], 0.25, kappa, dt_coarse / 4.0, 100)
    assert np.array_equal(amr, uni)


# ----------------------------------------------------------------- cluster
def art(rows):
    return [(c, r) for r, row in enumerate(rows) for c, ch in enumerate(row) if ch == "x"]


def _cover_disjoint(flags, boxes):
    for f in flags:
        n = sum(1 for lo, hi in boxes if lo[0] <= f[0] < hi[0] and lo[1] <= f[1] < hi[1])
        assert n == 1
    for i, (a_lo, a_hi) in enumerate(boxes):
        for b_lo, b_hi in boxes[i + 1 :]:
            assert not (a_lo[0] < b_hi[0] and b_lo[0] < a_hi[0] and a_lo[1] < b_hi[1] and b_lo[1] < a_hi[1])


def _cells(boxes):
    return sum((hi[0] - lo[0]) * (hi[1] - lo[1]) for lo, hi in boxes)


def test_cluster_reference_cases(orc):
    # src/geometry/amr/cluster.rs:320-430
    flags = art(["xxxx", "xxxx", "xxxx"])
    assert orc.cluster(flags, 0.8, 4) == [((0, 0), (4, 3))]
    flags = art(["xx....xxx", "xx....xxx", "......xxx"])
    boxes = orc.cluster(flags, 0.8, 4)
    assert len(boxes) == 2 and _cells(boxes) == len(flags)
    _cover_disjoint(flags, boxes)
    flags = art(["xx......"] * 4 + ["xxxxxxxx"] * 2)
    boxes = orc.cluster(flags, 0.95, 2)
    assert len(boxes) == 2 and _cells(boxes) == len(flags)
    _cover_disjoint(flags, boxes)
    flags = art(["x.....x", "x.....x", ".x...x.", ".x...x.", "..x.x..", "..xxx.."])
    boxes = orc.cluster(flags, 0.7, 1)
    _cover_disjoint(flags, boxes)
    assert len(flags) / _cells(boxes) >= 0.7
    diag = [(i, i) for i in range(8)]
    boxes = orc.cluster(diag, 0.8, 1)
    _cover_disjoint(diag, boxes)
    assert len(diag) / _cells(boxes) >= 0.8
    boxes = orc.cluster(diag, 0.8, 2)
    _cover_disjoint(diag, boxes)
    for lo, hi in boxes:
        inside = sum(1 for f in diag if lo[0] <= f[0] < hi[0] and lo[1] <= f[1] < hi[1])
        assert inside / ((hi[0] - lo[0]) * (hi[1] - lo[1])) >= 0.5


def test_cluster_output_is_independent_of_input_order(orc):
    # cluster.rs:452-478
    flags = art(["xx....xxx", "xx..x.xxx", "....x.xxx"])
    a = list(flags)
    b = list(flags)
    n = len(b)
    for i in range(n - 1, 0, -1):
        j = orc.splitmix64(orc.mix_key(9, [i])) % (i + 1)
        b[i], b[j] = b[j], b[i]
    assert sorted(orc.cluster(a, 0.8, 1)) == sorted(orc.cluster(b, 0.8, 1))


# ------------------------------------------------------------------ regrid
def gaussian(c):
    dx, dy = c[0] - 4.0, c[1] - 4.0
    return math.exp(-H._fma(dx, dx, dy * dy) / 0.4)


def _policy(orc, every=4, buffer=2, eff=0.7, minw=2, thr=0.05, field=0):
    return orc.RegridPolicy(every, buffer, eff, minw, thr, field)


def test_initial_regrid_refines_around_the_feature(orc):
    # tests/amr.rs:599-677
    base = base_grid(orc, [16, 16], [8, 8], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [])
    sim = orc.AmrSim(g, "heat", kappa=0.7, policy=_policy(orc))
    sim.grid.fill_by_center(sim.field(0), gaussian)
    dt = 0.2 * 0.25 * 0.25 / 0.7
    sim.step(dt)
    gr = sim.grid
    assert gr.num_levels == 2
    for cy in range(16):
        for cx in range(16):
            c = (H._fma(0.5, cx, 0.25), H._fma(0.5, cy, 0.25))
            gx = (gaussian((c[0] + 0.5, c[1])) - gaussian((c[0] - 0.5, c[1]))) / 2.0
            gy = (gaussian((c[0], c[1] + 0.5)) - gaussian((c[0], c[1] - 0.5))) / 2.0
            if max(abs(gx), abs(gy)) > 0.06:
                assert gr.find_patch(1, (cx * 2, cy * 2)) is not None
    sim.step(dt, 8)
    assert sim.grid.num_levels == 2
    f = sim.field(0)
    for b in range(sim.grid.num_blocks):
        v = sim.grid.view(f, b)[1:-1, 1:-1]
        assert np.all(np.isfinite(v)) and v.min() >= -1e-9 and v.max() <= 1.0


def test_adaptive_beats_coarse_against_fine_reference_golden(orc):
    # tests/amr.rs:682-788 — the reference's golden numbers.
    kappa = 0.7
    dt = 0.2 * 0.25 * 0.25 / kappa
    steps = 60
    fine = _uniform_heat(orc, gaussian, 32, [16, 16], 0.25, kappa, dt, steps)
    ref = np.zeros((16, 16))
    xs = (np.arange(32) + 0.5) * 0.25
    for y in range(32):
        for x in range(32):
            cx, cy = int(xs[x] / 0.5), int(xs[y] / 0.5)
            ref[cy, cx] = H._fma(fine[y, x], 0.25, ref[cy, cx])
    coarse = _uniform_heat(orc, gaussian, 16, [8, 8], 0.5, kappa, dt, steps)
    base = base_grid(orc, [16, 16], [8, 8], 0.5)
    g = orc.AmrGrid.from_patches(base, [2], [])
    sim = orc.AmrSim(g, "heat", kappa=kappa, policy=_policy(orc))
    sim.grid.fill_by_center(sim.field(0), gaussian)
    refined = 0
    for _ in range(steps):
        sim.step(dt)
        gr = sim.grid
        cells = 0
        for b in gr.blocks_at(1):
            _, lo, hi = gr.patch(b)
            cells += (hi[0] - lo[0]) * (hi[1] - lo[1])
        refined = max(refined, cells)
    gr = sim.grid
    f = sim.field(0)
    gr.restrict(f)
    amr = _gather_level(gr, f, 0, 16)
    err_adaptive = float(np.max(np.abs(amr - ref)))
    err_coarse = float(np.max(np.abs(coarse - ref)))
    assert abs(err_adaptive - 0.001680900) <= 1e-6, err_adaptive
    assert abs(err_coarse - 0.003716000) <= 1e-6, err_coarse
    assert refined == 400


def test_model_c_dendrite_interface_is_tracked_by_refinement(orc):
    # tests/amr.rs:917-988
    base = base_grid(orc, [48, 48], [24, 24], 0.4)
    g = orc.AmrGrid.from_patches(base, [2], [])
    p = orc.params(0.06, 3.19, 0.6267 * 3.19, 0.0)
    sim = orc.AmrSim(g, "modelc", params=p, seed=7, policy=_policy(orc, 4, 2, 0.8, 4, 0.15, 0))
    # ModelC::initialize on the base grid: one seed at (0.4, 0.4), r = 4, u0 = 0.7
    ob = base_grid(orc, [48, 48], [24, 24], 0.4)
    phi = np.zeros(orc.field_len(ob, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(ob, phi, u, np.array([[0.4, 0.4, 4.0]]), 0.7)
    sim.field(0)[:] = phi
    sim.field(1)[:] = u
    # Integrator::suggested_dt for a global-dt scheme: the finest level's
    # stable dt (src/integrators/mod.rs:87-92), h = 0.4 / 2.
    dt = 0.2 * 0.2**2 / p.d_thermal
    sim.step(dt, 150)
    gr = sim.grid
    assert gr.num_levels == 2
    refined = sum((hi[0] - lo[0]) * (hi[1] - lo[1]) for _, lo, hi in (gr.patch(b) for b in gr.blocks_at(1)))
    assert 0 < refined < 96 * 96
    f = sim.field(0)
    for b in range(gr.num_blocks):
        lev, lo, hi = gr.patch(b)
        v = gr.view(f, b)
        inner = v[1:-1, 1:-1]
        assert np.all(np.isfinite(inner)) and np.max(np.abs(inner)) < 1.2
        if lev == 0:
            for j in range(hi[1] - lo[1]):
                for i in range(hi[0] - lo[0]):
                    gx = 0.5 * (v[j + 1, i + 2] - v[j + 1, i])
                    gy = 0.5 * (v[j + 2, i + 1] - v[j, i + 1])
                    if max(abs(gx), abs(gy)) > 0.2:
                        assert gr.find_patch(1, ((lo[0] + i) * 2, (lo[1] + j) * 2)) is not None


def test_model_c_subcycled_dendrite_is_physical(orc):
    # tests/amr.rs:1181-1228 (subcycled adaptive Model C stays physical)
    base = base_grid(orc, [48, 48], [24, 24], 0.4)
    g = orc.AmrGrid.from_patches(base, [2], [])
    p = orc.params(0.06, 3.19, 0.6267 * 3.19, 0.0)
    sim = orc.AmrSim(g, "modelc", params=p, seed=7, integ="subcycle",
                     policy=_policy(orc, 4, 2, 0.8, 4, 0.15, 0))
    ob = base_grid(orc, [48, 48], [24, 24], 0.4)
    phi = np.zeros(orc.field_len(ob, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(ob, phi, u, np.array([[0.4, 0.4, 4.0]]), 0.7)
    sim.field(0)[:] = phi
    sim.field(1)[:] = u
    dt = 0.2 * 0.4**2 / p.d_thermal
    sim.step(dt, 40)
    gr = sim.grid
    assert gr.num_levels == 2
    f0, f1 = sim.field(0), sim.field(1)
    for b in range(gr.num_blocks):
        ph = gr.view(f0, b)[1:-1, 1:-1]
        uu = gr.view(f1, b)[1:-1, 1:-1]
        assert np.all(np.isfinite(ph)) and np.max(np.abs(ph)) < 1.2
        assert np.all(uu >= -0.75) and np.all(uu < 0.05)
