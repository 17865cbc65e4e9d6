"""Pins the oracle's three-level regrid and the reference AMR tests that pin the
transfers (tests/amr.rs), on the CPU.

Function → pin:
  nesting flags (adapt.rs:194-211)   hand-computed boxes for ratios 2/2, 3/2, 4/3 and a
                                     box at the domain origin; the least-cover property
                                     over random aligned boxes
  3-level regrid                     a case whose level-1 boxes come from the nesting
                                     flags alone (exact boxes, derived in helpers.py);
                                     proper nesting after every step of 3-level adaptive
                                     Model C and heat runs
  migrate / sample (adapt.rs:323-402) amr.rs:846-900: a forced box moved mid-run under a
                                     frozen model keeps a linear field to 1e-12
  global-dt stack on an AmrGrid      amr.rs:192: no patches ≡ the Cartesian run, bitwise
  interface run                      amr.rs:491-537: finite, maximum principle, decay
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import helpers as H


# ------------------------------------------------------------- nesting flags
@pytest.mark.parametrize("box, rf, rs, want", [
    # [48, 80) × [46, 74) at ratio 2 → need [23, 41) × [22, 38) → parents at ratio 2.
    (((48, 46), (80, 74)), 2, 2, ((11, 11), (21, 19))),
    # At the origin the grown box reaches −1: floor division keeps −1 (clipped later).
    (((0, 0), (8, 4)), 2, 2, ((-1, -1), (3, 2))),
    # rf = 3: need [2, 8) × [1, 5); rs = 2: floor(2/2) = 1, ceil(8/2) = 4, 0, ceil(5/2) = 3.
    (((9, 6), (21, 12)), 3, 2, ((1, 0), (4, 3))),
    # rf = 4: need [1, 5) × [0, 4); rs = 3: 0, ceil(5/3) = 2, 0, ceil(4/3) = 2.
    (((8, 4), (16, 12)), 4, 3, ((0, 0), (2, 2))),
])
def test_nesting_parents_known_answers(orc, box, rf, rs, want):
    assert orc.nesting_parents(box, rf, rs) == want


def test_nesting_parents_is_the_least_cover(orc):
    """For boxes aligned to rf (as the regrid builds them), the parents are the least
    box whose cells, refined by rs, cover the box coarsened by rf and grown by one."""
    rng = np.random.default_rng(3)
    for _ in range(400):
        rf, rs = int(rng.integers(2, 5)), int(rng.integers(2, 5))
        lo = rng.integers(0, 12, 2) * rf
        hi = lo + rng.integers(1, 7, 2) * rf
        (plo, phi) = orc.nesting_parents(((int(lo[0]), int(lo[1])), (int(hi[0]), int(hi[1]))), rf, rs)
        for d in range(2):
            need_lo, need_hi = lo[d] // rf - 1, hi[d] // rf + 1
            assert plo[d] * rs <= need_lo < (plo[d] + 1) * rs, (rf, rs, lo, hi)
            assert (phi[d] - 1) * rs < need_hi <= phi[d] * rs, (rf, rs, lo, hi)


# ------------------------------------------------------------ 3-level regrid
def _patches(g):
    return sorted(g.patch(b) for b in range(g.num_blocks) if g.patch(b)[0] > 0)


def stripes_sim(orc, model="heat"):
    S = H.STRIPES
    base = orc.grid(S["cells"], S["block"], [0, 0], [S["h"], S["h"]])
    g = orc.AmrGrid.from_patches(base, S["ratios"], S["levels"])
    f = H.stripes_field(g, g.field_len(1))
    sim = orc.AmrSim(g, "heat", kappa=0.05, policy=orc.RegridPolicy(*S["policy"]))
    sim.field(0)[:] = f
    return sim


def test_three_level_regrid_builds_level_one_from_nesting_flags(orc):
    """The level-2 proposal needs a level-1 box that no level-0 tag asks for: the
    regrid must build it from the nesting flags alone, exactly."""
    sim = stripes_sim(orc)
    sim.step(0.01)
    assert _patches(sim.grid) == H.STRIPES["expect"]
    assert H.proper_nesting_violations(sim.grid) == []


def _three_level_modelc(orc, seeds, subcycle, noise=0.0, cells=64, block=32, steps=24):
    h = 0.4
    base = orc.grid([cells, cells], [block, block], [0, 0], [h, h])
    p = orc.params(0.05, 3.19, 2.0, noise)
    pol = orc.RegridPolicy(4, 2, 0.8, 4, 0.1, 0)
    sim = orc.AmrSim(orc.AmrGrid.from_patches(base, [2, 2], []), "modelc", params=p, seed=5,
                     wiener=noise > 0, integ="subcycle" if subcycle else "global", policy=pol)
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, seeds, 0.55)
    sim.field(0)[:] = phi
    sim.field(1)[:] = u
    return sim


@pytest.mark.parametrize("subcycle", [False, True])
def test_three_level_adaptive_model_c_keeps_proper_nesting(orc, subcycle):
    seeds = np.array([[8.0, 8.0, 2.0], [17.0, 9.0, 2.0], [12.0, 18.0, 1.6]])
    sim = _three_level_modelc(orc, seeds, subcycle)
    h = 0.4
    dt = 0.2 * (h if subcycle else h / 4) ** 2 / 2.0
    deepest = 0
    for s in range(16):
        sim.step(dt)
        g = sim.grid
        deepest = max(deepest, g.num_levels)
        assert H.proper_nesting_violations(g) == [], f"step {s}"
    assert deepest == 3
    f0 = sim.field(0)
    for b in range(sim.grid.num_blocks):
        v = sim.grid.view(f0, b)[1:-1, 1:-1]
        assert np.all(np.isfinite(v)) and np.max(np.abs(v)) < 1.2


# -------------------------------------------------- ported reference tests
def lin(c):
    return H._fma(3.0, c[0], 2.0) - c[1]


def test_migration_preserves_linear_fields_exactly(orc):
    """tests/amr.rs:846-900: a box tagger forces [2,6)² then [4,8)² under a frozen
    model (heat with κ = 0: every tendency is 0·∇²u = 0); the moved patch holds the
    linear field through both migration paths (same-level copy, prolongation)."""
    base = orc.grid([16, 16], [8, 8], [0, 0], [0.5, 0.5])
    g = orc.AmrGrid.from_patches(base, [2], [])
    sim = orc.AmrSim(g, "heat", kappa=0.0, policy=orc.RegridPolicy(1, 0, 0.9, 2, 0.0, 0))
    calls = []

    def box_tagger(grid, level):
        if level != 0:
            return []
        shift = 2 if calls else 0
        calls.append(level)
        return [(x, y) for y in range(2 + shift, 6 + shift) for x in range(2 + shift, 6 + shift)]

    sim.set_tagger(box_tagger)
    sim.grid.fill_by_center(sim.field(0), lin)
    sim.step(1e-3)
    assert _patches(sim.grid) == [(1, (4, 4), (12, 12))]
    sim.step(1e-3)
    gr = sim.grid
    assert _patches(gr) == [(1, (8, 8), (16, 16))]
    f = sim.field(0)
    for b in gr.blocks_at(1):
        v = gr.view(f, b)
        for j in range(8):
            for i in range(8):
                want = lin(gr.cell_center(b, i, j))
                assert abs(v[j + 1, i + 1] - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("model", ["heat", "modelc"])
def test_simulation_on_amr_matches_cartesian_bitwise(orc, model):
    """tests/amr.rs:192: the Simulation stack on an AmrGrid without patches equals
    the Cartesian run bit for bit (global dt; Model C with noise, whose keys are the
    same block ids and cell indices)."""
    cells, block, h = [24, 16], [8, 8], 0.4
    base = orc.grid(cells, block, [0, 0], [h, h])
    if model == "heat":
        init = lambda c: math.cos(0.7 * c[0]) + math.sin(1.3 * c[1])
        sim = orc.AmrSim(orc.AmrGrid.from_patches(base, [], []), "heat", kappa=0.7)
        sim.grid.fill_by_center(sim.field(0), init)
        u = sim.field(0).copy()
        dt = 0.2 * h * h / 0.7
        sim.step(dt, 40)
        orc.heat_run(base, 0.7, u, 40, dt)
        m = np.zeros(u.size, bool)
        H.field_view(m, cells, block, 1)[(slice(None),) + H.interior_slices(2, 1)] = True
        assert np.array_equal(sim.field(0)[m], u[m])
        return
    p = orc.params(0.05, 3.19, 2.0, 0.02)
    sim = orc.AmrSim(orc.AmrGrid.from_patches(base, [], []), "modelc", params=p, seed=9, wiener=True)
    phi = np.zeros(orc.field_len(base, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(base, phi, u, np.array([[4.0, 3.0, 2.0], [7.0, 4.0, 1.2]]), 0.55)
    sim.field(0)[:] = phi
    sim.field(1)[:] = u
    sim.step(0.008, 30)
    orc.modelc_run(base, p, phi, u, 30, 0.008, seed=9, wiener=True)
    m = np.zeros(phi.size, bool)
    H.field_view(m, cells, block, 1)[(slice(None),) + H.interior_slices(2, 1)] = True
    assert np.array_equal(sim.field(0)[m], phi[m]) and np.array_equal(sim.field(1)[m], u[m])


def interface_gaussian(c):
    dx, dy = c[0] - 4.0, c[1] - 4.0
    return math.exp(-H._fma(dx, dx, dy * dy) / 1.5)


def test_interface_run_stays_physical(orc):
    """tests/amr.rs:491-537: one fine patch over a Gaussian bump, 200 global steps."""
    kappa, hf = 0.7, 0.25
    dt = 0.2 * hf * hf / kappa
    base = orc.grid([16, 16], [8, 8], [0, 0], [0.5, 0.5])
    sim = orc.AmrSim(orc.AmrGrid.from_patches(base, [2], [[((8, 8), (24, 24))]]), "heat", kappa=kappa)
    sim.grid.fill_by_center(sim.field(0), interface_gaussian)
    sim.step(dt, 200)
    lo, hi = math.inf, -math.inf
    f = sim.field(0)
    for b in range(sim.grid.num_blocks):
        v = sim.grid.view(f, b)[1:-1, 1:-1]
        assert np.all(np.isfinite(v))
        lo, hi = min(lo, v.min()), max(hi, v.max())
    assert hi < 0.9, hi
    assert hi > 0.0 and lo > -1e-9, (lo, hi)
