"""Pins the oracle's jump driver and multi-driver explicit update
(orc_apply_drivers, orc_jump_diffusion_run) before the device is compared with it.

- The reference's own behavioural tests of Driver::Jump and WienerJump
  (tests/jump.rs:174-281), ported: the thinned-Poisson mean, the exact
  count/jsum lockstep, per-cell intensity, zero intensity, the Wiener
  statistics in the mixed set, and scheduling independence.
- A single-step firing fraction at λ·dt = 1, where Bernoulli(1 − e^{−λ·dt})
  (0.632) is far from λ·dt (1.0) and from the reversed comparison (0.368).
- The key chain of each fire, restated with Python integers from the
  KAT-pinned SplitMix64 (driver.rs:55-61 stream tags, driver.rs:159-182), on a
  ghost-0 and a ghost-1 field moved by the same driver.
- The Time and Wiener arms equal the separately pinned axpy and apply_wiener.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

MU, SIGMA, PAYOFF = 0.1, 0.5, 2.0
M64 = (1 << 64) - 1


def run(orc, lam_of, paths, block, dt, steps, seed, nthreads=1):
    g = orc.grid([paths], [block], [0.0], [1.0])
    n = orc.field_len(g, 0)
    lam = np.array([lam_of(i) for i in range(n)], dtype=np.float64)
    w, count, jsum = np.zeros(n), np.zeros(n), np.zeros(n)
    orc.jump_diffusion_run(g, lam, w, count, jsum, steps, dt, seed, MU, SIGMA, PAYOFF, nthreads=nthreads)
    return w, count, jsum


def thinned_mean(lam, dt, steps):
    return steps * (1.0 - math.exp(-lam * dt))


def test_jump_count_matches_thinned_poisson_mean(orc):
    _, count, _ = run(orc, lambda i: 2.0, 50_000, 5000, 0.01, 100, 1)
    assert abs(count.mean() - thinned_mean(2.0, 0.01, 100)) < 0.03
    assert np.all((count == np.round(count)) & (count >= 0.0))


def test_coupled_fields_move_on_the_same_fire(orc):
    _, count, jsum = run(orc, lambda i: 3.0, 20_000, 5000, 0.01, 100, 5)
    assert np.array_equal(jsum, count * PAYOFF)


def test_intensity_is_read_per_cell(orc):
    n, lo, hi = 40_000, 1.0, 5.0
    _, count, _ = run(orc, lambda i: lo if i < n // 2 else hi, n, 5000, 0.01, 100, 9)
    lo_mean, hi_mean = count[: n // 2].mean(), count[n // 2:].mean()
    assert abs(lo_mean - thinned_mean(lo, 0.01, 100)) < 0.05
    assert abs(hi_mean - thinned_mean(hi, 0.01, 100)) < 0.05
    assert hi_mean > lo_mean


def test_zero_intensity_never_fires(orc):
    _, count, jsum = run(orc, lambda i: 0.0, 10_000, 5000, 0.01, 200, 3)
    assert not count.any() and not jsum.any()


def test_wiener_and_jump_coexist_in_a_mixed_set(orc):
    dt, steps = 0.01, 100
    w, _, _ = run(orc, lambda i: 3.0, 50_000, 5000, dt, steps, 7)
    t = dt * steps
    assert abs(w.mean() - MU * t) < 0.02
    assert abs(w.var() - SIGMA * SIGMA * t) < 0.02


def test_jump_paths_are_bitwise_scheduler_independent(orc):
    a = run(orc, lambda i: 2.5, 16_000, 4000, 0.01, 60, 42, nthreads=1)
    b = run(orc, lambda i: 2.5, 16_000, 4000, 0.01, 60, 42, nthreads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_one_step_fires_with_probability_one_minus_exp_of_minus_rate_dt(orc):
    n, lam, dt = 200_000, 20.0, 0.05
    _, count, _ = run(orc, lambda i: lam, n, 10_000, dt, 1, 11)
    p = 1.0 - math.exp(-lam * dt)
    se = math.sqrt(p * (1 - p) / n)
    assert abs(count.mean() - p) < 5 * se, (count.mean(), p)


def _py_fire_set(orc, g, nblocks, ncell, lam, dt, seed, salt, j):
    """Fires of Jump(j) restated from driver.rs:55-61, 159-182 with Python integers:
    stream = (2 << 32) | j; block key = mix_key(seed, [salt, stream, block]);
    fire where λ > 0 and unit_open(splitmix64(block_key ^ cell)) < −expm1(−λ·dt)."""
    fired = np.zeros(nblocks * ncell, dtype=bool)
    stream = (2 << 32) | j
    for b in range(nblocks):
        bk = orc.mix_key(seed, [salt, stream, b])
        for c in range(ncell):
            lmb = lam[b * ncell + c]
            if not lmb > 0.0:
                continue
            key = orc.splitmix64((bk ^ c) & M64)
            fired[b * ncell + c] = orc.unit_open(key) < -math.expm1(-lmb * dt)
    return fired


def test_fires_follow_the_key_chain_on_fields_of_two_ghost_widths(orc):
    cells, block = [12, 10], [6, 5]
    g = orc.grid(cells, block, [0.0, 0.0], [1.0, 1.0])
    nb, ncell = orc.num_blocks(g), 30
    rng = np.random.default_rng(3)
    lam = rng.uniform(-2.0, 30.0, nb * ncell)
    lam[::17] = 0.0
    lam[5] = np.nan
    dt, seed, salt, j = 0.05, 77, orc.time_bits(0.35), 1
    x0 = np.zeros(orc.field_len(g, 0))
    x1 = rng.standard_normal(orc.field_len(g, 1))
    inc0 = rng.uniform(0.5, 1.5, x0.size)
    inc1 = rng.uniform(0.5, 1.5, x1.size)
    y0, y1 = x0.copy(), x1.copy()
    orc.apply_drivers(g, [y0, y1], [0, 1], [(orc.DRIVER_JUMP, j, [inc0, inc1], lam)], dt, seed, salt)
    fired = _py_fire_set(orc, g, nb, ncell, lam, dt, seed, salt, j)
    assert 0 < fired.sum() < fired.size
    assert np.array_equal(y0, np.where(fired, x0 + inc0, x0))
    # The ghost-1 field moves on the same fires at its interior cells; ghosts never move.
    want1 = x1.copy()
    for off in range(x1.size):
        b, r = divmod(off, orc.block_len(g, 1))
        c = orc.cell_key(g, 1, r)
        if c is not None and fired[b * ncell + c]:
            want1[off] = x1[off] + inc1[off]
    assert np.array_equal(y1, want1)


def test_time_and_wiener_arms_equal_axpy_and_apply_wiener(orc):
    g = orc.grid([16, 12], [8, 6], [0.0, 0.0], [0.5, 0.5])
    rng = np.random.default_rng(8)
    n = orc.field_len(g, 1)
    x, k, a0, a1 = (rng.standard_normal(n) for _ in range(4))
    a1[::3] = 0.0
    dt, seed, salt = 0.02, 5, orc.time_bits(1.25)
    got = x.copy()
    orc.apply_drivers(g, [got], [1], [(orc.DRIVER_TIME, 0, [k], None),
                                      (orc.DRIVER_WIENER, 0, [a0], None),
                                      (orc.DRIVER_WIENER, 1, [a1], None)], dt, seed, salt)
    want = x.copy()
    orc.axpy(want, k, dt)
    orc.apply_wiener(g, want, a0, 1, dt, seed, salt, (1 << 32) | 0)
    orc.apply_wiener(g, want, a1, 1, dt, seed, salt, (1 << 32) | 1)
    assert np.array_equal(got, want)
    # Two Wiener drivers draw independent deviates (distinct stream tags).
    one = x.copy()
    orc.apply_drivers(g, [one], [1], [(orc.DRIVER_WIENER, 0, [a0], None)], dt, seed, salt)
    other = x.copy()
    orc.apply_drivers(g, [other], [1], [(orc.DRIVER_WIENER, 1, [a0], None)], dt, seed, salt)
    assert not np.array_equal(one, other)


def test_a_jump_driver_without_intensity_is_rejected(orc):
    g = orc.grid([8], [4], [0.0], [1.0])
    x = np.zeros(orc.field_len(g, 0))
    with pytest.raises(ValueError):
        orc.apply_drivers(g, [x], [0], [(orc.DRIVER_JUMP, 0, [x.copy()], None)], 0.1, 1, 0)
