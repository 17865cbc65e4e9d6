"""The binomial checkpointing schedule of ModelCCalibration (host arithmetic in the
HIP library, no device work): its forward-step count equals an independent
dynamic-programming optimum of the same reversal recursion, and the limiting
cases are the textbook ones (one slot: quadratic recomputation; a slot per
step: the forward sweep alone)."""
from __future__ import annotations

import functools

import pytest

lg = pytest.importorskip("legendre_amd")


@functools.lru_cache(maxsize=None)
def T(n: int, s: int) -> int:
    """Least forward steps to reverse n steps from a stored state with s free slots."""
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(m + T(n - m, s - 1) + T(m, s) for m in range(1, n))


def split(n: int, s: int) -> int:
    best, cost = 1, None
    for m in range(1, n):
        c = m + T(n - m, s - 1) + T(m, s)
        if cost is None or c < cost:
            best, cost = m, c
    return best


def total_forward(n: int, snapshots: int) -> int:
    snapshots = min(snapshots, n)
    chain, rem, free = 0, n, snapshots - 1
    while rem >= 2 and free >= 1:
        m = split(rem, free)
        chain += m
        rem -= m
        free -= 1
    return n + T(n, snapshots - 1) - chain


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 21, 34, 55])
@pytest.mark.parametrize("snapshots", [1, 2, 3, 4, 6])
def test_schedule_cost_is_the_dp_optimum(n, snapshots):
    assert lg.checkpoint_forward_steps(n, snapshots) == total_forward(n, snapshots)


def test_limits():
    for n in [1, 5, 40]:
        assert lg.checkpoint_forward_steps(n, 1) == n + n * (n - 1) // 2
        assert lg.checkpoint_forward_steps(n, n) == n
        assert lg.checkpoint_forward_steps(n, n + 7) == n


def test_calibration_config_cost():
    """2000 steps in 45 slots: about 2.4 forward sweeps (binomial repetition number 3)."""
    c = lg.checkpoint_forward_steps(2000, 45)
    assert 2000 < c < 3 * 2000
