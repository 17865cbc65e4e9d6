"""GPU parity of lgd_apply_drivers (State::apply_step_with with Driver::apply_slab's
Time, Wiener and Jump arms, src/core/state.rs:646-699, src/core/driver.rs:115-182)
against the oracle: the jump-diffusion model of tests/jump.rs stepped on the device
(fires and the jump-moved fields bitwise, the Wiener-moved field to the device
log/cos), fields of two ghost widths moved by one jump driver, and the error
contract."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")

MU, SIGMA, PAYOFF = 0.1, 0.5, 2.0


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def buf(ctx, arr):
    b = lg.DeviceBuffer(ctx, arr.size)
    b.upload(np.ascontiguousarray(arr, dtype=np.float64))
    return b


def device_jump_diffusion(ctx, lam, paths, block, dt, steps, seed):
    """tests/jump.rs's JumpDiffusion under EulerMaruyama, through the C ABI: the
    vector fields are constant, so they are written once; each step applies Time,
    Wiener(0) and Jump(0) in one launch with salt = t.to_bits()."""
    grid = lg.CartesianGrid(ctx, [paths], [block], [0.0], [1.0])
    n = grid.field_len(0)
    w, count, jsum = (lg.DeviceBuffer(ctx, n) for _ in range(3))
    drift, amp = buf(ctx, np.full(n, MU)), buf(ctx, np.full(n, SIGMA))
    one, pay = buf(ctx, np.ones(n)), buf(ctx, np.full(n, PAYOFF))
    rate = buf(ctx, lam)
    drivers = [lg.Driver(lg.LGD_DRIVER_TIME, 0, [drift, None, None]),
               lg.Driver(lg.LGD_DRIVER_WIENER, 0, [amp, None, None]),
               lg.Driver(lg.LGD_DRIVER_JUMP, 0, [None, one, pay], rate)]
    t = 0.0
    for _ in range(steps):
        grid.apply_drivers([w, count, jsum], [0, 0, 0], drivers, dt, seed, lg.time_bits(t))
        t = t + dt
    out = tuple(b.download() for b in (w, count, jsum))
    grid.close()
    return out


@pytest.mark.parametrize("lam_kind", ["uniform", "per_cell"])
def test_jump_diffusion_matches_oracle(ctx, orc, lam_kind):
    paths, block, dt, steps, seed = 50_000, 5000, 0.01, 100, 9
    say(f"jump diffusion {lam_kind} paths={paths} steps={steps}")
    if lam_kind == "uniform":
        lam = np.full(paths, 2.0)
    else:
        rng = np.random.default_rng(2)
        lam = rng.uniform(-1.0, 6.0, paths)
        lam[::7] = 0.0
    w, count, jsum = device_jump_diffusion(ctx, lam, paths, block, dt, steps, seed)
    og = orc.grid([paths], [block], [0.0], [1.0])
    ww, wc, wj = np.zeros(paths), np.zeros(paths), np.zeros(paths)
    orc.jump_diffusion_run(og, lam, ww, wc, wj, steps, dt, seed, MU, SIGMA, PAYOFF)
    assert np.array_equal(count, wc)
    assert np.array_equal(jsum, wj)
    assert np.array_equal(jsum, count * PAYOFF)
    # w carries 100 Box–Muller deviates: the device log/cos against glibc's.
    assert np.max(np.abs(w - ww)) <= 1e-12
    if lam_kind == "uniform":
        assert abs(count.mean() - steps * (1 - np.exp(-2.0 * dt))) < 0.03


def test_one_jump_driver_moves_fields_of_two_ghost_widths(ctx, orc):
    cells, block = [48, 40], [12, 10]
    say(f"jump on ghost-0 and ghost-1 fields cells={cells}")
    og = orc.grid(cells, block, [0.0, 0.0], [0.5, 0.5])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.5, 0.5])
    rng = np.random.default_rng(6)
    n0, n1 = orc.field_len(og, 0), orc.field_len(og, 1)
    lam = rng.uniform(-3.0, 25.0, n0)
    lam[::13] = 0.0
    lam[4] = np.nan
    s0, s1, s2 = rng.standard_normal(n0), rng.standard_normal(n1), rng.standard_normal(n1)
    inc0, inc1, k2 = rng.uniform(0.5, 1.5, n0), rng.uniform(0.5, 1.5, n1), rng.standard_normal(n1)
    dt, seed, times = 0.03, 1234, [0.0, 0.03, 0.06]
    x0, x1, x2 = s0.copy(), s1.copy(), s2.copy()
    for t in times:
        orc.apply_drivers(og, [x0, x1, x2], [0, 1, 1],
                          [(orc.DRIVER_TIME, 0, [None, None, k2], None),
                           (orc.DRIVER_JUMP, 2, [inc0, inc1, None], lam)], dt, seed, lg.time_bits(t))
    xs = [buf(ctx, s0), buf(ctx, s1), buf(ctx, s2)]
    drivers = [lg.Driver(lg.LGD_DRIVER_TIME, 0, [None, None, buf(ctx, k2)]),
               lg.Driver(lg.LGD_DRIVER_JUMP, 2, [buf(ctx, inc0), buf(ctx, inc1), None], buf(ctx, lam))]
    for t in times:
        grid.apply_drivers(xs, [0, 1, 1], drivers, dt, seed, lg.time_bits(t))
    assert np.array_equal(xs[0].download(), x0)
    assert np.array_equal(xs[1].download(), x1)
    assert np.array_equal(xs[2].download(), x2)
    assert not np.array_equal(x0, s0) and not np.array_equal(x1, s1)  # something fired
    grid.close()


def test_time_and_wiener_drivers_match_the_oracle(ctx, orc):
    cells, block = [32, 24], [16, 12]
    say("time + two wiener drivers, ghost 1")
    og = orc.grid(cells, block, [0.0, 0.0], [0.5, 0.5])
    grid = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.5, 0.5])
    rng = np.random.default_rng(12)
    n = orc.field_len(og, 1)
    x, y, k, a0, a1 = (rng.standard_normal(n) for _ in range(5))
    a1[::4] = 0.0
    dt, seed, salt = 0.02, 3, lg.time_bits(0.7)
    want_x, want_y = x.copy(), y.copy()
    orc.apply_drivers(og, [want_x, want_y], [1, 1], [(orc.DRIVER_TIME, 0, [k, None], None),
                                                      (orc.DRIVER_WIENER, 0, [a0, a0], None),
                                                      (orc.DRIVER_WIENER, 1, [None, a1], None)],
                      dt, seed, salt)
    bx, by = buf(ctx, x), buf(ctx, y)
    ba0 = buf(ctx, a0)
    grid.apply_drivers([bx, by], [1, 1], [lg.Driver(lg.LGD_DRIVER_TIME, 0, [buf(ctx, k), None]),
                                          lg.Driver(lg.LGD_DRIVER_WIENER, 0, [ba0, ba0]),
                                          lg.Driver(lg.LGD_DRIVER_WIENER, 1, [None, buf(ctx, a1)])],
                       dt, seed, salt)
    gx, gy = bx.download(), by.download()
    scale = np.max(np.abs(want_x))
    assert np.max(np.abs(gx - want_x)) <= 1e-13 * scale
    assert np.max(np.abs(gy - want_y)) <= 1e-13 * scale
    # Ghost entries move only under Time: the y field (no Time driver) keeps its ghosts.
    ghost = np.array([orc.cell_key(og, 1, i % orc.block_len(og, 1)) is None for i in range(n)])
    assert np.array_equal(gy[ghost], y[ghost])
    grid.close()


def test_driver_contract_errors(ctx, orc):
    say("apply_drivers error contract")
    grid = lg.CartesianGrid(ctx, [16], [8], [0.0], [1.0])
    n = grid.field_len(0)
    x, v = lg.DeviceBuffer(ctx, n), lg.DeviceBuffer(ctx, n)
    with pytest.raises(lg.LegendreError) as e:
        grid.apply_drivers([x], [0], [lg.Driver(lg.LGD_DRIVER_JUMP, 0, [v])], 0.1, 1, 0)
    assert e.value.status == 1
    with pytest.raises(lg.LegendreError) as e:
        grid.apply_drivers([x] * 9, [0] * 9, [lg.Driver(lg.LGD_DRIVER_TIME, 0, [v] * 9)], 0.1, 1, 0)
    assert e.value.status == 5
    with pytest.raises(lg.LegendreError) as e:
        grid.apply_drivers([x], [0], [lg.Driver(9, 0, [v])], 0.1, 1, 0)
    assert e.value.status == 1
    short = lg.DeviceBuffer(ctx, n - 1)
    with pytest.raises(lg.LegendreError) as e:
        grid.apply_drivers([x], [0], [lg.Driver(lg.LGD_DRIVER_TIME, 0, [short])], 0.1, 1, 0)
    assert e.value.status == 3
    grid.close()
