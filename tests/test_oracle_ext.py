"""The oracle's multi-grain Model C and RungeKutta4, pinned without retyping them:
the oriented vector field reduces bitwise to the axis-aligned one at θ₀ = 0 and
to rounding at θ₀ = π/2 (the 4-fold symmetry); RK4 follows the closed-form
eigenmode decay of the reference's heat test (tests/heat.rs:145-168); the
orientation pass equals a brute-force nearest seed; and a 45°-rotated grain grows
along the diagonals (tests/model_c.rs:186-280)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def orc():
    oracle.lib()
    return oracle


def state(orc, cells, block, grains, u0=0.7, h=0.4):
    g = orc.grid(cells, block, [0, 0], [h, h])
    n = orc.field_len(g, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(g, phi, u, np.array(grains)[:, :3], u0)
    orc.fill_ghosts(g, phi, 1)
    orc.fill_ghosts(g, u, 1)
    return g, phi, u


def test_oriented_field_at_zero_angle_is_the_axis_aligned_field_bitwise(orc):
    g, phi, u = state(orc, [48, 40], [24, 20], [[5.0, 4.0, 2.5, 0.0], [13.0, 11.0, 1.5, 0.0]])
    p = orc.params(0.06, 3.19, 2.0)
    n = phi.size
    k1, k2 = np.zeros(n), np.zeros(n)
    orc.modelc_rhs(g, p, phi, u, k1, k2)
    theta = np.zeros(orc.field_len(g, 0))
    o1, o2 = np.zeros(n), np.zeros(n)
    orc.modelc_rhs_oriented(g, p, phi, u, theta, o1, o2)
    assert np.array_equal(o1, k1) and np.array_equal(o2, k2)


def test_quarter_turn_orientation_is_the_four_fold_identity(orc):
    """A(θ − θ₀) is π/2-periodic in θ₀: θ₀ = π/2 reproduces θ₀ = 0 up to the rounding
    of cos(π/2); θ₀ = π/4 does not (it swaps the fast and slow directions)."""
    g, phi, u = state(orc, [48, 40], [24, 20], [[5.0, 4.0, 2.5, 0.0], [13.0, 11.0, 1.5, 0.0]])
    p = orc.params(0.06, 3.19, 2.0)
    n = phi.size
    base = np.zeros(n), np.zeros(n)
    orc.modelc_rhs_oriented(g, p, phi, u, np.zeros(orc.field_len(g, 0)), *base)
    quarter = np.zeros(n), np.zeros(n)
    orc.modelc_rhs_oriented(g, p, phi, u, np.full(orc.field_len(g, 0), math.pi / 2), *quarter)
    eighth = np.zeros(n), np.zeros(n)
    orc.modelc_rhs_oriented(g, p, phi, u, np.full(orc.field_len(g, 0), math.pi / 4), *eighth)
    scale = np.max(np.abs(base[0]))
    assert np.max(np.abs(quarter[0] - base[0])) <= 1e-12 * scale
    assert np.max(np.abs(eighth[0] - base[0])) > 1e-3 * scale


def test_orientation_pass_equals_brute_force_nearest_seed(orc):
    rng = np.random.default_rng(4)
    g = orc.grid([40, 36], [10, 12], [0, 0], [0.5, 0.5])
    grains = np.column_stack([rng.uniform(0, 20, 9), rng.uniform(0, 18, 9), np.full(9, 1.0),
                              rng.uniform(-1, 1, 9)])
    theta = np.zeros(orc.field_len(g, 0))
    orc.initialize_orientations(g, theta, grains)
    nb, (nx, ny) = orc.num_blocks(g), (10, 12)
    want = np.zeros_like(theta)
    for b in range(nb):
        for j in range(ny):
            for i in range(nx):
                x, y = orc.cell_center(g, b, (i, j))[:2]
                d = [math.fma((y - gy), (y - gy), (x - gx) * (x - gx)) if hasattr(math, "fma")
                     else (y - gy) ** 2 + (x - gx) ** 2 for gx, gy in grains[:, :2]]
                want[b * nx * ny + j * nx + i] = grains[int(np.argmin(d)), 3]
    assert np.array_equal(theta, want)


def test_rk4_follows_the_discrete_eigenmode_decay(orc):
    """tests/heat.rs:145-168 on the oracle: u = cos(kx·x) on a 32² Neumann grid decays by
    R(z)^n with R the RK4 amplification 1 − z + z²/2 − z³/6 + z⁴/24."""
    N, h, kappa, steps = 32, 0.1, 0.7, 200
    g = orc.grid([N, N], [N // 4, N // 4], [0, 0], [h, h])
    kx = math.pi / (N * h)
    u = np.zeros(orc.field_len(g, 1))
    nb, nbk = orc.num_blocks(g), N // 4
    view = u.reshape(nb, nbk + 2, nbk + 2)
    xs = []
    for b in range(nb):
        for j in range(nbk):
            for i in range(nbk):
                x = orc.cell_center(g, b, (i, j))[0]
                view[b, j + 1, i + 1] = math.cos(kx * x)
                xs.append((b, i, j, x))
    dt = 0.25 * h * h / kappa
    orc.heat_rk4_run(g, kappa, u, steps, dt)
    lam = math.fma(2.0, -math.cos(kx * h), 2.0) if hasattr(math, "fma") else 2.0 - 2.0 * math.cos(kx * h)
    lam /= h * h
    z = dt * kappa * lam
    factor = (1 - z + z * z / 2 - z ** 3 / 6 + z ** 4 / 24) ** steps
    for b, i, j, x in xs:
        want = factor * math.cos(kx * x)
        assert abs(view[b, j + 1, i + 1] - want) <= max(1e-10 * abs(want), 1e-12)


def test_forward_euler_and_rk4_differ_and_rk4_is_deterministic(orc):
    g, phi, u = state(orc, [32, 32], [16, 16], [[6.4, 6.4, 2.0, 0.0]])
    p = orc.params(0.06, 3.19, 2.0)
    a = phi.copy(), u.copy()
    b = phi.copy(), u.copy()
    c = phi.copy(), u.copy()
    orc.modelc_rk4_run(g, p, *a, 10, 0.008)
    orc.modelc_rk4_run(g, p, *b, 10, 0.008)
    orc.modelc_run(g, p, *c, 10, 0.008)
    assert np.array_equal(a[0], b[0])
    assert not np.array_equal(a[0], c[0])
    assert np.max(np.abs(a[0] - c[0])) < 1e-2  # both approximate the same flow


def _arm_lengths(orc, g, phi, center, nb, n):
    wedge = math.radians(12.0)
    axes = [0.0, math.pi / 2, math.pi, -math.pi / 2]
    diags = [math.pi / 4, 3 * math.pi / 4, -math.pi / 4, -3 * math.pi / 4]
    view = phi.reshape(nb, n + 2, n + 2)
    solid = []
    for b in range(nb):
        for j in range(n):
            for i in range(n):
                if view[b, j + 1, i + 1] > 0.0:
                    x, y = orc.cell_center(g, b, (i, j))[:2]
                    dx, dy = x - center, y - center
                    solid.append((math.atan2(dy, dx), math.hypot(dx, dy)))

    def arm(dirs):
        best = 0.0
        for ang, r in solid:
            for d in dirs:
                diff = abs(ang - d) % (2 * math.pi)
                if diff > math.pi:
                    diff = 2 * math.pi - diff
                if diff < wedge:
                    best = max(best, r)
                    break
        return best

    return arm(axes), arm(diags)


@pytest.mark.parametrize("theta0", [0.0, math.pi / 4])
def test_rotated_grain_grows_along_rotated_axes(orc, theta0):
    """tests/model_c.rs:186-280 on the oracle (96², 700 forward-Euler steps)."""
    M, H = 96, 0.4
    g = orc.grid([M, M], [M // 2, M // 2], [0, 0], [H, H])
    center = M * H / 2
    grains = np.array([[center, center, 10.0 * H, theta0]])
    n = orc.field_len(g, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(g, phi, u, grains[:, :3], 0.7)
    theta = np.zeros(orc.field_len(g, 0))
    orc.initialize_orientations(g, theta, grains)
    eps4, lam = 0.06, 3.19
    p = orc.params(eps4, lam, 0.6267 * lam)
    dt = 0.2 * H * H / p.d_thermal
    orc.modelc_oriented_run(g, p, phi, u, theta, 700, dt)
    axis, diag = _arm_lengths(orc, g, phi, center, orc.num_blocks(g), M // 2)
    if theta0 == 0.0:
        assert axis > 1.02 * diag, (axis, diag)
    else:
        assert diag > 1.02 * axis, (axis, diag)
