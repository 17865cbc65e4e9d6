"""The reference's integrator convergence tests (tests/convergence.rs:86-107) on the
device. The semi-discrete heat equation's eigenmode u₀ = cos(k·x) on a mirror-faced
grid evolves exactly as exp(−κ·λ_h·t)·u₀ in time (λ_h the discrete Laplacian's
eigenvalue), so halving dt must shrink the time-integration error by 2 (forward
Euler) or 16 (RK4). Euler–Maruyama with a zero-amplitude Wiener driver is forward
Euler bit for bit."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


# Mode 16 of 64 cells: λ_h·dt large enough for RK4's error to stand above rounding,
# with both schemes inside their stability limits for every mode of the grid.
N, B, H, KAPPA, T_END, MODE = 64, 16, 0.1, 1.0, 0.04, 16


def heat_error(ctx, integrator: str, dt: float) -> float:
    grid = lg.CartesianGrid(ctx, [N], [B], [0.0], [H])
    k = MODE * math.pi / (N * H)
    lam = (2.0 - 2.0 * math.cos(k * H)) / (H * H)
    u0 = np.zeros(grid.field_len(1))
    view = u0.reshape(N // B, B + 2)
    x = (np.arange(N) + 0.5) * H
    view[:, 1:-1] = np.cos(k * x).reshape(N // B, B)
    sim = lg.Simulation.heat(grid, KAPPA)
    sim.set_integrator(integrator)
    sim.field(0).upload(u0)
    steps = int(round(T_END / dt))
    sim.step(dt, steps)
    got = sim.field(0).download().reshape(N // B, B + 2)[:, 1:-1].reshape(-1)
    want = math.exp(-KAPPA * lam * steps * dt) * np.cos(k * x)
    sim.close()
    grid.close()
    return float(np.max(np.abs(got - want)))


def observed_order(ctx, integrator: str, dt: float) -> float:
    return math.log2(heat_error(ctx, integrator, dt) / heat_error(ctx, integrator, dt / 2))


def test_forward_euler_is_first_order(ctx):
    say("forward Euler order")
    order = observed_order(ctx, "euler", 0.001)
    assert abs(order - 1.0) < 0.1, order


def test_rk4_is_fourth_order(ctx):
    say("RK4 order")
    order = observed_order(ctx, "rk4", 0.004)
    assert abs(order - 4.0) < 0.3, order


def test_euler_maruyama_without_noise_is_forward_euler(ctx):
    say("EM with zero amplitude == FE")
    grid = lg.CartesianGrid(ctx, [128, 64], [64, 32], [0.0, 0.0], [0.4, 0.4])
    out = []
    for wiener in (True, False):
        for mode in (lg.LGD_MODE_FUSED, lg.LGD_MODE_PER_INTENT):
            sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0, 0.0), wiener=wiener,
                                        seed=11, mode=mode)
            sim.initialize_grains(np.array([[12.0, 12.0, 3.0], [35.0, 14.0, 2.0]]), 0.55)
            sim.step(0.008, 30)
            out.append((sim.field(0).download(), sim.field(1).download()))
            sim.close()
    m = np.zeros((4, 34, 66), dtype=bool)
    m[:, 1:-1, 1:-1] = True
    m = m.reshape(-1)
    for phi, u in out[1:]:
        assert np.array_equal(phi[m], out[0][0][m]) and np.array_equal(u[m], out[0][1][m])
    grid.close()
