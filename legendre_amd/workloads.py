"""Synthetic workloads of the benchmark configurations (setup only, not timed).

heat3d   128³ cells as 4×4×4 blocks of 32³, periodic, D = 1, dx = 1/128,
         dt = 0.9·dx²/(6D), u₀ = U[0,1) from the counter RNG + a Gaussian bump.
model_c  10080² cells as 80×80 blocks of 126², h = 0.4, dt = 0.008, λ = 3.19,
         D = 2.0, ε₄ = 0.05, u = −0.55, 64 seeds of radius 5h placed by the
         seed_positions recipe (examples/model_c.rs:175-206), mirror faces.
adjoint  1024² cells as 16×16 blocks of 64², b = 0, 2000 steps, L = ½‖φ_T − φ_target‖²
         with φ_target from a forward run at (λ, D, ε₄) = (3.4, 2.1, 0.06); gradient at
         (3.19, 2.0, 0.05) w.r.t. (λ, D, ε₄) and the initial fields; 45 snapshots;
         16 seeds of radius 5h by the same seed_positions recipe, u = −0.55.
ensemble 512 independent Model C realizations, side ∈ {256, 384, 512, 768, 1024}
         (square), b ∈ U[0.005, 0.02], Δ ∈ U[0.45, 0.65], ε₄ ∈ U[0.02, 0.06], all drawn
         as unit_open(mix_key(seed, [r, k])), blocks of 64², h = 0.4, dt = 0.008,
         λ = 3.19, D = 2.0, one seed of radius 5h at the centre, noise seed 1000 + r,
         2000 steps.
amr      Model C on a Berger–Oliger hierarchy: 2048² base in 64² blocks (h = 0.4), ratios
         [2, 2] (3 levels), subcycled, coarse dt = 0.2·h²/D, gradient tagging
         max_d |½(φ[+]−φ[−])| > 0.1, regrid every 20 coarse steps (buffer 2, efficiency
         0.8, minimum width 16), 16 seeds of radius 5h (seed_positions recipe), b = 0,
         500 coarse steps.
"""
from __future__ import annotations

import dataclasses

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_np(z: np.ndarray) -> np.ndarray:
    """Vectorised SplitMix64 finaliser (src/util/rng.rs:13-18), wrapping uint64."""
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z += np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def mix_key_np(seed: int, parts: list) -> np.ndarray:
    """mix_key (src/util/rng.rs:23-29) with array-valued parts."""
    s = splitmix64_np(np.array([seed], dtype=np.uint64))
    for p in parts:
        s = splitmix64_np(s ^ np.asarray(p, dtype=np.uint64))
    return s


def unit_open_np(bits: np.ndarray) -> np.ndarray:
    return ((bits >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass(frozen=True)
class Heat3dConfig:
    cells: int = 128
    block: int = 32
    kappa: float = 1.0
    steps: int = 100
    seed: int = 42

    @property
    def dx(self) -> float:
        return 1.0 / self.cells

    @property
    def dt(self) -> float:
        return 0.9 * self.dx * self.dx / (6.0 * self.kappa)

    def initial_global(self) -> np.ndarray:
        """u₀ over the global grid, numpy axes [z][y][x]."""
        n = self.cells
        lin = np.arange(n**3, dtype=np.uint64)
        bits = splitmix64_np(mix_key_np(self.seed, [lin]))
        u = (bits >> np.uint64(11)).astype(np.float64) * 2.0**-53  # U[0,1)
        u = u.reshape(n, n, n)
        c = (np.arange(n) + 0.5) * self.dx
        r2 = (c[None, None, :] - 0.5) ** 2 + (c[None, :, None] - 0.5) ** 2 + (c[:, None, None] - 0.5) ** 2
        return u + np.exp(-r2 / (2.0 * 0.1**2))


@dataclasses.dataclass(frozen=True)
class ModelCConfig:
    cells: int = 10080
    block: int = 126
    h: float = 0.4
    dt: float = 0.008
    eps4: float = 0.05
    lambda_: float = 3.19
    d_thermal: float = 2.0
    delta: float = 0.55
    nseeds: int = 64
    seed: int = 42
    noise: float = 0.0
    steps: int = 1000

    @property
    def interior_cells(self) -> int:
        return self.cells * self.cells

    def grains(self) -> np.ndarray:
        """seed_positions (examples/model_c.rs:192-205) with radius 5h."""
        radius = 5.0 * self.h
        extent = self.cells * self.h
        margin = 2.0 * radius
        span = np.fma(2.0, -margin, extent) if hasattr(np, "fma") else _fma(2.0, -margin, extent)
        out = np.empty((self.nseeds, 3))
        for i in range(self.nseeds):
            ux = unit_open_np(mix_key_np(self.seed, [i, 0]))[0]
            uy = unit_open_np(mix_key_np(self.seed, [i, 1]))[0]
            out[i] = (margin + ux * span, margin + uy * span, radius)
        return out


def _fma(a: float, b: float, c: float) -> float:
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def to_slabs(u: np.ndarray, cells, block, ghost: int = 1) -> np.ndarray:
    """A global interior array (numpy axes reversed: [z][y][x]) as a field buffer of
    ghost-inclusive block slabs, block-major, dimension 0 fastest (ghosts zero)."""
    dim = len(cells)
    nbs = [c // b for c, b in zip(cells, block)]
    shape = tuple(b + 2 * ghost for b in reversed(block))
    out = np.zeros((int(np.prod(nbs)),) + shape)
    for bid in range(out.shape[0]):
        co, r = [], bid
        for nb in nbs:
            co.append(r % nb)
            r //= nb
        src = tuple(slice(co[d] * block[d], (co[d] + 1) * block[d]) for d in reversed(range(dim)))
        dst = (bid,) + tuple(slice(ghost, ghost + block[d]) for d in reversed(range(dim)))
        out[dst] = u[src]
    return out.reshape(-1)


HEAT3D = Heat3dConfig()
MODEL_C = ModelCConfig()


@dataclasses.dataclass(frozen=True)
class AdjointConfig:
    cells: int = 1024
    block: int = 64
    h: float = 0.4
    dt: float = 0.008
    steps: int = 2000
    snapshots: int = 45
    lambda_: float = 3.19
    d_thermal: float = 2.0
    eps4: float = 0.05
    target_lambda: float = 3.4
    target_d: float = 2.1
    target_eps4: float = 0.06
    delta: float = 0.55
    nseeds: int = 16
    seed: int = 42

    @property
    def interior_cells(self) -> int:
        return self.cells * self.cells

    def grains(self) -> np.ndarray:
        return ModelCConfig(cells=self.cells, block=self.block, h=self.h, nseeds=self.nseeds,
                            seed=self.seed).grains()


@dataclasses.dataclass(frozen=True)
class EnsembleConfig:
    n: int = 512
    sides: tuple = (256, 384, 512, 768, 1024)
    block: int = 64
    h: float = 0.4
    dt: float = 0.008
    steps: int = 2000
    lambda_: float = 3.19
    d_thermal: float = 2.0
    seed: int = 42

    def members(self) -> list:
        """Per realization: dict(side, b, delta, eps4, seed)."""
        out = []
        for r in range(self.n):
            u = [float(unit_open_np(mix_key_np(self.seed, [r, k]))[0]) for k in range(4)]
            side = self.sides[min(int(u[0] * len(self.sides)), len(self.sides) - 1)]
            out.append(dict(side=side, b=0.005 + 0.015 * u[1], delta=0.45 + 0.2 * u[2],
                            eps4=0.02 + 0.04 * u[3], seed=1000 + r))
        return out

    def grains(self, side: int) -> np.ndarray:
        c = side * self.h / 2.0
        return np.array([[c, c, 5.0 * self.h]])


@dataclasses.dataclass(frozen=True)
class AmrConfig:
    cells: int = 2048
    block: int = 64
    h: float = 0.4
    ratios: tuple = (2, 2)
    eps4: float = 0.05
    lambda_: float = 3.19
    d_thermal: float = 2.0
    delta: float = 0.55
    nseeds: int = 16
    seed: int = 42
    coarse_steps: int = 500
    every: int = 20
    buffer: int = 2
    efficiency: float = 0.8
    min_width: int = 16
    threshold: float = 0.1

    @property
    def dt(self) -> float:
        return 0.2 * self.h * self.h / self.d_thermal

    def grains(self) -> np.ndarray:
        return ModelCConfig(cells=self.cells, block=self.block, h=self.h, nseeds=self.nseeds,
                            seed=self.seed).grains()

