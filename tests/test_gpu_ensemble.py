"""GPU parity of the batched ensemble: every realization of an uneven-sized
ensemble advanced in one launch per step equals its own Simulation run on the
oracle (bitwise without noise; within tolerance with noise), and equals the
per-block-launch dispatch bit for bit."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def make_ensemble(ctx, sides, noise, block=32, h=0.4):
    reals = []
    rng = np.random.default_rng(3)
    for i, s in enumerate(sides):
        eps4 = float(rng.uniform(0.02, 0.06))
        b = float(rng.uniform(0.005, 0.02)) if noise else 0.0
        reals.append(lg.Realization((s[0], s[1]), eps4, 3.19, 2.0, b, seed=100 + i))
    ens = lg.Ensemble(ctx, reals, block_cells=(block, block), spacing=(h, h))
    deltas = rng.uniform(0.45, 0.65, size=len(sides))
    for i, s in enumerate(sides):
        ens.initialize_grains(i, np.array([[s[0] * h / 2, s[1] * h / 2, 5 * h]]), float(deltas[i]))
    return ens, reals, deltas


SIDES = [(64, 64), (96, 32), (32, 128), (160, 96), (32, 32)]


@pytest.mark.parametrize("noise", [False, True])
def test_ensemble_members_match_their_own_oracle_runs(ctx, orc, noise):
    say(f"ensemble noise={noise}")
    ens, reals, deltas = make_ensemble(ctx, SIDES, noise)
    dt, steps = 0.008, 60
    ens.step(dt, steps)
    for i, (s, r) in enumerate(zip(SIDES, reals)):
        g = orc.grid(list(s), [32, 32], [0, 0], [0.4, 0.4])
        phi = np.zeros(orc.field_len(g, 1))
        u = np.zeros_like(phi)
        orc.initialize_grains(g, phi, u, np.array([[s[0] * 0.2, s[1] * 0.2, 2.0]]), float(deltas[i]))
        p = orc.params(r.eps4, r.lambda_, r.d_thermal, r.noise_amplitude)
        orc.modelc_run(g, p, phi, u, steps, dt, seed=r.seed, wiener=noise, nthreads=4)
        gp = H.field_to_global(ens.realization_field(i, 0), list(s), [32, 32], 1)
        gu = H.field_to_global(ens.realization_field(i, 1), list(s), [32, 32], 1)
        wp = H.field_to_global(phi, list(s), [32, 32], 1)
        wu = H.field_to_global(u, list(s), [32, 32], 1)
        if noise:
            assert np.max(np.abs(gp - wp)) <= 1e-10 and np.max(np.abs(gu - wu)) <= 1e-10
        else:
            assert np.array_equal(gp, wp) and np.array_equal(gu, wu)
    frac = ens.solid_fraction()
    for i, s in enumerate(SIDES):
        gp = H.field_to_global(ens.realization_field(i, 0), list(s), [32, 32], 1)
        assert frac[i] == np.count_nonzero(gp > 0.0) / gp.size


def test_batched_equals_per_block_launches(ctx, orc):
    say("ensemble batched vs per-block")
    a, _, _ = make_ensemble(ctx, SIDES, True)
    b, _, _ = make_ensemble(ctx, SIDES, True)
    a.step(0.008, 20)
    b.step(0.008, 20, per_block=True)
    assert np.array_equal(a.field(0).download(), b.field(0).download())
    assert np.array_equal(a.field(1).download(), b.field(1).download())


PAIRED_SIDES = [(64, 64), (128, 32), (64, 128), (192, 96), (64, 32)]  # even blocks per row


def test_block_pairs_equal_single_blocks_and_the_oracle(ctx, orc):
    """Two x-neighbouring blocks per wave (every realization's block rows pair up):
    bitwise equal to one block per wave with noise, and to the oracle without."""
    say("ensemble block pairs")
    outs = []
    try:
        for kernel in (3, 2):
            ctx.set_tuning("modelc_kernel", kernel)
            ens, _, _ = make_ensemble(ctx, PAIRED_SIDES, True)
            ens.step(0.008, 30)
            outs.append((ens.field(0).download(), ens.field(1).download()))
            ens.close()
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        ctx.set_tuning("modelc_kernel", 3)
        ens, reals, deltas = make_ensemble(ctx, PAIRED_SIDES, False)
        ens.step(0.008, 40)
        for i, (s, r) in enumerate(zip(PAIRED_SIDES, reals)):
            g = orc.grid(list(s), [32, 32], [0, 0], [0.4, 0.4])
            phi = np.zeros(orc.field_len(g, 1))
            u = np.zeros_like(phi)
            orc.initialize_grains(g, phi, u, np.array([[s[0] * 0.2, s[1] * 0.2, 2.0]]), float(deltas[i]))
            p = orc.params(r.eps4, r.lambda_, r.d_thermal, 0.0)
            orc.modelc_run(g, p, phi, u, 40, 0.008, nthreads=4)
            assert np.array_equal(H.field_to_global(ens.realization_field(i, 0), list(s), [32, 32], 1),
                                  H.field_to_global(phi, list(s), [32, 32], 1))
            assert np.array_equal(H.field_to_global(ens.realization_field(i, 1), list(s), [32, 32], 1),
                                  H.field_to_global(u, list(s), [32, 32], 1))
        ens.close()
    finally:
        ctx.set_tuning("modelc_kernel", 0)
