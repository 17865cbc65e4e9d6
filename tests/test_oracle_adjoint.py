"""Pins the oracle's Model C derivatives (oracle/legendre_oracle_adjoint.cpp), the
checkers of the device adjoint.  The reference has no derivative code, so the pins
are mathematical, and each derivation is checked against one it does not share
code with:

  tangent-linear run   values bitwise equal to the forward run (same operation
                       sequence); tangents vs central finite differences of the
                       forward run (the forward alone, no derivative code)
  adjoint step         transpose identity against the tangent-linear step:
                       <P', T δ> = <Tᵀ P', δ> for random P', δ (state, θ)
  full gradient        directional derivative = the tangent-linear run's
                       dL, and central finite differences of L in λ, D, ε₄
"""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H

orc = pytest.importorskip("oracle")

CASES = [
    # cells, block, periodic, grains (x, y, r)
    ([32, 24], [16, 12], [False, False], [[6.0, 5.0, 2.0]]),
    ([24, 24], [8, 12], [True, True], [[4.8, 4.8, 1.6], [2.0, 7.0, 1.2]]),
    ([20, 28], [20, 14], [True, False], [[7.9, 0.2, 1.8]]),
]
DT = 0.008


def setup(case, eps4=0.05, lam=3.19, d=2.0, steps=30):
    cells, block, periodic, grains = CASES[case]
    g = orc.grid(cells, block, [0, 0], [0.4, 0.4], periodic)
    p = orc.params(eps4, lam, d)
    n = orc.field_len(g, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(g, phi, u, np.array(grains), 0.55)
    orc.modelc_run(g, p, phi, u, steps, DT)  # a developed interface
    return g, p, phi, u, cells, block


def interior(field, cells, block):
    return H.field_to_global(field, cells, block, 1)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_tlm_values_equal_forward_bitwise(case):
    g, p, phi, u, cells, block = setup(case)
    a, b = phi.copy(), u.copy()
    orc.modelc_run(g, p, a, b, 25, DT)
    c, d = phi.copy(), u.copy()
    orc.modelc_tlm_run(g, p, c, d, np.zeros_like(c), np.zeros_like(d), [1.0, 1.0, 1.0], 25, DT)
    assert np.array_equal(interior(a, cells, block), interior(c, cells, block))
    assert np.array_equal(interior(b, cells, block), interior(d, cells, block))


def _run(g, p, phi, u, steps):
    a, b = phi.copy(), u.copy()
    orc.modelc_run(g, p, a, b, steps, DT)
    return a, b


@pytest.mark.parametrize("case", range(len(CASES)))
def test_tlm_tangents_match_central_differences(case):
    g, p, phi, u, cells, block = setup(case)
    steps = 15
    rng = np.random.default_rng(case)
    w = rng.standard_normal(cells[::-1])  # a fixed functional of the final phi
    base = dict(eps4=p.eps4, lambda_=p.lambda_, d_thermal=p.d_thermal)

    def functional(params, ph0, u0):
        a, _ = _run(g, params, ph0, u0, steps)
        return float(np.sum(w * interior(a, cells, block)))

    # parameter directions
    for k, name in enumerate(["lambda_", "d_thermal", "eps4"]):
        h = 1e-6 * max(1.0, abs(base[name]))
        pp = dict(base)
        pm = dict(base)
        pp[name] += h
        pm[name] -= h
        fd = (functional(orc.params(pp["eps4"], pp["lambda_"], pp["d_thermal"]), phi, u)
              - functional(orc.params(pm["eps4"], pm["lambda_"], pm["d_thermal"]), phi, u)) / (2 * h)
        c, d = phi.copy(), u.copy()
        dp, du = np.zeros_like(c), np.zeros_like(d)
        th = np.zeros(3)
        th[k] = 1.0
        orc.modelc_tlm_run(g, p, c, d, dp, du, th, steps, DT)
        tl = float(np.sum(w * interior(dp, cells, block)))
        assert abs(tl - fd) <= 1e-6 * max(abs(tl), 1e-3), (name, tl, fd)
    # a field direction on (phi0, u0)
    vp = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    vu = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    h = 1e-7
    fd = (functional(p, phi + h * vp, u + h * vu) - functional(p, phi - h * vp, u - h * vu)) / (2 * h)
    c, d = phi.copy(), u.copy()
    dp, du = vp.copy(), vu.copy()
    orc.modelc_tlm_run(g, p, c, d, dp, du, np.zeros(3), steps, DT)
    tl = float(np.sum(w * interior(dp, cells, block)))
    assert abs(tl - fd) <= 1e-6 * max(abs(tl), 1e-3), (tl, fd)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_adjoint_step_is_the_transpose_of_the_tangent_step(case):
    g, p, phi, u, cells, block = setup(case)
    rng = np.random.default_rng(10 + case)
    for trial in range(3):
        dphi = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
        du = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
        dth = rng.standard_normal(3)
        pphi = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
        pu = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
        # tangent step
        c, d = phi.copy(), u.copy()
        tp, tu = dphi.copy(), du.copy()
        orc.modelc_tlm_run(g, p, c, d, tp, tu, dth, 1, DT)
        lhs = float(np.sum(interior(pphi, cells, block) * interior(tp, cells, block))
                    + np.sum(interior(pu, cells, block) * interior(tu, cells, block)))
        # adjoint step
        ap, au = pphi.copy(), pu.copy()
        grad = np.zeros(3)
        orc.modelc_adjoint_step(g, p, phi, u, DT, ap, au, grad)
        rhs = float(np.sum(interior(ap, cells, block) * interior(dphi, cells, block))
                    + np.sum(interior(au, cells, block) * interior(du, cells, block))
                    + grad @ dth)
        scale = float(np.sum(np.abs(interior(pphi, cells, block) * interior(tp, cells, block)))
                      + np.sum(np.abs(interior(pu, cells, block) * interior(tu, cells, block))))
        assert abs(lhs - rhs) <= 1e-12 * scale, (trial, lhs, rhs)


def test_adjoint_parameter_terms_are_each_nonzero_and_transposed():
    """Each θ component alone satisfies the identity (a dropped term in one of
    the λ, D, ε₄ sums fails here even if another compensates in the combined test)."""
    g, p, phi, u, cells, block = setup(0)
    rng = np.random.default_rng(99)
    pphi = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    pu = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    ap, au = pphi.copy(), pu.copy()
    grad = np.zeros(3)
    orc.modelc_adjoint_step(g, p, phi, u, DT, ap, au, grad)
    for k in range(3):
        th = np.zeros(3)
        th[k] = 1.0
        c, d = phi.copy(), u.copy()
        tp, tu = np.zeros_like(c), np.zeros_like(d)
        orc.modelc_tlm_run(g, p, c, d, tp, tu, th, 1, DT)
        lhs = float(np.sum(interior(pphi, cells, block) * interior(tp, cells, block))
                    + np.sum(interior(pu, cells, block) * interior(tu, cells, block)))
        assert lhs != 0.0
        assert abs(lhs - grad[k]) <= 1e-11 * abs(lhs), (k, lhs, grad[k])


@pytest.mark.parametrize("case", [0, 1])
def test_full_gradient_matches_tlm_and_finite_differences(case):
    g, p, phi, u, cells, block = setup(case, steps=10)
    steps = 40
    # target: a forward run at other parameters (the calibration setup)
    tp_params = orc.params(0.06, 3.4, 2.1)
    target, _ = _run(g, tp_params, phi, u, steps)
    loss, grad, gphi, gu = orc.modelc_gradient(g, p, phi, u, target, steps, DT)
    fin, _ = _run(g, p, phi, u, steps)
    r = interior(fin, cells, block) - interior(target, cells, block)
    assert loss == pytest.approx(0.5 * float(np.sum(r * r)), rel=1e-14)
    rng = np.random.default_rng(5)
    vp = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    vu = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    th = rng.standard_normal(3)
    c, d = phi.copy(), u.copy()
    dp, du = vp.copy(), vu.copy()
    orc.modelc_tlm_run(g, p, c, d, dp, du, th, steps, DT)
    tlm = float(np.sum(r * interior(dp, cells, block)))
    adj = float(grad @ th + np.sum(interior(gphi, cells, block) * interior(vp, cells, block))
                + np.sum(interior(gu, cells, block) * interior(vu, cells, block)))
    assert abs(tlm - adj) <= 1e-10 * abs(tlm)

    def L(pr):
        a, _ = _run(g, pr, phi, u, steps)
        rr = interior(a, cells, block) - interior(target, cells, block)
        return 0.5 * float(np.sum(rr * rr))

    vals = dict(eps4=p.eps4, lambda_=p.lambda_, d_thermal=p.d_thermal)
    for k, name in enumerate(["lambda_", "d_thermal", "eps4"]):
        h = 1e-6
        up, dn = dict(vals), dict(vals)
        up[name] += h
        dn[name] -= h
        fd = (L(orc.params(up["eps4"], up["lambda_"], up["d_thermal"]))
              - L(orc.params(dn["eps4"], dn["lambda_"], dn["d_thermal"]))) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * abs(grad[k]), (name, fd, grad[k])


def test_tlm_several_directions_equal_single_runs():
    """Four tangents carried at once equal four single-direction runs bitwise."""
    g, p, phi, u, cells, block = setup(0)
    rng = np.random.default_rng(7)
    n = phi.size
    dth = rng.standard_normal((4, 3))
    dp0 = np.stack([H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1) for _ in range(4)])
    du0 = np.stack([H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1) for _ in range(4)])
    c, d = phi.copy(), u.copy()
    dp, du = dp0.copy(), du0.copy()
    orc.modelc_tlm_run(g, p, c, d, dp, du, dth, 5, DT)
    for k in range(4):
        c1, d1 = phi.copy(), u.copy()
        a, b = dp0[k].copy(), du0[k].copy()
        orc.modelc_tlm_run(g, p, c1, d1, a, b, dth[k], 5, DT)
        assert np.array_equal(interior(a, cells, block), interior(dp[k], cells, block))
        assert np.array_equal(interior(b, cells, block), interior(du[k], cells, block))
        assert np.array_equal(interior(c1, cells, block), interior(c, cells, block))
    assert n == dp.shape[1]


def test_branch_crossings_count_the_faces_between_two_tolerances(orc):
    """One spike φ = c at cell (1, 1) of a 4×4 block: the four faces of that cell and
    the facing face of each of its four neighbours have |∇φ|⁴ = c⁴ (normal difference
    ±c, transverse 0); every other face has at most c⁴/256.  Between fv_tol = c⁴/2 and
    2c⁴ exactly those 8 evaluations change branch in the first step's right-hand
    side; tolerances below c⁴/256 on both sides change none."""
    g = orc.grid([4, 4], [4, 4], [0, 0], [0.4, 0.4])
    n = orc.field_len(g, 1)
    phi, u = np.zeros(n), np.zeros(n)
    c = 0.1
    phi[(1 + 1) * 6 + (1 + 1)] = c
    c4 = c ** 4

    def p(tol):
        return orc.params(0.05, 3.19, 2.0, 0.0, anisotropy_tol=1e-12, fv_tol=tol)

    assert orc.modelc_branch_crossings(g, p(0.5 * c4), p(2.0 * c4), phi, u, 1, 0.008) == 8
    assert orc.modelc_branch_crossings(g, p(1e-3 * c4), p(2e-3 * c4), phi, u, 1, 0.008) == 0
    assert orc.modelc_branch_crossings(g, p(0.5 * c4), p(0.5 * c4), phi, u, 5, 0.008) == 0
