"""GPU parity of the Model C adjoint: one reverse step against the oracle's
hand-written adjoint (itself pinned by the transpose identity against the
tangent-linear model), the checkpointed calibration gradient against the
oracle's all-states gradient, bitwise independence of the snapshot count,
step counts equal to the schedule, and central finite differences of the GPU
loss itself."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")

CASES = [
    # cells, block, periodic, grains (x, y, r)
    ([64, 48], [32, 24], [False, False], [[9.0, 8.0, 2.4], [17.0, 12.0, 1.6]]),
    ([48, 48], [16, 24], [True, True], [[4.8, 4.8, 2.0], [12.0, 15.0, 1.6]]),
    ([40, 56], [40, 28], [True, False], [[7.9, 0.3, 2.0]]),
]
DT = 0.008


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel_close(got, want, rel):
    scale = max(float(np.max(np.abs(want))), 1e-300)
    return float(np.max(np.abs(got - want))) <= rel * scale


def state(orc, case, steps=20):
    cells, block, periodic, grains = CASES[case]
    og = orc.grid(cells, block, [0, 0], [0.4, 0.4], periodic)
    op = orc.params(0.05, 3.19, 2.0)
    n = orc.field_len(og, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(og, phi, u, np.array(grains), 0.55)
    orc.modelc_run(og, op, phi, u, steps, DT)
    return og, op, phi, u


def gpu_grid(ctx, case):
    cells, block, periodic, _ = CASES[case]
    return lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4], periodic)


def upload(ctx, host):
    b = lg.DeviceBuffer(ctx, host.size)
    b.upload(host)
    return b


# Reverse-step paths: the three passes, and the streaming kernel with the default
# strip height and with strip heights that split blocks raggedly.
PATHS = {"passes": (1, 0), "stream": (2, 0), "stream_rows3": (2, 3), "stream_rows5": (2, 5)}


@pytest.fixture
def adjoint_path(ctx, request):
    kernel, rows = PATHS[request.param]
    ctx.set_tuning("adjoint_kernel", kernel)
    ctx.set_tuning("adjoint_rows", rows)
    yield request.param
    ctx.set_tuning("adjoint_kernel", 0)
    ctx.set_tuning("adjoint_rows", 0)


@pytest.mark.parametrize("adjoint_path", list(PATHS), indirect=True)
@pytest.mark.parametrize("case", range(len(CASES)))
def test_adjoint_step_matches_oracle(ctx, orc, case, adjoint_path):
    cells, block = CASES[case][0], CASES[case][1]
    say(f"adjoint step cells={cells} block={block} path={adjoint_path}")
    og, op, phi, u = state(orc, case)
    rng = np.random.default_rng(case)
    pphi = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    pu = H.global_to_field(rng.standard_normal(cells[::-1]), cells, block, 1)
    grid = gpu_grid(ctx, case)
    adj = lg.Adjoint(grid)
    params = lg.ModelCParams(0.05, 3.19, 2.0)
    dphi, du, dpp, dpu = upload(ctx, phi), upload(ctx, u), upload(ctx, pphi), upload(ctx, pu)
    adj.step(params, dphi, du, DT, dpp, dpu)
    grad = adj.grad()
    wp, wu, wg = pphi.copy(), pu.copy(), np.zeros(3)
    orc.modelc_adjoint_step(og, op, phi, u, DT, wp, wu, wg)
    gp = H.field_to_global(dpp.download(), cells, block, 1)
    gu = H.field_to_global(dpu.download(), cells, block, 1)
    assert rel_close(gp, H.field_to_global(wp, cells, block, 1), 1e-11)
    assert rel_close(gu, H.field_to_global(wu, cells, block, 1), 1e-11)
    assert np.allclose(grad, wg, rtol=1e-11, atol=0.0), (grad, wg)
    # the accumulator sums over steps and resets to zero
    adj.step(params, dphi, du, DT, dpp, dpu)
    assert not np.array_equal(adj.grad(), grad)
    adj.reset()
    assert np.array_equal(adj.grad(), np.zeros(3))
    adj.close()


def _calibration_inputs(orc, case, steps):
    og, op, phi, u = state(orc, case, steps=10)
    target, tu = phi.copy(), u.copy()
    orc.modelc_run(og, orc.params(0.06, 3.4, 2.1), target, tu, steps, DT)
    return og, op, phi, u, target


@pytest.mark.parametrize("case", [0, 1])
def test_calibration_gradient_matches_oracle_for_every_snapshot_count(ctx, orc, case):
    cells, block = CASES[case][0], CASES[case][1]
    steps = 30
    say(f"calibration cells={cells} steps={steps}")
    og, op, phi, u, target = _calibration_inputs(orc, case, steps)
    wl, wg, wgp, wgu = orc.modelc_gradient(og, op, phi, u, target, steps, DT)
    grid = gpu_grid(ctx, case)
    params = lg.ModelCParams(0.05, 3.19, 2.0)
    dphi, du, dt_ = upload(ctx, phi), upload(ctx, u), upload(ctx, target)
    ref = None
    for snaps in [1, 2, 5, steps]:
        cal = lg.ModelCCalibration(grid, steps, DT, snapshots=snaps)
        gp, gu = grid.alloc_field(), grid.alloc_field()
        loss = cal.forward(params, dphi, du, dt_)
        grad = cal.backward(gp, gu)
        fwd, rev = cal.counts()
        assert fwd == lg.checkpoint_forward_steps(steps, snaps)
        assert rev == steps
        out = (loss, grad, H.field_to_global(gp.download(), cells, block, 1),
               H.field_to_global(gu.download(), cells, block, 1))
        if ref is None:
            ref = out
            assert loss == pytest.approx(wl, rel=1e-12)
            assert np.allclose(grad, wg, rtol=1e-9, atol=0.0), (grad, wg)
            assert rel_close(out[2], H.field_to_global(wgp, cells, block, 1), 1e-9)
            assert rel_close(out[3], H.field_to_global(wgu, cells, block, 1), 1e-9)
        else:  # the schedule changes only what is recomputed, never a bit of the result
            assert out[0] == ref[0]
            assert np.array_equal(out[1], ref[1])
            assert np.array_equal(out[2], ref[2]) and np.array_equal(out[3], ref[3])
        cal.close()


@pytest.mark.parametrize("case", [0, 2])
def test_calibration_streaming_and_three_pass_agree(ctx, orc, case):
    """The streaming reverse step and the three passes are the same derivative; they
    sum in different orders, so they agree to rounding, not bitwise."""
    cells, block = CASES[case][0], CASES[case][1]
    steps = 12
    say(f"calibration paths cells={cells} steps={steps}")
    og, op, phi, u, target = _calibration_inputs(orc, case, steps)
    grid = gpu_grid(ctx, case)
    params = lg.ModelCParams(0.05, 3.19, 2.0)
    dphi, du, dt_ = upload(ctx, phi), upload(ctx, u), upload(ctx, target)
    outs = []
    try:
        for kernel in (1, 2):
            ctx.set_tuning("adjoint_kernel", kernel)
            cal = lg.ModelCCalibration(grid, steps, DT, snapshots=3)
            gp, gu = grid.alloc_field(), grid.alloc_field()
            cal.forward(params, dphi, du, dt_)
            grad = cal.backward(gp, gu)
            outs.append((grad, H.field_to_global(gp.download(), cells, block, 1),
                         H.field_to_global(gu.download(), cells, block, 1)))
            cal.close()
    finally:
        ctx.set_tuning("adjoint_kernel", 0)
    (g1, p1, u1), (g2, p2, u2) = outs
    assert np.allclose(g2, g1, rtol=1e-11, atol=0.0), (g1, g2)
    assert rel_close(p2, p1, 1e-11) and rel_close(u2, u1, 1e-11)


def _fd_check(ctx, orc, case, steps, hs_by_param):
    """Central differences of the device loss against the device gradient at the step
    sizes given per parameter, each with the oracle's count of face evaluations whose
    flat branch |∇φ|⁴ <= tol (finite_volume.rs:129-142) differs between the θ − h and
    θ + h runs: returns [(name, h, crossings, rel err)]."""
    og, op, phi, u, target = _calibration_inputs(orc, case, steps)
    grid = gpu_grid(ctx, case)
    dphi, du, dt_ = upload(ctx, phi), upload(ctx, u), upload(ctx, target)
    cal = lg.ModelCCalibration(grid, steps, DT, snapshots=8)
    base = dict(eps4=0.05, lambda_=3.19, d_thermal=2.0)

    def at(**kw):
        pr = dict(base)
        pr.update(kw)
        return pr

    def loss(pr):
        return cal.forward(lg.ModelCParams(pr["eps4"], pr["lambda_"], pr["d_thermal"]), dphi, du, dt_)

    loss(base)
    grad = cal.backward()
    out = []
    for k, name in enumerate(["lambda_", "d_thermal", "eps4"]):
        for h in hs_by_param[name]:
            lo, hi = at(**{name: base[name] - h}), at(**{name: base[name] + h})
            cross = orc.modelc_branch_crossings(og, orc.params(lo["eps4"], lo["lambda_"], lo["d_thermal"]),
                                                orc.params(hi["eps4"], hi["lambda_"], hi["d_thermal"]),
                                                phi, u, steps, DT, 8)
            fd = (loss(hi) - loss(lo)) / (2 * h)
            out.append((name, h, cross, abs(fd - grad[k]) / abs(grad[k])))
    cal.close()
    return out


def test_calibration_gradient_matches_central_differences_at_h_1e_6(ctx, orc):
    """FD at h = 1e-6 (no other step size) on a run where no face changes branch
    between θ ± h: every parameter within 1e-6."""
    say("calibration FD h=1e-6 steps=30")
    res = _fd_check(ctx, orc, 0, 30, {n: [1e-6] for n in ("lambda_", "d_thermal", "eps4")})
    for name, h, cross, err in res:
        assert cross == 0, (name, cross)
        assert err <= 1e-6, (name, h, err)


def test_central_differences_fail_only_where_a_face_changes_branch(ctx, orc):
    """60 steps: at h = 1e-6 one ε₄ face evaluation takes the other branch of the
    face function between θ − h and θ + h, and that central difference carries the
    jump; every step size with no crossing agrees within 1e-6."""
    say("calibration FD branch crossings steps=60")
    res = _fd_check(ctx, orc, 0, 60, {"lambda_": [1e-6], "d_thermal": [1e-6], "eps4": [1e-6, 1e-7]})
    for name, h, cross, err in res:
        print(f"{name} h={h:g}: crossings {cross}, FD rel err {err:.2e}", flush=True)
        if cross == 0:
            assert err <= 1e-6, (name, h, err)
    assert any(cross > 0 for name, h, cross, err in res if name == "eps4" and h == 1e-6)
    assert any(cross == 0 for name, h, cross, err in res if name == "eps4" and h == 1e-7)


def test_closing_the_context_first_releases_everything():
    """Objects that outlive their context are released by its close (no library
    call on a destroyed context afterwards)."""
    say("context close releases dependents")
    c2 = lg.Context(0)
    grid = lg.CartesianGrid(c2, [32, 32], [16, 16], [0.0, 0.0], [0.4, 0.4])
    buf = grid.alloc_field()
    adj = lg.Adjoint(grid)
    cal = lg.ModelCCalibration(grid, 4, DT, snapshots=2)
    sim = lg.Simulation.model_c(grid, lg.ModelCParams(0.05, 3.19, 2.0))
    c2.close()
    for o in (buf, adj, cal, sim, grid):
        assert not o.handle
    del buf, adj, cal, sim, grid
