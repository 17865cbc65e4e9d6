"""The allocation-free reverse step (lgd_modelc_adjoint_step_noalloc), called through
the C ABI, against lgd_modelc_adjoint_step with its lgd_adjoint workspace on the
same inputs: (pphi, pu) bit for bit and the step's dL/dθ equal, on two calls in a
row, the second on a smaller grid (the static scratch keeps the first call's
contents)."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lg = pytest.importorskip("legendre_amd")

DT = 0.008
CASES = [
    # cells, block, periodic, grains (x, y, r): the larger grid first
    ([256, 192], [64, 48], [False, False], [[30.0, 24.0, 2.4], [70.0, 50.0, 2.0], [12.0, 60.0, 1.6]]),
    ([96, 64], [32, 32], [True, False], [[9.0, 8.0, 2.4], [27.0, 19.0, 1.6]]),
]


def say(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def upload(ctx, host):
    b = lg.DeviceBuffer(ctx, host.size)
    b.upload(host)
    return b


def case_inputs(orc, case):
    cells, block, periodic, grains = CASES[case]
    og = orc.grid(cells, block, [0, 0], [0.4, 0.4], periodic)
    n = orc.field_len(og, 1)
    phi, u = np.zeros(n), np.zeros(n)
    orc.initialize_grains(og, phi, u, np.array(grains), 0.55)
    orc.modelc_run(og, orc.params(0.05, 3.19, 2.0), phi, u, 20, DT)
    rng = np.random.default_rng(11 + case)
    return phi, u, rng.standard_normal(n), rng.standard_normal(n)


def test_noalloc_reverse_step_matches_the_workspace_step(ctx, orc):
    params = lg.ModelCParams(0.05, 3.19, 2.0, 0.0)
    pc = params.c()
    for case in range(len(CASES)):
        cells, block, periodic, _ = CASES[case]
        say(f"noalloc adjoint step case {case} cells={cells} block={block}")
        g = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4], periodic)
        phi, u, pphi0, pu0 = case_inputs(orc, case)
        dphi, du = upload(ctx, phi), upload(ctx, u)
        # The existing entry point with its workspace.
        adj = lg.Adjoint(g)
        adj.reset()
        wp, wu = upload(ctx, pphi0), upload(ctx, pu0)
        adj.step(params, dphi, du, DT, wp, wu)
        want_grad = adj.grad()
        want_p, want_u = wp.download(), wu.download()
        adj.close()
        # The allocation-free one.
        gp, gu = upload(ctx, pphi0), upload(ctx, pu0)
        grad = lg.DeviceBuffer(ctx, 3)
        rc = ctx._lib.lgd_modelc_adjoint_step_noalloc(ctx.handle, g.handle, C.byref(pc), dphi.handle,
                                                      du.handle, DT, gp.handle, gu.handle, grad.handle)
        lg._abi.check(rc, ctx.handle)
        got_grad = grad.download()
        assert np.array_equal(gp.download(), want_p) and np.array_equal(gu.download(), want_u)
        assert np.array_equal(got_grad, want_grad), (got_grad, want_grad)
        assert np.all(np.isfinite(got_grad)) and np.any(got_grad != 0.0)
        for b in (dphi, du, wp, wu, gp, gu, grad):
            b.free()
        g.close()


def test_noalloc_reverse_step_rejects_grids_above_its_scratch(ctx):
    say("noalloc adjoint step: size limit")
    cells, block = [1024, 1040], [64, 65]  # 1,064,960 cells > 2^20
    g = lg.CartesianGrid(ctx, cells, block, [0.0, 0.0], [0.4, 0.4])
    bufs = [g.alloc_field(1) for _ in range(4)]
    grad = lg.DeviceBuffer(ctx, 3)
    pc = lg.ModelCParams(0.05, 3.19, 2.0, 0.0).c()
    rc = ctx._lib.lgd_modelc_adjoint_step_noalloc(ctx.handle, g.handle, C.byref(pc), bufs[0].handle,
                                                  bufs[1].handle, DT, bufs[2].handle, bufs[3].handle,
                                                  grad.handle)
    assert rc == 5  # LGD_ERR_UNSUPPORTED
    for b in bufs + [grad]:
        b.free()
    g.close()
