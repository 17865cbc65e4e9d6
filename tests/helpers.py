"""Shared test helpers: slab layout conversions and independent global-array
restatements (no blocks, no ghost sweeps) used to cross-check the oracle."""
from __future__ import annotations

import math

import numpy as np


def slab_shape(block_cells, ghost):
    """numpy shape of one slab, dimension 0 fastest (last numpy axis = dim 0)."""
    return tuple(n + 2 * ghost for n in reversed(block_cells))


def field_view(field: np.ndarray, cells, block_cells, ghost):
    """View a flat field buffer as (nblocks, *slab_shape)."""
    nb = 1
    for c, b in zip(cells, block_cells):
        nb *= c // b
    return field.reshape((nb,) + slab_shape(block_cells, ghost))


def block_coords(b, cells, block_cells):
    out = []
    for c, n in zip(cells, block_cells):
        nb = c // n
        out.append(b % nb)
        b //= nb
    return out


def interior_slices(dim, ghost):
    return tuple(slice(ghost, -ghost if ghost else None) for _ in range(dim))


def global_to_field(u: np.ndarray, cells, block_cells, ghost, fill=0.0) -> np.ndarray:
    """Global interior array (numpy axes reversed: [z][y][x]) → flat field buffer."""
    dim = len(cells)
    nb = 1
    for c, b in zip(cells, block_cells):
        nb *= c // b
    f = np.full((nb,) + slab_shape(block_cells, ghost), fill, dtype=np.float64)
    for b in range(nb):
        co = block_coords(b, cells, block_cells)
        sl = tuple(
            slice(co[d] * block_cells[d], (co[d] + 1) * block_cells[d]) for d in reversed(range(dim))
        )
        f[(b,) + interior_slices(dim, ghost)] = u[sl]
    return f.ravel()


def field_to_global(field: np.ndarray, cells, block_cells, ghost) -> np.ndarray:
    dim = len(cells)
    fv = field_view(field, cells, block_cells, ghost)
    out = np.empty(tuple(reversed(cells)), dtype=np.float64)
    for b in range(fv.shape[0]):
        co = block_coords(b, cells, block_cells)
        sl = tuple(
            slice(co[d] * block_cells[d], (co[d] + 1) * block_cells[d]) for d in reversed(range(dim))
        )
        out[sl] = fv[(b,) + interior_slices(dim, ghost)]
    return out


def resolve(gi: int, n: int, periodic: bool) -> int:
    """Reflect (mirror) or wrap (periodic) one global coordinate into [0, n)."""
    if gi < 0:
        return gi + n if periodic else -1 - gi
    if gi >= n:
        return gi - n if periodic else 2 * n - 1 - gi
    return gi


def expected_copy_ghosts(u: np.ndarray, cells, block_cells, ghost, periodic) -> np.ndarray:
    """Brute force: every slab entry (ghosts included) = the global cell at the
    per-dimension reflected/wrapped coordinate — the value the dimension sweep
    produces for copy-type faces (mirror, exchange, periodic)."""
    dim = len(cells)
    nb = 1
    for c, b in zip(cells, block_cells):
        nb *= c // b
    shape = slab_shape(block_cells, ghost)
    f = np.empty((nb,) + shape, dtype=np.float64)
    for b in range(nb):
        co = block_coords(b, cells, block_cells)
        for idx in np.ndindex(*shape):
            g = []
            for ax, s in enumerate(idx):
                d = dim - 1 - ax
                gi = co[d] * block_cells[d] + s - ghost
                g.append(resolve(gi, cells[d], periodic[d]))
            f[(b,) + idx] = u[tuple(g)]
    return f.ravel()


def global_laplacian(u: np.ndarray, h, periodic) -> np.ndarray:
    """(2D+1)-point Laplacian on the global array with mirror/periodic padding,
    accumulated in the reference's order: acc = 0 + ((u+ − (c+c)) + u−)·(1/h²) per dim."""
    dim = u.ndim
    mode = ["wrap" if periodic[d] else "symmetric" for d in range(dim)]
    acc = np.zeros_like(u)
    c = u
    for d in range(dim):
        ax = dim - 1 - d
        pad = [(0, 0)] * dim
        pad[ax] = (1, 1)
        up = np.pad(u, pad, mode=mode[d])
        sl_p = [slice(None)] * dim
        sl_m = [slice(None)] * dim
        sl_p[ax] = slice(2, None)
        sl_m[ax] = slice(0, -2)
        inv = 1.0 / (h[d] * h[d])
        acc = acc + (((up[tuple(sl_p)] - (c + c)) + up[tuple(sl_m)]) * inv)
    return acc


def kr_flux_global(p: np.ndarray, eps4: float, h: float, tol: float = 1e-8, periodic=(False, False)):
    """Karma–Rappel flux divergence over a global 2-D array (axes [y][x]) with
    mirror/periodic padding, vectorised, in the reference's operation order."""
    a_bar = math.fma(3.0, -eps4, 1.0) if hasattr(math, "fma") else _fma(3.0, -eps4, 1.0)
    eps_p = 3.0 * eps4 / a_bar
    a12 = 4.0 * a_bar * eps_p
    inv_h2 = 1.0 / (h * h)
    q = 0.25
    P = p
    if periodic[0]:
        P = np.pad(P, [(0, 0), (1, 1)], mode="wrap")
    else:
        P = np.pad(P, [(0, 0), (1, 1)], mode="symmetric")
    if periodic[1]:
        P = np.pad(P, [(1, 1), (0, 0)], mode="wrap")
    else:
        P = np.pad(P, [(1, 1), (0, 0)], mode="symmetric")
    ny, nx = p.shape

    def at(di, dj):
        return P[1 + dj : 1 + dj + ny, 1 + di : 1 + di + nx]

    derx_r = at(1, 0) - at(0, 0)
    derx_l = at(0, 0) - at(-1, 0)
    derx_t = q * (((at(1, 1) - at(-1, 1)) + at(1, 0)) - at(-1, 0))
    derx_b = q * (((at(1, 0) - at(-1, 0)) + at(1, -1)) - at(-1, -1))
    dery_t = at(0, 1) - at(0, 0)
    dery_b = at(0, 0) - at(0, -1)
    dery_r = q * (((at(1, 1) - at(1, -1)) + at(0, 1)) - at(0, -1))
    dery_l = q * (((at(0, 1) - at(0, -1)) + at(-1, 1)) - at(-1, -1))

    def face(dx, dy):
        dx2 = dx * dx
        dy2 = dy * dy
        s = dx2 + dy2
        m = s * s
        flat = m <= tol
        with np.errstate(divide="ignore", invalid="ignore"):
            a = a_bar * (1.0 + (eps_p * (dx2 * dx2 + dy2 * dy2)) / m)
            da = ((((-a12) * dx) * dy) * (dx2 - dy2)) / m
        return np.where(flat, a_bar, a), np.where(flat, 0.0, da)

    a_r, da_r = face(derx_r, dery_r)
    a_l, da_l = face(derx_l, dery_l)
    a_t, da_t = face(derx_t, dery_t)
    a_b, da_b = face(derx_b, dery_b)
    j_r = a_r * ((a_r * derx_r) - (da_r * dery_r))
    j_l = a_l * ((a_l * derx_l) - (da_l * dery_l))
    j_t = a_t * ((a_t * dery_t) + (da_t * derx_t))
    j_b = a_b * ((a_b * dery_b) + (da_b * derx_b))
    return (((j_r - j_l) + j_t) - j_b) * inv_h2


def _fma(a: float, b: float, c: float) -> float:
    """Correctly rounded a*b+c via exact rationals (Python 3.10 has no math.fma)."""
    from fractions import Fraction

    r = Fraction(a) * Fraction(b) + Fraction(c)
    return float(r) if r != 0 else (0.0 if (a * b + c) >= 0 else -0.0)


def within(got, want, rel=1e-10, eps=1e-12) -> bool:
    """approx::assert_relative_eq!(got, want, max_relative = rel, epsilon = eps)."""
    got = np.asarray(got)
    want = np.asarray(want)
    diff = np.abs(got - want)
    tol = np.maximum(eps, rel * np.maximum(np.abs(got), np.abs(want)))
    return bool(np.all(diff <= tol))


def amr_interior_mask(amr) -> np.ndarray:
    """Interior entries of every patch of a device AmrGrid's field layout (ghost
    entries of a simulation's state are scratch after a step)."""
    m = np.zeros(amr.field_len, dtype=bool)
    for b in range(amr.num_blocks):
        _, lo, hi = amr.patch(b)
        nx, ny = hi[0] - lo[0], hi[1] - lo[1]
        o = amr.offset(b)
        m[o:o + (nx + 2) * (ny + 2)].reshape(ny + 2, nx + 2)[1:-1, 1:-1] = True
    return m


# A 3-level case whose level-1 boxes come from the nesting flags alone
# (src/geometry/amr/adapt.rs:194-211): base 32² in 16² blocks (h = 0.5), ratios
# [2, 2], one level-1 patch [16, 48)²; level 1 holds x-stripes of period 4
# (1, 1, 0, 0) on [24, 40) × [24, 36) and 0.5 elsewhere, level 0 holds 0.5.
# Undivided central differences are 0.5 inside the stripes and at most 0.25 at
# their edges, and the restriction of the stripes is a period-2 pattern whose
# central differences vanish, so at threshold 0.3 only level-1 cells are tagged:
# [25, 39) × [24, 36).  With buffer 1 they cluster into the level-1 box
# [24, 40) × [23, 37), i.e. level 2 = [48, 80) × [46, 74).  Coarsened to level 1
# and grown by one that is [23, 41) × [22, 38), whose level-0 parents are
# [11, 21) × [11, 19): level 1 = [22, 42) × [22, 38).
STRIPES = dict(cells=[32, 32], block=[16, 16], h=0.5, ratios=[2, 2],
               levels=[[((16, 16), (48, 48))]], policy=(4, 1, 0.7, 2, 0.3, 0),
               expect=[(1, (22, 22), (42, 38)), (2, (48, 46), (80, 74))])


def stripes_field(amr, n: int) -> np.ndarray:
    """The striped field on a hierarchy with AmrGrid's patch/offset (ghost 1)."""
    f = np.zeros(n)
    for b in range(amr.num_blocks):
        lev, lo, hi = amr.patch(b)
        nx, ny = hi[0] - lo[0], hi[1] - lo[1]
        o = amr.offset(b)
        v = f[o:o + (nx + 2) * (ny + 2)].reshape(ny + 2, nx + 2)
        v[:] = 0.5
        if lev == 1:
            for j in range(ny):
                for i in range(nx):
                    x, y = lo[0] + i, lo[1] + j
                    if 24 <= x < 40 and 24 <= y < 36:
                        v[j + 1, i + 1] = 1.0 if x % 4 < 2 else 0.0
    return f


def proper_nesting_violations(amr) -> list:
    """Cells of every level-l box (l >= 2) coarsened to level l-1 and grown by one,
    clipped to the level domain, that no level-(l-1) patch covers — the proper
    nesting the regrid's nesting flags guarantee (adapt.rs:194-211).  ratios are
    read off the hierarchy's boxes: this helper takes r = 2 between levels."""
    bad = []
    boxes = [amr.patch(b) for b in range(amr.num_blocks)]
    dom0 = None
    for lev, lo, hi in boxes:
        if lev == 0:
            dom0 = (max(dom0[0], hi[0]), max(dom0[1], hi[1])) if dom0 else hi
    for lev, lo, hi in boxes:
        if lev < 2:
            continue
        w, h = dom0[0] * 2 ** (lev - 1), dom0[1] * 2 ** (lev - 1)
        for y in range(max(lo[1] // 2 - 1, 0), min(hi[1] // 2 + 1, h)):
            for x in range(max(lo[0] // 2 - 1, 0), min(hi[0] // 2 + 1, w)):
                if amr.find_patch(lev - 1, (x, y)) is None:
                    bad.append((lev, x, y))
    return bad
