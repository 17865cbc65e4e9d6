"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liblegendre_oracle.so, the C++ restatement of the
reference crate's serial semantics (see legendre_oracle.h).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package; the product (legendre_amd) never does.

Parity status of each restated function (how it is pinned) is listed in
tests/test_oracle.py; every function there is pinned by a closed form, a
known-answer vector, an invariant or a brute-force second implementation.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess
import struct
from typing import Optional, Sequence

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liblegendre_oracle.so"

MIRROR, DIRICHLET, FLUX, EXTRAPOLATE = 0, 1, 2, 3


class Grid(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("cells", C.c_int64 * 3),
        ("block_cells", C.c_int64 * 3),
        ("origin", C.c_double * 3),
        ("spacing", C.c_double * 3),
    ]


class FaceBc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("value", C.c_double)]


class ModelCParams(C.Structure):
    _fields_ = [
        ("eps4", C.c_double),
        ("lambda_", C.c_double),
        ("d_thermal", C.c_double),
        ("noise_amplitude", C.c_double),
        ("anisotropy_tol", C.c_double),
        ("fv_tol", C.c_double),
    ]


DP = C.POINTER(C.c_double)
GP = C.POINTER(Grid)

_SIGS = {
    "orc_splitmix64": (C.c_uint64, [C.c_uint64]),
    "orc_mix_key": (C.c_uint64, [C.c_uint64, C.POINTER(C.c_uint64), C.c_size_t]),
    "orc_unit_open": (C.c_double, [C.c_uint64]),
    "orc_standard_normal": (C.c_double, [C.c_uint64]),
    "orc_grid_check": (C.c_int, [GP]),
    "orc_num_blocks": (C.c_int64, [GP]),
    "orc_block_len": (C.c_int64, [GP, C.c_uint32]),
    "orc_face_neighbor": (C.c_int64, [GP, C.c_int64, C.c_int32, C.c_int32]),
    "orc_cell_key": (C.c_int64, [GP, C.c_uint32, C.c_int64]),
    "orc_cell_center": (None, [GP, C.c_int64, C.POINTER(C.c_int64), DP]),
    "orc_fill_ghosts": (C.c_int, [GP, DP, C.c_uint32, C.POINTER(FaceBc)]),
    "orc_central_laplacian": (None, [GP, DP, DP, C.c_uint32]),
    "orc_karma_rappel_flux": (None, [GP, DP, DP, C.c_double, C.c_double]),
    "orc_modelc_rhs": (None, [GP, C.POINTER(ModelCParams), DP, DP, DP, DP]),
    "orc_modelc_amplitude": (None, [GP, C.POINTER(ModelCParams), DP]),
    "orc_heat_rhs": (None, [GP, C.c_double, DP, DP]),
    "orc_axpy": (None, [DP, DP, C.c_double, C.c_size_t]),
    "orc_apply_time": (None, [DP, DP, C.c_double, C.c_size_t]),
    "orc_apply_wiener": (None, [GP, DP, DP, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64]),
    "orc_modelc_run": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, C.c_int64, C.c_double, DP,
                                 C.c_uint64, C.c_int32, C.c_int32]),
    "orc_heat_run": (C.c_int, [GP, C.c_double, DP, C.c_int64, C.c_double, DP, C.POINTER(FaceBc), C.c_int32]),
    "orc_initialize_grains": (None, [GP, DP, DP, DP, C.c_size_t, C.c_double]),
    "orc_seed_positions": (None, [C.c_uint64, C.c_size_t, C.c_double, C.c_double, DP]),
}

_lib = None


def build() -> pathlib.Path:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def grid(cells, block_cells, origin, spacing, periodic=None) -> Grid:
    g = Grid()
    g.dim = len(cells)
    for d in range(len(cells)):
        g.cells[d] = int(cells[d])
        g.block_cells[d] = int(block_cells[d])
        g.origin[d] = float(origin[d])
        g.spacing[d] = float(spacing[d])
        g.periodic[d] = int(bool(periodic[d])) if periodic else 0
    return g


def bcs_array(bcs: Optional[Sequence]):
    """bcs: sequence of (kind, value) pairs, 2*dim entries [d*2 + (side > 0)]."""
    if bcs is None:
        return None
    arr = (FaceBc * len(bcs))()
    for i, (k, v) in enumerate(bcs):
        arr[i].kind = k
        arr[i].value = v
    return arr


def params(eps4, lambda_, d_thermal, noise=0.0, anisotropy_tol=1e-8, fv_tol=1e-8) -> ModelCParams:
    p = ModelCParams()
    p.eps4, p.lambda_, p.d_thermal = eps4, lambda_, d_thermal
    p.noise_amplitude, p.anisotropy_tol, p.fv_tol = noise, anisotropy_tol, fv_tol
    return p


# ---- rng -------------------------------------------------------------------
def splitmix64(z: int) -> int:
    return int(lib().orc_splitmix64(z))


def mix_key(seed: int, parts: Sequence[int]) -> int:
    arr = (C.c_uint64 * len(parts))(*parts)
    return int(lib().orc_mix_key(seed, arr, len(parts)))


def unit_open(bits: int) -> float:
    return float(lib().orc_unit_open(bits))


def standard_normal(key: int) -> float:
    return float(lib().orc_standard_normal(key))


# ---- grid ------------------------------------------------------------------
def num_blocks(g: Grid) -> int:
    return int(lib().orc_num_blocks(C.byref(g)))


def block_len(g: Grid, ghost: int) -> int:
    return int(lib().orc_block_len(C.byref(g), ghost))


def field_len(g: Grid, ghost: int) -> int:
    return num_blocks(g) * block_len(g, ghost)


def face_neighbor(g: Grid, block: int, dim: int, direction: int) -> Optional[int]:
    r = int(lib().orc_face_neighbor(C.byref(g), block, dim, direction))
    return None if r < 0 else r


def cell_key(g: Grid, ghost: int, offset: int) -> Optional[int]:
    r = int(lib().orc_cell_key(C.byref(g), ghost, offset))
    return None if r < 0 else r


def cell_center(g: Grid, block: int, idx) -> tuple:
    ia = (C.c_int64 * 3)(*(list(idx) + [0] * (3 - len(idx))))
    out = (C.c_double * 3)()
    lib().orc_cell_center(C.byref(g), block, ia, out)
    return tuple(out[d] for d in range(g.dim))


# ---- kernels ---------------------------------------------------------------
def fill_ghosts(g: Grid, field: np.ndarray, ghost: int = 1, bcs=None) -> None:
    rc = lib().orc_fill_ghosts(C.byref(g), _dp(field), ghost, bcs_array(bcs))
    if rc != 0:
        raise ValueError("ghost width exceeds a block extent")


def central_laplacian(g: Grid, inp: np.ndarray, out: np.ndarray, ghost: int = 1) -> None:
    lib().orc_central_laplacian(C.byref(g), _dp(inp), _dp(out), ghost)


def karma_rappel_flux(g: Grid, phi: np.ndarray, out: np.ndarray, eps4: float, tol: float = 1e-8):
    lib().orc_karma_rappel_flux(C.byref(g), _dp(phi), _dp(out), eps4, tol)


def modelc_rhs(g: Grid, p: ModelCParams, phi, u, kphi, ku) -> None:
    lib().orc_modelc_rhs(C.byref(g), C.byref(p), _dp(phi), _dp(u), _dp(kphi), _dp(ku))


def modelc_amplitude(g: Grid, p: ModelCParams, amp) -> None:
    lib().orc_modelc_amplitude(C.byref(g), C.byref(p), _dp(amp))


def heat_rhs(g: Grid, kappa: float, u, k) -> None:
    lib().orc_heat_rhs(C.byref(g), kappa, _dp(u), _dp(k))


def axpy(x: np.ndarray, y: np.ndarray, alpha: float) -> None:
    lib().orc_axpy(_dp(x), _dp(y), alpha, x.size)


def apply_wiener(g: Grid, x, amp, ghost, dt, seed, salt, stream) -> None:
    lib().orc_apply_wiener(C.byref(g), _dp(x), _dp(amp), ghost, dt, seed, salt, stream)


def modelc_run(g: Grid, p: ModelCParams, phi, u, nsteps, dt, t0=0.0, seed=0, wiener=False,
               nthreads=1) -> float:
    t = C.c_double(t0)
    rc = lib().orc_modelc_run(C.byref(g), C.byref(p), _dp(phi), _dp(u), nsteps, dt, C.byref(t),
                              seed, int(wiener), nthreads)
    if rc != 0:
        raise ValueError("orc_modelc_run rejected the configuration")
    return t.value


def heat_run(g: Grid, kappa, u, nsteps, dt, t0=0.0, bcs=None, nthreads=1) -> float:
    t = C.c_double(t0)
    rc = lib().orc_heat_run(C.byref(g), kappa, _dp(u), nsteps, dt, C.byref(t), bcs_array(bcs), nthreads)
    if rc != 0:
        raise ValueError("orc_heat_run rejected the configuration")
    return t.value


def initialize_grains(g: Grid, phi, u, grains: np.ndarray, u0: float) -> None:
    gr = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 3)
    lib().orc_initialize_grains(C.byref(g), _dp(phi), _dp(u), _dp(gr), gr.shape[0], u0)


def seed_positions(seed: int, nseeds: int, extent: float, radius: float) -> np.ndarray:
    out = np.empty((nseeds, 3), dtype=np.float64)
    lib().orc_seed_positions(seed, nseeds, extent, radius, _dp(out))
    return out


def time_bits(t: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", t))[0]
