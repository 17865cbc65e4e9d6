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

DRIVER_TIME, DRIVER_WIENER, DRIVER_JUMP = 0, 1, 2


class Driver(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_uint32), ("amp", C.POINTER(DP)), ("rate", DP)]

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
    "orc_modelc_branch_crossings": (C.c_int, [GP, C.POINTER(ModelCParams), C.POINTER(ModelCParams), DP, DP,
                                              C.c_int64, C.c_double, C.c_int32, C.POINTER(C.c_int64)]),
    "orc_initialize_grains": (None, [GP, DP, DP, DP, C.c_size_t, C.c_double]),
    "orc_seed_positions": (None, [C.c_uint64, C.c_size_t, C.c_double, C.c_double, DP]),
    "orc_modelc_rhs_oriented": (None, [GP, C.POINTER(ModelCParams), DP, DP, DP, DP, DP]),
    "orc_modelc_oriented_run": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, DP, C.c_int64,
                                          C.c_double, DP]),
    "orc_modelc_rk4_run": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, DP, C.c_int64, C.c_double,
                                     DP]),
    "orc_heat_rk4_run": (C.c_int, [GP, C.c_double, DP, C.c_int64, C.c_double, DP, C.POINTER(FaceBc)]),
    "orc_initialize_orientations": (None, [GP, DP, DP, C.c_size_t]),
    "orc_apply_drivers": (C.c_int, [GP, C.c_size_t, C.POINTER(DP), C.POINTER(C.c_uint32), C.c_size_t,
                                    C.POINTER(Driver), C.c_double, C.c_uint64, C.c_uint64, C.c_int32]),
    "orc_jump_diffusion_run": (C.c_int, [GP, DP, DP, DP, DP, C.c_int64, C.c_double, DP, C.c_uint64,
                                         C.c_double, C.c_double, C.c_double, C.c_int32]),
    "orc_modelc_tlm_run": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, C.c_int32, DP, DP, DP,
                                     C.c_int64, C.c_double]),
    "orc_modelc_adjoint_step": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, C.c_double, DP, DP,
                                          DP]),
    "orc_modelc_gradient": (C.c_int, [GP, C.POINTER(ModelCParams), DP, DP, DP, C.c_int64, C.c_double,
                                      DP, DP, DP, DP]),
}

_lib = None


def build() -> pathlib.Path:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def build_native(outdir: str) -> pathlib.Path:
    """The same sources built for this host's CPU (-march=native; still
    -ffp-contract=off, so the same bits) into outdir, for timing the CPU baseline on
    the machine that runs it."""
    out = pathlib.Path(outdir) / "liblegendre_oracle.so"
    subprocess.run(["make", "-s", "-C", str(HERE), "native", f"NATIVE_OUT={out}"], check=True)
    return out


def load(path) -> None:
    """Use the oracle library at `path` for every later call (before or after lib())."""
    global _lib
    h = C.CDLL(str(path))
    for name, (res, args) in _SIGS.items():
        f = getattr(h, name)
        f.restype = res
        f.argtypes = args
    _lib = h


def lib() -> C.CDLL:
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        load(LIB_PATH)
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


def modelc_branch_crossings(g: Grid, p_minus: ModelCParams, p_plus: ModelCParams, phi0, u0, nsteps,
                            dt, nthreads=1) -> int:
    """Face evaluations (four faces and the centre per cell and step) whose flat
    branch |∇φ|⁴ <= tol differs between the runs at p_minus and p_plus (b = 0)."""
    n = C.c_int64()
    rc = lib().orc_modelc_branch_crossings(C.byref(g), C.byref(p_minus), C.byref(p_plus), _dp(phi0),
                                           _dp(u0), nsteps, dt, nthreads, C.byref(n))
    if rc != 0:
        raise ValueError("orc_modelc_branch_crossings rejected the configuration")
    return int(n.value)


def heat_run(g: Grid, kappa, u, nsteps, dt, t0=0.0, bcs=None, nthreads=1) -> float:
    t = C.c_double(t0)
    rc = lib().orc_heat_run(C.byref(g), kappa, _dp(u), nsteps, dt, C.byref(t), bcs_array(bcs), nthreads)
    if rc != 0:
        raise ValueError("orc_heat_run rejected the configuration")
    return t.value


def initialize_grains(g: Grid, phi, u, grains: np.ndarray, u0: float) -> None:
    gr = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 3)
    lib().orc_initialize_grains(C.byref(g), _dp(phi), _dp(u), _dp(gr), gr.shape[0], u0)


def _dp_or_null(a):
    return None if a is None else _dp(a)


def modelc_rhs_oriented(g: Grid, p: ModelCParams, phi, u, theta, kphi, ku) -> None:
    """Multi-grain vector field (finite_volume.rs:188-235, phasefield.rs:378-392);
    theta in ghost-0 slabs."""
    lib().orc_modelc_rhs_oriented(C.byref(g), C.byref(p), _dp(phi), _dp(u), _dp(theta), _dp(kphi),
                                  _dp(ku))


def modelc_oriented_run(g: Grid, p: ModelCParams, phi, u, theta, nsteps, dt, t0=0.0) -> float:
    t = C.c_double(t0)
    if lib().orc_modelc_oriented_run(C.byref(g), C.byref(p), _dp(phi), _dp(u), _dp(theta), nsteps,
                                     dt, C.byref(t)) != 0:
        raise ValueError("orc_modelc_oriented_run rejected the configuration")
    return t.value


def modelc_rk4_run(g: Grid, p: ModelCParams, phi, u, nsteps, dt, theta=None, t0=0.0) -> float:
    """RungeKutta4 (src/integrators/rk4.rs) on Model C; theta (ghost-0 slabs) for the
    multi-grain model."""
    t = C.c_double(t0)
    if lib().orc_modelc_rk4_run(C.byref(g), C.byref(p), _dp(phi), _dp(u), _dp_or_null(theta),
                                nsteps, dt, C.byref(t)) != 0:
        raise ValueError("orc_modelc_rk4_run rejected the configuration")
    return t.value


def heat_rk4_run(g: Grid, kappa, u, nsteps, dt, t0=0.0, bcs=None) -> float:
    t = C.c_double(t0)
    if lib().orc_heat_rk4_run(C.byref(g), kappa, _dp(u), nsteps, dt, C.byref(t), bcs_array(bcs)) != 0:
        raise ValueError("orc_heat_rk4_run rejected the configuration")
    return t.value


def initialize_orientations(g: Grid, theta, grains: np.ndarray) -> None:
    """The nearest-seed orientation pass of initialize_grains; grains [cx, cy, r, θ₀]."""
    gr = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 4)
    lib().orc_initialize_orientations(C.byref(g), _dp(theta), _dp(gr), gr.shape[0])


def apply_drivers(g: Grid, xs, ghosts, drivers, dt, seed, salt, nthreads=1) -> None:
    """State::apply_step_with over the field buffers xs. drivers: (kind, index, amps, rate)
    with amps one array or None per field and rate the jump intensity (ghost-0 slabs)."""
    nf = len(xs)
    xarr = (DP * nf)(*[_dp(x) for x in xs])
    garr = (C.c_uint32 * nf)(*ghosts)
    keep = []
    darr = (Driver * len(drivers))()
    for i, (kind, index, amps, rate) in enumerate(drivers):
        a = (DP * nf)(*[_dp(v) if v is not None else DP() for v in amps])
        keep.append(a)
        darr[i] = Driver(kind, index, a, _dp(rate) if rate is not None else DP())
    rc = lib().orc_apply_drivers(C.byref(g), nf, xarr, garr, len(drivers), darr, dt, seed, salt,
                                 nthreads)
    if rc != 0:
        raise ValueError("orc_apply_drivers rejected the drivers")


def jump_diffusion_run(g: Grid, lam, w, count, jsum, nsteps, dt, seed, mu=0.1, sigma=0.5, payoff=2.0,
                       t0=0.0, nthreads=1) -> float:
    t = C.c_double(t0)
    rc = lib().orc_jump_diffusion_run(C.byref(g), _dp(lam), _dp(w), _dp(count), _dp(jsum), nsteps, dt,
                                      C.byref(t), seed, mu, sigma, payoff, nthreads)
    if rc != 0:
        raise ValueError("orc_jump_diffusion_run rejected the configuration")
    return t.value


def seed_positions(seed: int, nseeds: int, extent: float, radius: float) -> np.ndarray:
    out = np.empty((nseeds, 3), dtype=np.float64)
    lib().orc_seed_positions(seed, nseeds, extent, radius, _dp(out))
    return out


# ---- Model C derivatives (checkers for the device adjoint) ------------------------
def modelc_tlm_run(g: Grid, p: ModelCParams, phi, u, dphi, du, dtheta, nsteps, dt) -> None:
    """Forward-Euler Model C (b = 0) with tangents: values bitwise those of
    modelc_run, tangents the exact derivative.  One direction: dphi, du are field
    buffers and dtheta = (dλ, dD, dε₄); up to four at once: dphi, du of shape
    (ndir, len) and dtheta of shape (ndir, 3)."""
    th = np.ascontiguousarray(dtheta, dtype=np.float64)
    ndir = 1 if th.ndim == 1 else th.shape[0]
    assert dphi.flags["C_CONTIGUOUS"] and du.flags["C_CONTIGUOUS"]
    assert dphi.size == ndir * phi.size and du.size == ndir * u.size
    rc = lib().orc_modelc_tlm_run(C.byref(g), C.byref(p), _dp(phi), _dp(u), ndir,
                                  dphi.ctypes.data_as(DP), du.ctypes.data_as(DP), _dp(th), nsteps, dt)
    if rc != 0:
        raise ValueError("orc_modelc_tlm_run rejected the configuration")


def modelc_adjoint_step(g: Grid, p: ModelCParams, phi, u, dt, pphi, pu, grad) -> None:
    """One reverse step: (pphi, pu) <- (dS/dx)^T (pphi, pu); grad (λ, D, ε₄) += (dS/dθ)^T."""
    rc = lib().orc_modelc_adjoint_step(C.byref(g), C.byref(p), _dp(phi), _dp(u), dt, _dp(pphi),
                                       _dp(pu), _dp(grad))
    if rc != 0:
        raise ValueError("orc_modelc_adjoint_step rejected the configuration")


def modelc_gradient(g: Grid, p: ModelCParams, phi0, u0, target, nsteps, dt):
    """L = ½‖φ_T − target‖² and its gradient w.r.t. (λ, D, ε₄), φ₀ and u₀ (all states kept)."""
    loss = C.c_double(0.0)
    grad = np.zeros(3)
    gphi = np.zeros_like(phi0)
    gu = np.zeros_like(u0)
    rc = lib().orc_modelc_gradient(C.byref(g), C.byref(p), _dp(phi0), _dp(u0), _dp(target), nsteps,
                                   dt, C.byref(loss), _dp(grad), _dp(gphi), _dp(gu))
    if rc != 0:
        raise ValueError("orc_modelc_gradient rejected the configuration")
    return loss.value, grad, gphi, gu


def time_bits(t: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", t))[0]


# ---- AMR (2-D) -----------------------------------------------------------------
class RegridPolicy(C.Structure):
    _fields_ = [
        ("every", C.c_int64),
        ("buffer", C.c_int64),
        ("efficiency", C.c_double),
        ("min_width", C.c_int64),
        ("threshold", C.c_double),
        ("tag_field", C.c_int32),
    ]


_AMR_SIGS = {
    "orc_amr_create": (C.c_int, [GP, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_int64),
                                 C.c_size_t, C.POINTER(C.c_void_p)]),
    "orc_amr_destroy": (None, [C.c_void_p]),
    "orc_amr_num_blocks": (C.c_int64, [C.c_void_p]),
    "orc_amr_num_levels": (C.c_int32, [C.c_void_p]),
    "orc_amr_patch": (None, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "orc_amr_block_offset": (C.c_int64, [C.c_void_p, C.c_int64, C.c_uint32]),
    "orc_amr_field_len": (C.c_int64, [C.c_void_p, C.c_uint32]),
    "orc_amr_find_patch": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64]),
    "orc_amr_cell_center": (None, [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, DP]),
    "orc_amr_restrict_level": (None, [C.c_void_p, DP, C.c_uint32, C.c_int32]),
    "orc_amr_restrict": (None, [C.c_void_p, DP, C.c_uint32]),
    "orc_amr_fill_level": (None, [C.c_void_p, DP, DP, C.c_double, C.c_uint32, C.c_int32]),
    "orc_amr_fill_ghosts": (None, [C.c_void_p, DP, C.c_uint32]),
    "orc_cluster": (C.c_size_t, [C.POINTER(C.c_int64), C.c_size_t, C.c_double, C.c_size_t,
                                 C.POINTER(C.c_int64), C.c_size_t]),
    "orc_amr_sim_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ModelCParams), C.c_double,
                                     C.c_int32, C.c_uint64, C.c_int32, C.POINTER(RegridPolicy),
                                     C.POINTER(C.c_void_p)]),
    "orc_amr_sim_destroy": (None, [C.c_void_p]),
    "orc_amr_sim_grid": (C.c_void_p, [C.c_void_p]),
    "orc_amr_sim_field": (DP, [C.c_void_p, C.c_int32]),
    "orc_amr_sim_time": (C.c_double, [C.c_void_p]),
    "orc_amr_sim_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int64]),
    "orc_amr_sim_set_tagger": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_tag_sink_push": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_size_t]),
    "orc_nesting_parents": (None, [C.POINTER(C.c_int64), C.c_int64, C.c_int64, C.POINTER(C.c_int64)]),
}

# int tag_level(void *user, const orc_amr *grid, int32_t level, orc_tag_sink *sink)
TAG_LEVEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p)


def nesting_parents(box, rf, rs):
    """The source-level cells the nesting flags of one box cover (adapt.rs:194-211),
    unclipped: box = ((lo0, lo1), (hi0, hi1)) in its own level's cells."""
    (lo, hi) = box
    b = (C.c_int64 * 4)(lo[0], lo[1], hi[0], hi[1])
    out = (C.c_int64 * 4)()
    _amr_lib().orc_nesting_parents(b, rf, rs, out)
    return (int(out[0]), int(out[1])), (int(out[2]), int(out[3]))


def _amr_lib():
    h = lib()
    if not getattr(h, "_amr_bound", False):
        for name, (res, args) in _AMR_SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        h._amr_bound = True
    return h


class AmrGrid:
    """AmrGrid<2> (src/geometry/amr/grid.rs) — a view over an oracle handle."""

    def __init__(self, handle, owned=True):
        self.h = handle
        self.owned = owned

    @classmethod
    def from_patches(cls, base: Grid, ratios, levels):
        """levels[l] = list of ((lo0, lo1), (hi0, hi1)) boxes of level l+1."""
        L = _amr_lib()
        boxes = []
        for l, bl in enumerate(levels):
            for (lo, hi) in bl:
                boxes += [l + 1, lo[0], lo[1], hi[0], hi[1]]
        r = (C.c_uint32 * max(1, len(ratios)))(*ratios)
        b = (C.c_int64 * max(1, len(boxes)))(*boxes)
        h = C.c_void_p()
        rc = L.orc_amr_create(C.byref(base), r, len(ratios), b, len(boxes) // 5, C.byref(h))
        if rc != 0:
            raise ValueError(f"AmrGrid::from_patches error {rc}")
        return cls(h)

    def __del__(self):
        if getattr(self, "owned", False) and self.h:
            _amr_lib().orc_amr_destroy(self.h)
            self.h = None

    @property
    def num_blocks(self):
        return int(_amr_lib().orc_amr_num_blocks(self.h))

    @property
    def num_levels(self):
        return int(_amr_lib().orc_amr_num_levels(self.h))

    def patch(self, b):
        out = (C.c_int64 * 5)()
        _amr_lib().orc_amr_patch(self.h, b, out)
        return int(out[0]), (int(out[1]), int(out[2])), (int(out[3]), int(out[4]))

    def blocks_at(self, level):
        return [b for b in range(self.num_blocks) if self.patch(b)[0] == level]

    def offset(self, b, ghost=1):
        return int(_amr_lib().orc_amr_block_offset(self.h, b, ghost))

    def field_len(self, ghost=1):
        return int(_amr_lib().orc_amr_field_len(self.h, ghost))

    def find_patch(self, level, cell):
        r = int(_amr_lib().orc_amr_find_patch(self.h, level, cell[0], cell[1]))
        return None if r < 0 else r

    def cell_center(self, b, i, j):
        out = (C.c_double * 2)()
        _amr_lib().orc_amr_cell_center(self.h, b, i, j, out)
        return out[0], out[1]

    def view(self, field: np.ndarray, b: int, ghost: int = 1) -> np.ndarray:
        """(ny+2g, nx+2g) view of block b's slab."""
        _, lo, hi = self.patch(b)
        nx, ny = hi[0] - lo[0], hi[1] - lo[1]
        o = self.offset(b, ghost)
        return field[o : o + (nx + 2 * ghost) * (ny + 2 * ghost)].reshape(ny + 2 * ghost, nx + 2 * ghost)

    def fill_by_center(self, field, f, ghost=1):
        for b in range(self.num_blocks):
            v = self.view(field, b, ghost)
            ny, nx = v.shape[0] - 2 * ghost, v.shape[1] - 2 * ghost
            for j in range(ny):
                for i in range(nx):
                    v[j + ghost, i + ghost] = f(self.cell_center(b, i, j))

    def restrict(self, field, ghost=1):
        _amr_lib().orc_amr_restrict(self.h, _dp(field), ghost)

    def restrict_level(self, field, fine_level, ghost=1):
        _amr_lib().orc_amr_restrict_level(self.h, _dp(field), ghost, fine_level)

    def fill_ghosts(self, field, ghost=1):
        _amr_lib().orc_amr_fill_ghosts(self.h, _dp(field), ghost)

    def fill_level(self, field, level, old=None, alpha=0.0, ghost=1):
        _amr_lib().orc_amr_fill_level(self.h, _dp(field), None if old is None else _dp(old), alpha,
                                      ghost, level)


def cluster(flags, efficiency, min_width):
    """Berger–Rigoutsos (cluster.rs:144-155); returns boxes ((lo0, lo1), (hi0, hi1))."""
    L = _amr_lib()
    n = len(flags)
    f = (C.c_int64 * max(1, 2 * n))(*[v for c in flags for v in c])
    out = (C.c_int64 * max(4, 4 * n))()
    k = int(L.orc_cluster(f, n, efficiency, min_width, out, max(1, n)))
    return [((out[4 * i], out[4 * i + 1]), (out[4 * i + 2], out[4 * i + 3])) for i in range(k)]


class AmrSim:
    """Simulation on an AmrGrid (model 'heat' or 'modelc'; integ 'global' or 'subcycle')."""

    def __init__(self, grid: AmrGrid, model="heat", params=None, kappa=0.0, integ="global",
                 seed=0, wiener=False, policy=None):
        L = _amr_lib()
        grid.owned = False  # ownership moves to the simulation
        h = C.c_void_p()
        pol = None
        if policy is not None:
            pol = RegridPolicy(*policy) if not isinstance(policy, RegridPolicy) else policy
        rc = L.orc_amr_sim_create(grid.h, {"heat": 0, "modelc": 1}[model],
                                  C.byref(params) if params is not None else None, kappa,
                                  {"global": 0, "subcycle": 1}[integ], seed, int(wiener),
                                  C.byref(pol) if pol is not None else None, C.byref(h))
        if rc != 0:
            raise ValueError("orc_amr_sim_create failed")
        self.h = h
        self.nfields = 2 if model == "modelc" else 1

    def __del__(self):
        if getattr(self, "h", None):
            _amr_lib().orc_amr_sim_destroy(self.h)
            self.h = None

    @property
    def grid(self) -> AmrGrid:
        return AmrGrid(C.c_void_p(_amr_lib().orc_amr_sim_grid(self.h)), owned=False)

    def field(self, f=0) -> np.ndarray:
        """Live numpy view of the current field buffer (valid until the next step)."""
        n = self.grid.field_len(1)
        ptr = _amr_lib().orc_amr_sim_field(self.h, f)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def time(self):
        return float(_amr_lib().orc_amr_sim_time(self.h))

    def step(self, dt, n=1):
        if _amr_lib().orc_amr_sim_step(self.h, dt, n) != 0:
            raise ValueError("regrid failed: the tagger failed or the proposal is not a valid hierarchy")

    def set_tagger(self, fn):
        """A caller's TagCells::tag_level in place of the GradientTagger (prepare is
        the trait default): fn(grid: AmrGrid, level) -> iterable of (x, y) cells."""
        if fn is None:
            self._tagger = None
            _amr_lib().orc_amr_sim_set_tagger(self.h, None, None)
            return

        def cb(user, grid, level, sink):
            try:
                cells = [c for xy in fn(AmrGrid(C.c_void_p(grid), owned=False), int(level)) for c in xy]
            except Exception:  # noqa: BLE001 - a failing tagger aborts the regrid
                return 1
            buf = (C.c_int64 * max(1, len(cells)))(*cells)
            return _amr_lib().orc_tag_sink_push(sink, buf, len(cells) // 2)

        self._tagger = TAG_LEVEL_FN(cb)  # kept alive with the simulation
        _amr_lib().orc_amr_sim_set_tagger(self.h, C.cast(self._tagger, C.c_void_p), None)
