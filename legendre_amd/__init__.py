"""legendre_amd — Python handle on the MI355X backend's C ABI (include/legendre_gpu.h).

Thin plumbing over ctypes: every call goes to liblegendre_gpu.so, whose kernels
run on the GPU.  Names follow the reference crate (CartesianGrid, FaceBc,
ModelC, Simulation, ...) so that tests read like the reference's own tests.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import struct
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import (
    LGD_BC_DIRICHLET,
    LGD_BC_EXTRAPOLATE,
    LGD_BC_FLUX,
    LGD_BC_MIRROR,
    LGD_DRIVER_JUMP,
    LGD_DRIVER_TIME,
    LGD_DRIVER_WIENER,
    LGD_INTEGRATOR_EULER,
    LGD_INTEGRATOR_RK4,
    LGD_MODE_FUSED,
    LGD_MODE_PER_INTENT,
    LegendreError,
)

__all__ = [
    "A2",
    "CartesianGrid",
    "Context",
    "DeviceBuffer",
    "FaceBc",
    "LegendreError",
    "ModelCParams",
    "Simulation",
    "AmrGrid",
    "RegridPolicy",
    "time_bits",
    "LGD_MODE_FUSED",
    "LGD_MODE_PER_INTENT",
    "LGD_INTEGRATOR_EULER",
    "LGD_INTEGRATOR_RK4",
    "LGD_DRIVER_TIME",
    "LGD_DRIVER_WIENER",
    "LGD_DRIVER_JUMP",
    "Driver",
    "ParquetObserver",
    "write_parquet",
]

# Karma–Rappel thin-interface constant a₂ (src/physics/phasefield.rs:44).
A2 = 0.6267


def time_bits(t: float) -> int:
    """f64::to_bits — the noise salt of a step starting at time t."""
    return struct.unpack("<Q", struct.pack("<d", t))[0]


class Context:
    """A device + stream (lgd_ctx).  `stream` may be an existing hipStream_t
    handle (an int, e.g. torch.cuda.Stream().cuda_stream)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self._lib = _abi.lib()
        h = C.c_void_p()
        st = self._lib.lgd_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h))
        if st != 0:
            raise LegendreError(st, f"could not create a context on device {device}")
        self.handle = h
        self._children = []  # weak references, creation order

    def _adopt(self, obj) -> None:
        """Register an object that holds device resources of this context; closing
        the context closes it first (no library object may outlive its context)."""
        self._children.append(weakref.ref(obj))

    def sync(self) -> None:
        _abi.check(self._lib.lgd_ctx_sync(self.handle), self.handle)

    @property
    def stream(self) -> int:
        return self._lib.lgd_ctx_stream(self.handle) or 0

    def set_tuning(self, key: str, value: int) -> None:
        """Launch-shape knob (results are bitwise independent of it)."""
        _abi.check(self._lib.lgd_ctx_set_tuning(self.handle, key.encode(), int(value)), self.handle)

    def autotune_modelc_step(self, grid: "CartesianGrid", params: "ModelCParams", phi: "DeviceBuffer",
                             u: "DeviceBuffer", dt: float, wiener: bool = False, reps: int = 5) -> dict:
        """Time every launch shape of the fused Model C step on (phi, u) and keep
        the fastest for this grid shape (used while no knob is set)."""
        r = _abi.TuningResultC()
        pc = params.c()
        _abi.check(self._lib.lgd_autotune_modelc_step(self.handle, grid.handle, C.byref(pc), phi.handle,
                                                      u.handle, dt, int(wiener), int(reps), C.byref(r)),
                   self.handle)
        return {"kernel": {1: "percell", 2: "stream", 3: "pair"}[r.kernel], "strips": r.strips,
                "prefetch": r.prefetch, "ms": r.ms, "default_ms": r.default_ms,
                "candidates": r.candidates}

    def close(self) -> None:
        if self.handle:
            for ref in reversed(self._children):  # dependents before what they use
                obj = ref()
                if obj is not None:
                    try:
                        obj.close()
                    except Exception:
                        pass
            self._children.clear()
            self._lib.lgd_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device storage from the DeviceAllocator (zero-initialised f64)."""

    def __init__(self, ctx: Context, length: int):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _abi.check(self._lib.lgd_buf_alloc(ctx.handle, int(length), C.byref(h)), ctx.handle)
        self.handle = h
        self.length = int(length)
        self.owned = True
        ctx._adopt(self)

    @classmethod
    def _borrow(cls, ctx: Context, handle: C.c_void_p) -> "DeviceBuffer":
        self = cls.__new__(cls)
        self.ctx = ctx
        self._lib = ctx._lib
        self.handle = handle
        self.length = int(ctx._lib.lgd_buf_len(handle))
        self.owned = False
        return self

    def __len__(self) -> int:
        return self.length

    def upload(self, host: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(host, dtype=np.float64).ravel()
        _abi.check(
            self._lib.lgd_buf_upload(
                self.ctx.handle, self.handle, offset, a.ctypes.data_as(_abi.DP), a.size
            ),
            self.ctx.handle,
        )

    def download(self, offset: int = 0, n: Optional[int] = None) -> np.ndarray:
        n = self.length - offset if n is None else n
        out = np.empty(n, dtype=np.float64)
        _abi.check(
            self._lib.lgd_buf_download(
                self.ctx.handle, self.handle, offset, out.ctypes.data_as(_abi.DP), n
            ),
            self.ctx.handle,
        )
        return out

    def fill_zero(self) -> None:
        _abi.check(self._lib.lgd_buf_fill_zero(self.ctx.handle, self.handle), self.ctx.handle)

    def copy_from(self, other: "DeviceBuffer") -> None:
        _abi.check(self._lib.lgd_buf_copy(self.ctx.handle, self.handle, other.handle), self.ctx.handle)

    def free(self) -> None:
        if self.owned and self.handle:
            self._lib.lgd_buf_free(self.handle)
        self.handle = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclasses.dataclass(frozen=True)
class FaceBc:
    """Per-face physical boundary condition (src/geometry/boundary.rs:36-54)."""

    kind: int = LGD_BC_MIRROR
    value: float = 0.0

    @staticmethod
    def mirror() -> "FaceBc":
        return FaceBc(LGD_BC_MIRROR, 0.0)

    @staticmethod
    def dirichlet(v: float) -> "FaceBc":
        return FaceBc(LGD_BC_DIRICHLET, v)

    @staticmethod
    def flux(g: float) -> "FaceBc":
        return FaceBc(LGD_BC_FLUX, g)

    @staticmethod
    def extrapolate() -> "FaceBc":
        return FaceBc(LGD_BC_EXTRAPOLATE, 0.0)


def _bcs_array(bcs: Optional[Sequence[FaceBc]]):
    if bcs is None:
        return None
    arr = (_abi.FaceBcC * len(bcs))()
    for i, b in enumerate(bcs):
        arr[i].kind = b.kind
        arr[i].value = b.value
    return arr


class CartesianGrid:
    """CartesianGrid<D> (src/geometry/cartesian.rs:28-97) as a device block table."""

    def __init__(self, ctx: Context, cells, block_cells, origin, spacing, periodic=None):
        dim = len(cells)
        self.ctx = ctx
        self._lib = ctx._lib
        self.dim = dim
        self.cells = tuple(int(c) for c in cells)
        self.block_cells = tuple(int(c) for c in block_cells)
        self.origin = tuple(float(o) for o in origin)
        self.spacing = tuple(float(s) for s in spacing)
        self.periodic = tuple(bool(p) for p in (periodic or [False] * dim))
        d = _abi.CartesianDesc()
        d.dim = dim
        for i in range(dim):
            d.cells[i] = self.cells[i]
            d.block_cells[i] = self.block_cells[i]
            d.origin[i] = self.origin[i]
            d.spacing[i] = self.spacing[i]
            d.periodic[i] = int(self.periodic[i])
        self.desc = d
        h = C.c_void_p()
        _abi.check(self._lib.lgd_grid_create_cartesian(ctx.handle, C.byref(d), C.byref(h)), ctx.handle)
        self.handle = h
        ctx._adopt(self)

    @property
    def num_blocks(self) -> int:
        return int(self._lib.lgd_grid_num_blocks(self.handle))

    def block_len(self, ghost: int) -> int:
        return int(self._lib.lgd_grid_block_len(self.handle, ghost))

    def field_len(self, ghost: int) -> int:
        return self.num_blocks * self.block_len(ghost)

    def face_neighbor(self, block: int, dim: int, direction: int) -> Optional[int]:
        r = int(self._lib.lgd_grid_face_neighbor(self.handle, block, dim, direction))
        return None if r < 0 else r

    def cell_key(self, ghost: int, offset: int) -> Optional[int]:
        r = int(self._lib.lgd_grid_cell_key(self.handle, ghost, offset))
        return None if r < 0 else r

    def alloc_field(self, ghost: int = 1) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, self.field_len(ghost))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.lgd_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- per-intent kernels -------------------------------------------------
    def fill_ghosts(self, field: DeviceBuffer, ghost: int = 1, bcs=None) -> None:
        arr = _bcs_array(bcs)
        _abi.check(
            self._lib.lgd_fill_ghosts(self.ctx.handle, self.handle, field.handle, ghost, arr),
            self.ctx.handle,
        )

    def central_laplacian(self, inp: DeviceBuffer, out: DeviceBuffer, ghost: int = 1) -> None:
        _abi.check(
            self._lib.lgd_central_laplacian(self.ctx.handle, self.handle, inp.handle, out.handle, ghost),
            self.ctx.handle,
        )

    def karma_rappel_flux(self, phi: DeviceBuffer, out: DeviceBuffer, eps4: float, tol: float = 1e-8):
        _abi.check(
            self._lib.lgd_karma_rappel_flux(self.ctx.handle, self.handle, phi.handle, out.handle, eps4, tol),
            self.ctx.handle,
        )

    def heat_rhs(self, kappa: float, u: DeviceBuffer, k: DeviceBuffer) -> None:
        _abi.check(
            self._lib.lgd_heat_rhs(self.ctx.handle, self.handle, kappa, u.handle, k.handle),
            self.ctx.handle,
        )

    def modelc_rhs(self, params: "ModelCParams", phi, u, kphi, ku) -> None:
        p = params.c()
        _abi.check(
            self._lib.lgd_modelc_rhs(
                self.ctx.handle, self.handle, C.byref(p), phi.handle, u.handle, kphi.handle, ku.handle
            ),
            self.ctx.handle,
        )

    def modelc_rhs_oriented(self, params: "ModelCParams", phi, u, cos0, sin0, kphi, ku) -> None:
        """The multi-grain vector field; cos0, sin0 = cos θ₀, sin θ₀ in ghost-0 slabs."""
        p = params.c()
        _abi.check(
            self._lib.lgd_modelc_rhs_oriented(self.ctx.handle, self.handle, C.byref(p), phi.handle,
                                              u.handle, cos0.handle, sin0.handle, kphi.handle,
                                              ku.handle),
            self.ctx.handle,
        )

    def modelc_amplitude(self, params: "ModelCParams", amp: DeviceBuffer) -> None:
        p = params.c()
        _abi.check(
            self._lib.lgd_modelc_amplitude(self.ctx.handle, self.handle, C.byref(p), amp.handle),
            self.ctx.handle,
        )

    def axpy(self, x: DeviceBuffer, y: DeviceBuffer, alpha: float) -> None:
        _abi.check(self._lib.lgd_axpy(self.ctx.handle, x.handle, y.handle, alpha), self.ctx.handle)

    def apply_wiener(self, x, amp, ghost, dt, seed, salt, stream) -> None:
        _abi.check(
            self._lib.lgd_apply_wiener(
                self.ctx.handle, self.handle, x.handle, amp.handle, ghost, dt, seed, salt, stream
            ),
            self.ctx.handle,
        )

    def modelc_step(self, params, phi, u, phi_next, u_next, dt, seed, salt, wiener) -> None:
        p = params.c()
        _abi.check(
            self._lib.lgd_modelc_step(
                self.ctx.handle, self.handle, C.byref(p), phi.handle, u.handle, phi_next.handle,
                u_next.handle, dt, seed, salt, int(wiener)
            ),
            self.ctx.handle,
        )

    def heat_step(self, kappa, u, u_next, dt, bcs=None) -> None:
        if bcs is None:
            rc = self._lib.lgd_heat_step(self.ctx.handle, self.handle, kappa, u.handle, u_next.handle, dt)
        else:
            rc = self._lib.lgd_heat_step_bc(self.ctx.handle, self.handle, kappa, _bcs_array(bcs),
                                            u.handle, u_next.handle, dt)
        _abi.check(rc, self.ctx.handle)

    def apply_drivers(self, xs, ghosts, drivers, dt, seed, salt) -> None:
        """State::apply_step_with over the field buffers xs (ghost widths ghosts):
        drivers is a sequence of Driver, applied in order."""
        nf = len(xs)
        xarr = (C.c_void_p * nf)(*[x.handle for x in xs])
        garr = (C.c_uint32 * nf)(*ghosts)
        keep = []
        darr = (_abi.DriverC * len(drivers))()
        for i, d in enumerate(drivers):
            if len(d.fields) != nf:
                raise ValueError("a driver names one buffer (or None) per field")
            a = (C.c_void_p * nf)(*[f.handle if f is not None else None for f in d.fields])
            keep.append(a)
            darr[i] = _abi.DriverC(d.kind, d.index, C.cast(a, _abi.PP),
                                   d.rate.handle if d.rate is not None else None)
        _abi.check(
            self._lib.lgd_apply_drivers(self.ctx.handle, self.handle, nf, C.cast(xarr, _abi.PP), garr,
                                        len(drivers), darr, dt, seed, salt),
            self.ctx.handle,
        )


@dataclasses.dataclass
class Driver:
    """One driver of an explicit update (src/core/driver.rs:35-50): kind LGD_DRIVER_*,
    index j of Wiener(j) / Jump(j), per field the evaluated buffer (dY/dt, amplitude
    or jump increment) or None, and a jump driver's per-cell intensity (ghost 0)."""

    kind: int
    index: int
    fields: Sequence[Optional["DeviceBuffer"]]
    rate: Optional["DeviceBuffer"] = None


@dataclasses.dataclass
class ModelCParams:
    """ModelC public fields + FiniteVolume::tol (src/physics/phasefield.rs:55-116)."""

    eps4: float
    lambda_: float
    d_thermal: float
    noise_amplitude: float = 0.0
    anisotropy_tol: float = 1e-8
    fv_tol: float = 1e-8

    @staticmethod
    def new(eps4: float, lambda_: float, noise_amplitude: float) -> "ModelCParams":
        """ModelC::new: D = a₂λ (phasefield.rs:100-116)."""
        return ModelCParams(eps4, lambda_, A2 * lambda_, noise_amplitude)

    @staticmethod
    def classic() -> "ModelCParams":
        """ModelC::classic: ε₄ = 0.06, λ = 3.19, no noise (phasefield.rs:93-95)."""
        return ModelCParams.new(0.06, 3.19, 0.0)

    def c(self) -> _abi.ModelCParamsC:
        p = _abi.ModelCParamsC()
        p.eps4 = self.eps4
        p.lambda_ = self.lambda_
        p.d_thermal = self.d_thermal
        p.noise_amplitude = self.noise_amplitude
        p.anisotropy_tol = self.anisotropy_tol
        p.fv_tol = self.fv_tol
        return p

    def stable_dt(self, h: float) -> float:
        """ModelC::stable_dt (phasefield.rs:446-449)."""
        return 0.2 * h**2 / self.d_thermal


class Simulation:
    """Simulation (src/core/simulation.rs) driven by the GPU backend."""

    def __init__(self, grid: Optional[CartesianGrid], handle: C.c_void_p, nfields: int, ctx=None):
        self.grid = grid
        self.ctx = grid.ctx if grid is not None else ctx
        self._lib = self.ctx._lib
        self.handle = handle
        self.nfields = nfields
        self.ctx._adopt(self)

    @classmethod
    def model_c(cls, grid: CartesianGrid, params: ModelCParams, wiener: bool = False,
                seed: int = 0, mode: int = LGD_MODE_FUSED) -> "Simulation":
        p = params.c()
        h = C.c_void_p()
        _abi.check(
            grid._lib.lgd_sim_create_modelc(grid.ctx.handle, grid.handle, C.byref(p), int(wiener),
                                            seed, mode, C.byref(h)),
            grid.ctx.handle,
        )
        return cls(grid, h, 2)

    @classmethod
    def heat(cls, grid: CartesianGrid, kappa: float, bcs=None, mode: int = LGD_MODE_FUSED):
        arr = _bcs_array(bcs)
        h = C.c_void_p()
        _abi.check(
            grid._lib.lgd_sim_create_heat(grid.ctx.handle, grid.handle, kappa, arr, mode, C.byref(h)),
            grid.ctx.handle,
        )
        return cls(grid, h, 1)

    def field(self, index: int) -> DeviceBuffer:
        h = C.c_void_p()
        _abi.check(self._lib.lgd_sim_field(self.handle, index, C.byref(h)), self.ctx.handle)
        return DeviceBuffer._borrow(self.ctx, h)

    def step(self, dt: float, nsteps: int = 1) -> None:
        _abi.check(self._lib.lgd_sim_step(self.handle, dt, nsteps), self.ctx.handle)

    def regrid_stats(self) -> tuple:
        """(regrids, seconds by phase: tag field to host, proposal, new hierarchy,
        migration)."""
        n = C.c_int64()
        sec = np.zeros(4)
        _abi.check(self._lib.lgd_sim_regrid_stats(self.handle, C.byref(n), sec.ctypes.data_as(_abi.DP)),
                   self.ctx.handle)
        return int(n.value), sec

    def autotune_amr(self, dt: float, reps: int = 5) -> dict:
        """Time the AMR launch shapes on the current hierarchy and state (streaming
        level-stage strip heights and per-cell tiles, fill / restriction widths, level-0
        kernel and strips) and keep the fastest in the context; the simulation's results
        do not change."""
        r = _abi.AmrTuningResultC()
        _abi.check(self._lib.lgd_sim_autotune_amr(self.handle, dt, int(reps), C.byref(r)), self.ctx.handle)
        return {"stage": "streaming" if r.stage_rows else "per-cell", "stage_rows": r.stage_rows,
                "tile": (r.tile_x, r.tile_y), "wg": r.wg, "level0_kernel": r.level0_kernel,
                "level0_strips": r.level0_strips, "level0_prefetch": r.level0_prefetch,
                "candidates": r.candidates, "ms": r.ms, "default_ms": r.default_ms}

    def set_tagger(self, fn) -> None:
        """A caller's TagCells::tag_level in place of the policy's GradientTagger
        (adaptive AMR only; prepare is the trait default): fn(amr: AmrGrid, level)
        returns the level's cells needing refinement as (x, y) pairs.  None restores
        the GradientTagger."""
        if fn is None:
            self._tagger = None
            _abi.check(self._lib.lgd_sim_set_tagger(self.handle, None, None), self.ctx.handle)
            return
        lib = self._lib
        ctx = self.ctx

        def cb(user, grid, level, sink):
            try:
                cells = [c for xy in fn(AmrGrid(ctx, C.c_void_p(grid), owned=False), int(level)) for c in xy]
            except Exception:  # noqa: BLE001 - a failing tagger fails the step
                return 1
            buf = (C.c_int64 * max(1, len(cells)))(*cells)
            return lib.lgd_tag_sink_push(sink, buf, len(cells) // 2)

        self._tagger = _abi.TAG_LEVEL_FN(cb)  # kept alive with the simulation
        _abi.check(self._lib.lgd_sim_set_tagger(self.handle, C.cast(self._tagger, C.c_void_p), None),
                   self.ctx.handle)

    def set_integrator(self, name: str) -> None:
        """'euler' (EulerMaruyama / ForwardEuler) or 'rk4' (RungeKutta4, deterministic
        models only)."""
        code = {"euler": LGD_INTEGRATOR_EULER, "rk4": LGD_INTEGRATOR_RK4}[name]
        _abi.check(self._lib.lgd_sim_set_integrator(self.handle, code), self.ctx.handle)

    def set_orientation(self, theta0: Optional[np.ndarray]) -> None:
        """Multi-grain Model C: θ₀ per interior cell (ghost-0 slabs; None = 0)."""
        if theta0 is None:
            _abi.check(self._lib.lgd_sim_set_orientation(self.handle, None, 0), self.ctx.handle)
            return
        t = np.ascontiguousarray(theta0, dtype=np.float64)
        _abi.check(self._lib.lgd_sim_set_orientation(self.handle, t.ctypes.data_as(_abi.DP), t.size),
                   self.ctx.handle)

    def initialize_grains_oriented(self, grains: np.ndarray, u0: float) -> None:
        """Grains [cx, cy, r, θ₀]: φ, u as initialize_grains, θ₀ by nearest seed."""
        g = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 4)
        _abi.check(self._lib.lgd_sim_initialize_grains_oriented(self.handle, g.ctypes.data_as(_abi.DP),
                                                                g.shape[0], u0), self.ctx.handle)

    def attach_observer(self, observer: "ParquetObserver") -> None:
        """Simulation::attach_observer: the observer sees every completed step."""
        _abi.check(self._lib.lgd_sim_attach_observer(self.handle, observer.handle), self.ctx.handle)
        self._observers = getattr(self, "_observers", []) + [observer]
        observer._sims.append(weakref.ref(self))

    def detach_observer(self, observer: "ParquetObserver") -> None:
        _abi.check(self._lib.lgd_sim_detach_observer(self.handle, observer.handle), self.ctx.handle)
        self._observers = [o for o in getattr(self, "_observers", []) if o is not observer]

    def set_graph_steps(self, steps: int) -> None:
        """Run `steps` whole steps at a time as one replayed HIP graph (0: eager
        launches).  Results are bitwise those of eager steps."""
        _abi.check(self._lib.lgd_sim_set_graph_steps(self.handle, int(steps)), self.ctx.handle)

    @property
    def time(self) -> float:
        return float(self._lib.lgd_sim_time(self.handle))

    @property
    def step_index(self) -> int:
        return int(self._lib.lgd_sim_step_index(self.handle))

    def initialize_grains(self, grains: np.ndarray, u0: float) -> None:
        g = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 3)
        _abi.check(
            self._lib.lgd_sim_initialize_grains(self.handle, g.ctypes.data_as(_abi.DP), g.shape[0], u0),
            self.ctx.handle,
        )

    def close(self) -> None:
        if self.handle:
            self._lib.lgd_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParquetObserver:
    """ParquetObserver behind an AsyncObserver snapshot ring (src/io/parquet.rs,
    src/core/observer.rs): fields = [(column name, simulation field)], statics the
    same for per-epoch static columns; snapshots on step 1 and every `every` steps."""

    def __init__(self, ctx: Context, directory: str, fields, every: int = 1, ring: int = 2,
                 statics=(), blocking: bool = False):
        self.ctx = ctx
        self._lib = ctx._lib
        self.directory = str(directory)
        self.names = [n for n, _ in fields]
        self.static_names = [n for n, _ in statics]
        self._keep = [n.encode() for n, _ in fields] + [n.encode() for n, _ in statics]
        farr = (_abi.ObservedFieldC * len(fields))(*[_abi.ObservedFieldC(n.encode(), int(i))
                                                      for n, i in fields])
        sarr = (_abi.ObservedFieldC * max(1, len(statics)))(*[_abi.ObservedFieldC(n.encode(), int(i))
                                                              for n, i in statics])
        h = C.c_void_p()
        _abi.check(self._lib.lgd_observer_create(ctx.handle, self.directory.encode(), int(every),
                                                 int(ring), farr, len(fields), sarr, len(statics),
                                                 C.byref(h)), ctx.handle)
        self.handle = h
        self._sims = []  # simulations it is attached to (weak), detached on close
        ctx._adopt(self)
        if blocking:
            _abi.check(self._lib.lgd_observer_set_blocking(h, 1), ctx.handle)

    def observe(self, grid: "CartesianGrid", step: int, t: float, epoch: int, fields, ghosts=None,
                statics=(), static_ghosts=None) -> None:
        """Observer::observe with device buffers in declaration order."""
        nf, ns = len(fields), len(statics)
        ghosts = list(ghosts) if ghosts is not None else [1] * nf
        static_ghosts = list(static_ghosts) if static_ghosts is not None else [0] * ns
        f = (C.c_void_p * max(1, nf))(*[b.handle for b in fields])
        g = (C.c_uint32 * max(1, nf))(*ghosts)
        sf = (C.c_void_p * max(1, ns))(*[b.handle for b in statics])
        sg = (C.c_uint32 * max(1, ns))(*static_ghosts)
        _abi.check(self._lib.lgd_observer_observe(self.handle, grid.handle, int(step), float(t), int(epoch),
                                                  C.cast(f, _abi.PP), g, C.cast(sf, _abi.PP), sg),
                   self.ctx.handle)

    def observe_amr(self, amr: "AmrGrid", step: int, t: float, epoch: int, fields) -> None:
        f = (C.c_void_p * max(1, len(fields)))(*[b.handle for b in fields])
        _abi.check(self._lib.lgd_observer_observe_amr(self.handle, amr.handle, int(step), float(t),
                                                      int(epoch), C.cast(f, _abi.PP)), self.ctx.handle)

    def flush(self) -> None:
        _abi.check(self._lib.lgd_observer_flush(self.handle), self.ctx.handle)

    def pause(self, paused: bool = True) -> None:
        _abi.check(self._lib.lgd_observer_pause(self.handle, int(paused)), self.ctx.handle)

    def stats(self) -> dict:
        w, d, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        sec = C.c_double()
        _abi.check(self._lib.lgd_observer_stats(self.handle, C.byref(w), C.byref(d), C.byref(f),
                                                C.byref(sec)), self.ctx.handle)
        buf = C.create_string_buffer(512)
        self._lib.lgd_observer_last_error(self.handle, buf, 512)
        return {"written": w.value, "dropped": d.value, "failed": f.value, "write_seconds": sec.value,
                "last_error": buf.value.decode()}

    def close(self) -> None:
        if self.handle:
            for ref in self._sims:
                sim = ref()
                if sim is not None and sim.handle:
                    self._lib.lgd_sim_detach_observer(sim.handle, self.handle)
            self._sims = []
            self._lib.lgd_observer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_parquet(path: str, columns, group_rows=None) -> None:
    """Host columns (name -> float64 / int64 array, or a Python float / int for a
    constant column) as one Parquet file in the snapshot layout."""
    names = list(columns)
    arrays, types, consts = [], [], []
    nrows = None
    for n in names:
        v = columns[n]
        if isinstance(v, (int, np.integer)) and not isinstance(v, bool):
            types.append(1)
            consts.append(int(v) & ((1 << 64) - 1))
            arrays.append(None)
        elif isinstance(v, (float, np.floating)):
            types.append(0)
            consts.append(struct.unpack("<Q", struct.pack("<d", float(v)))[0])
            arrays.append(None)
        else:
            a = np.ascontiguousarray(v)
            if a.dtype not in (np.float64, np.int64):
                raise ValueError(f"column {n}: float64 or int64 expected")
            types.append(0 if a.dtype == np.float64 else 1)
            consts.append(0)
            arrays.append(a)
            nrows = a.size if nrows is None else nrows
            if a.size != nrows:
                raise ValueError("columns differ in length")
    nrows = nrows or 0
    groups = list(group_rows) if group_rows is not None else ([nrows] if nrows else [])
    lib = _abi.lib()
    nc = len(names)
    name_arr = (C.c_char_p * nc)(*[n.encode() for n in names])
    type_arr = (C.c_int32 * nc)(*types)
    col_arr = (C.c_void_p * nc)(*[a.ctypes.data if a is not None else None for a in arrays])
    const_arr = (C.c_uint64 * nc)(*consts)
    g_arr = (C.c_int64 * max(1, len(groups)))(*groups)
    err = C.create_string_buffer(512)
    st = lib.lgd_parquet_write_file(str(path).encode(), nc, name_arr, type_arr, C.cast(col_arr, _abi.PP),
                                    const_arr, nrows, g_arr, len(groups), err, 512)
    if st != 0:
        raise LegendreError(st, err.value.decode())


@dataclasses.dataclass(frozen=True)
class RegridPolicy:
    """BergerOliger + GradientTagger knobs (src/geometry/amr/adapt.rs:55-129)."""

    every: int = 4
    buffer: int = 2
    efficiency: float = 0.8
    min_width: int = 4
    threshold: float = 0.1
    tag_field: int = 0

    def c(self) -> _abi.RegridPolicyC:
        return _abi.RegridPolicyC(self.every, self.buffer, self.efficiency, self.min_width,
                                  self.threshold, self.tag_field)


class AmrGrid:
    """AmrGrid<2> (src/geometry/amr/grid.rs) on the device."""

    def __init__(self, ctx: Context, handle, owned: bool = True):
        self.ctx = ctx
        self._lib = ctx._lib
        self.handle = handle
        self.owned = owned
        if owned:
            ctx._adopt(self)

    @classmethod
    def from_patches(cls, ctx: Context, cells, block_cells, origin, spacing, ratios, levels):
        """levels[l] = list of ((lo0, lo1), (hi0, hi1)) boxes of level l+1."""
        d = _abi.CartesianDesc()
        d.dim = 2
        for i in range(2):
            d.cells[i], d.block_cells[i] = int(cells[i]), int(block_cells[i])
            d.origin[i], d.spacing[i] = float(origin[i]), float(spacing[i])
        boxes = [v for l, bl in enumerate(levels) for (lo, hi) in bl for v in (l + 1, lo[0], lo[1], hi[0], hi[1])]
        r = (C.c_uint32 * max(1, len(ratios)))(*ratios)
        b = (C.c_int64 * max(1, len(boxes)))(*boxes)
        h = C.c_void_p()
        _abi.check(ctx._lib.lgd_amr_create(ctx.handle, C.byref(d), r, len(ratios), b, len(boxes) // 5,
                                           C.byref(h)), ctx.handle)
        return cls(ctx, h)

    def close(self) -> None:
        if self.owned and getattr(self, "handle", None):
            self._lib.lgd_amr_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_blocks(self) -> int:
        return int(self._lib.lgd_amr_num_blocks(self.handle))

    @property
    def num_levels(self) -> int:
        return int(self._lib.lgd_amr_num_levels(self.handle))

    @property
    def field_len(self) -> int:
        return int(self._lib.lgd_amr_field_len(self.handle))

    def patch(self, b: int):
        out = (C.c_int64 * 5)()
        _abi.check(self._lib.lgd_amr_patch(self.handle, b, out))
        return int(out[0]), (int(out[1]), int(out[2])), (int(out[3]), int(out[4]))

    def offset(self, b: int) -> int:
        return int(self._lib.lgd_amr_block_offset(self.handle, b))

    def find_patch(self, level: int, cell) -> Optional[int]:
        r = int(self._lib.lgd_amr_find_patch(self.handle, level, cell[0], cell[1]))
        return None if r < 0 else r

    def alloc_field(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, self.field_len)

    def restrict_level(self, field: DeviceBuffer, fine_level: int) -> None:
        _abi.check(self._lib.lgd_amr_restrict_level(self.ctx.handle, self.handle, field.handle, fine_level),
                   self.ctx.handle)

    def fill_level(self, field: DeviceBuffer, level: int, old: Optional[DeviceBuffer] = None,
                   alpha: float = 0.0) -> None:
        _abi.check(self._lib.lgd_amr_fill_level(self.ctx.handle, self.handle, field.handle,
                                                old.handle if old is not None else None, alpha, level),
                   self.ctx.handle)


def amr_simulation(amr: AmrGrid, model: str = "heat", params: Optional[ModelCParams] = None,
                   kappa: float = 0.0, wiener: bool = False, seed: int = 0, subcycle: bool = False,
                   policy: Optional[RegridPolicy] = None) -> Simulation:
    """Simulation::adaptive on an AmrGrid (takes ownership of `amr`)."""
    h = C.c_void_p()
    p = params.c() if params is not None else None
    pol = policy.c() if policy is not None else None
    _abi.check(amr._lib.lgd_sim_create_amr(amr.ctx.handle, amr.handle, {"heat": 0, "modelc": 1}[model],
                                           C.byref(p) if p is not None else None, kappa, int(wiener), seed,
                                           int(subcycle), C.byref(pol) if pol is not None else None,
                                           C.byref(h)), amr.ctx.handle)
    amr.owned = False
    sim = Simulation(None, h, 2 if model == "modelc" else 1, ctx=amr.ctx)
    return sim


def _sim_amr(self) -> AmrGrid:
    """The simulation's current hierarchy (replaced by a regrid)."""
    return AmrGrid(self.ctx, C.c_void_p(self._lib.lgd_sim_amr(self.handle)), owned=False)


Simulation.amr = property(_sim_amr)


@dataclasses.dataclass(frozen=True)
class Realization:
    """One member of an ensemble: its own Simulation of ModelC on a CartesianGrid."""

    cells: tuple
    eps4: float
    lambda_: float
    d_thermal: float
    noise_amplitude: float = 0.0
    seed: int = 0


class Ensemble:
    """Independent Model C realizations of uneven sizes advanced together
    (one launch per step over every block of every realization)."""

    def __init__(self, ctx: Context, realizations, block_cells=(64, 64), spacing=(0.4, 0.4),
                 anisotropy_tol: float = 1e-8, fv_tol: float = 1e-8):
        self.ctx = ctx
        self._lib = ctx._lib
        self.realizations = list(realizations)
        arr = (_abi.RealizationC * len(self.realizations))()
        for i, r in enumerate(self.realizations):
            arr[i].cells[0], arr[i].cells[1] = int(r.cells[0]), int(r.cells[1])
            arr[i].eps4, arr[i].lambda_, arr[i].d_thermal = r.eps4, r.lambda_, r.d_thermal
            arr[i].noise_amplitude, arr[i].seed = r.noise_amplitude, r.seed
        bc = (C.c_int64 * 2)(*block_cells)
        sp = (C.c_double * 2)(*spacing)
        h = C.c_void_p()
        _abi.check(self._lib.lgd_ensemble_create(ctx.handle, arr, len(self.realizations), bc, sp,
                                                 anisotropy_tol, fv_tol, C.byref(h)), ctx.handle)
        self.handle = h
        self.block_cells = tuple(block_cells)
        ctx._adopt(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.lgd_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_blocks(self) -> int:
        return int(self._lib.lgd_ensemble_num_blocks(self.handle))

    def block_start(self, r: int) -> int:
        return int(self._lib.lgd_ensemble_block_start(self.handle, r))

    def field(self, index: int) -> DeviceBuffer:
        h = C.c_void_p()
        _abi.check(self._lib.lgd_ensemble_field(self.handle, index, C.byref(h)), self.ctx.handle)
        return DeviceBuffer._borrow(self.ctx, h)

    def initialize_grains(self, r: int, grains: np.ndarray, u0: float) -> None:
        g = np.ascontiguousarray(grains, dtype=np.float64).reshape(-1, 3)
        _abi.check(self._lib.lgd_ensemble_initialize_grains(self.handle, r, g.ctypes.data_as(_abi.DP),
                                                            g.shape[0], u0), self.ctx.handle)

    def step(self, dt: float, nsteps: int = 1, per_block: bool = False) -> None:
        _abi.check(self._lib.lgd_ensemble_step(self.handle, dt, nsteps, int(per_block)), self.ctx.handle)

    @property
    def time(self) -> float:
        return float(self._lib.lgd_ensemble_time(self.handle))

    def solid_fraction(self) -> np.ndarray:
        out = np.empty(len(self.realizations))
        _abi.check(self._lib.lgd_ensemble_solid_fraction(self.handle, out.ctypes.data_as(_abi.DP)),
                   self.ctx.handle)
        return out

    def realization_field(self, r: int, index: int) -> np.ndarray:
        """Download realization r's slabs of one field (block-major)."""
        nx, ny = self.block_cells
        slab = (nx + 2) * (ny + 2)
        b0, b1 = self.block_start(r), self.block_start(r + 1)
        return self.field(index).download(b0 * slab, (b1 - b0) * slab)


# ---------------------------------------------------------------- Model C adjoint
class Adjoint:
    """Reverse steps of the explicit Model C step (b = 0) on one grid: no
    reference counterpart (the reference has no derivative code)."""

    def __init__(self, grid: CartesianGrid):
        self.grid = grid
        self.ctx = grid.ctx
        self._lib = grid._lib
        h = C.c_void_p()
        _abi.check(self._lib.lgd_adjoint_create(self.ctx.handle, grid.handle, C.byref(h)), self.ctx.handle)
        self.handle = h
        self.ctx._adopt(self)

    def step(self, params: ModelCParams, phi: DeviceBuffer, u: DeviceBuffer, dt: float,
             pphi: DeviceBuffer, pu: DeviceBuffer) -> None:
        """(pphi, pu) <- transpose of d(step)/d(phi, u) applied to (pphi, pu); the
        step's dL/d(λ, D, ε₄) is accumulated."""
        pc = params.c()
        _abi.check(self._lib.lgd_modelc_adjoint_step(self.handle, C.byref(pc), phi.handle, u.handle,
                                                     dt, pphi.handle, pu.handle), self.ctx.handle)

    def reset(self) -> None:
        _abi.check(self._lib.lgd_adjoint_reset(self.handle), self.ctx.handle)

    def grad(self) -> np.ndarray:
        out = np.zeros(3)
        _abi.check(self._lib.lgd_adjoint_grad(self.handle, out.ctypes.data_as(_abi.DP)), self.ctx.handle)
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.lgd_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ModelCCalibration:
    """L(θ, φ₀, u₀) = ½‖φ_T − target‖² after `nsteps` Model C steps, with
    forward() and backward() in the autograd-Function role: forward runs the
    sweep and keeps the checkpoints, backward returns dL/d(λ, D, ε₄) and
    dL/d(φ₀, u₀) by binomial checkpointing in `snapshots` state slots."""

    def __init__(self, grid: CartesianGrid, nsteps: int, dt: float, snapshots: int = 45):
        self.grid = grid
        self.ctx = grid.ctx
        self._lib = grid._lib
        self.nsteps = int(nsteps)
        self.dt = float(dt)
        self.snapshots = int(snapshots)
        h = C.c_void_p()
        _abi.check(self._lib.lgd_calibration_create(self.ctx.handle, grid.handle, self.nsteps, self.dt,
                                                    self.snapshots, C.byref(h)), self.ctx.handle)
        self.handle = h
        self.ctx._adopt(self)

    def forward(self, params: ModelCParams, phi0: DeviceBuffer, u0: DeviceBuffer,
                target: DeviceBuffer) -> float:
        loss = C.c_double(0.0)
        pc = params.c()
        _abi.check(self._lib.lgd_calibration_forward(self.handle, C.byref(pc), phi0.handle, u0.handle,
                                                     target.handle, C.byref(loss)), self.ctx.handle)
        return loss.value

    def backward(self, gphi0: Optional[DeviceBuffer] = None, gu0: Optional[DeviceBuffer] = None):
        grad = np.zeros(3)
        _abi.check(self._lib.lgd_calibration_backward(
            self.handle, grad.ctypes.data_as(_abi.DP), gphi0.handle if gphi0 is not None else None,
            gu0.handle if gu0 is not None else None), self.ctx.handle)
        return grad

    def counts(self):
        f, r = C.c_int64(0), C.c_int64(0)
        _abi.check(self._lib.lgd_calibration_counts(self.handle, C.byref(f), C.byref(r)), self.ctx.handle)
        return f.value, r.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.lgd_calibration_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def checkpoint_forward_steps(nsteps: int, snapshots: int) -> int:
    """Forward steps of one forward + backward pass of ModelCCalibration."""
    return int(_abi.lib().lgd_checkpoint_forward_steps(int(nsteps), int(snapshots)))

