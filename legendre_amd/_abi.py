"""ctypes binding of the C ABI declared in include/legendre_gpu.h.

The shared library is built in-tree (legendre_amd/liblegendre_gpu.so) by
legendre_amd/build.py.  Loading fails loudly when it is missing: there is no
CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "liblegendre_gpu.so"

LGD_OK = 0
STATUS_NAMES = {
    0: "LGD_OK",
    1: "LGD_ERR_INVALID_ARGUMENT",
    2: "LGD_ERR_GHOST_TOO_WIDE",
    3: "LGD_ERR_SIZE_MISMATCH",
    4: "LGD_ERR_UNDRIVEN_DRIVER",
    5: "LGD_ERR_UNSUPPORTED",
    6: "LGD_ERR_OUT_OF_MEMORY",
    7: "LGD_ERR_HIP",
    8: "LGD_ERR_IO",
}

LGD_BC_MIRROR, LGD_BC_DIRICHLET, LGD_BC_FLUX, LGD_BC_EXTRAPOLATE = 0, 1, 2, 3
LGD_MODE_PER_INTENT, LGD_MODE_FUSED = 0, 1
LGD_INTEGRATOR_EULER, LGD_INTEGRATOR_RK4 = 0, 1
LGD_DRIVER_TIME, LGD_DRIVER_WIENER, LGD_DRIVER_JUMP = 0, 1, 2


class CartesianDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("cells", C.c_int64 * 3),
        ("block_cells", C.c_int64 * 3),
        ("origin", C.c_double * 3),
        ("spacing", C.c_double * 3),
    ]


class FaceBcC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("value", C.c_double)]


class ModelCParamsC(C.Structure):
    _fields_ = [
        ("eps4", C.c_double),
        ("lambda_", C.c_double),
        ("d_thermal", C.c_double),
        ("noise_amplitude", C.c_double),
        ("anisotropy_tol", C.c_double),
        ("fv_tol", C.c_double),
    ]


class RegridPolicyC(C.Structure):
    _fields_ = [
        ("every", C.c_int64),
        ("buffer", C.c_int64),
        ("efficiency", C.c_double),
        ("min_width", C.c_int64),
        ("threshold", C.c_double),
        ("tag_field", C.c_int32),
    ]


class RealizationC(C.Structure):
    _fields_ = [
        ("cells", C.c_int64 * 2),
        ("eps4", C.c_double),
        ("lambda_", C.c_double),
        ("d_thermal", C.c_double),
        ("noise_amplitude", C.c_double),
        ("seed", C.c_uint64),
    ]


class TuningResultC(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32),
        ("strips", C.c_int32),
        ("prefetch", C.c_int32),
        ("candidates", C.c_int32),
        ("ms", C.c_double),
        ("default_ms", C.c_double),
    ]


class AmrTuningResultC(C.Structure):
    _fields_ = [
        ("tile_x", C.c_int32),
        ("tile_y", C.c_int32),
        ("wg", C.c_int32),
        ("level0_strips", C.c_int32),
        ("level0_prefetch", C.c_int32),
        ("candidates", C.c_int32),
        ("stage_rows", C.c_int32),
        ("level0_kernel", C.c_int32),
        ("ms", C.c_double),
        ("default_ms", C.c_double),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)


class ObservedFieldC(C.Structure):
    _fields_ = [("name", C.c_char_p), ("field", C.c_int32)]


class DriverC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_uint32), ("amp", PP), ("rate", P)]

I32, I64, U32, U64, SZ, DBL = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_size_t, C.c_double
DP = C.POINTER(C.c_double)

# lgd_tag_level_fn: int (*)(void *user, const lgd_amr *grid, int32_t level, lgd_tag_sink *sink)
TAG_LEVEL_FN = C.CFUNCTYPE(C.c_int, P, P, I32, P)

# name -> (restype, argtypes); the exact list include/legendre_gpu.h declares.
SIGNATURES = {
    "lgd_abi_version": (U32, []),
    "lgd_ctx_create": (C.c_int, [I32, P, PP]),
    "lgd_ctx_destroy": (C.c_int, [P]),
    "lgd_ctx_sync": (C.c_int, [P]),
    "lgd_ctx_stream": (P, [P]),
    "lgd_ctx_last_error": (C.c_char_p, [P]),
    "lgd_ctx_set_tuning": (C.c_int, [P, C.c_char_p, I64]),
    "lgd_autotune_modelc_step": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P, P, DBL, I32, I32,
                                           C.POINTER(TuningResultC)]),
    "lgd_stream_copy": (C.c_int, [P, P, P, I32]),
    "lgd_buf_alloc": (C.c_int, [P, SZ, PP]),
    "lgd_buf_free": (C.c_int, [P]),
    "lgd_buf_len": (SZ, [P]),
    "lgd_buf_device_ptr": (P, [P]),
    "lgd_buf_upload": (C.c_int, [P, P, SZ, DP, SZ]),
    "lgd_buf_download": (C.c_int, [P, P, SZ, DP, SZ]),
    "lgd_buf_fill_zero": (C.c_int, [P, P]),
    "lgd_buf_copy": (C.c_int, [P, P, P]),
    "lgd_grid_create_cartesian": (C.c_int, [P, C.POINTER(CartesianDesc), PP]),
    "lgd_grid_destroy": (C.c_int, [P]),
    "lgd_grid_num_blocks": (I64, [P]),
    "lgd_grid_block_len": (I64, [P, U32]),
    "lgd_grid_face_neighbor": (I64, [P, I64, I32, I32]),
    "lgd_grid_cell_key": (I64, [P, U32, I64]),
    "lgd_fill_ghosts": (C.c_int, [P, P, P, U32, C.POINTER(FaceBcC)]),
    "lgd_central_laplacian": (C.c_int, [P, P, P, P, U32]),
    "lgd_karma_rappel_flux": (C.c_int, [P, P, P, P, DBL, DBL]),
    "lgd_modelc_rhs": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P, P, P, P]),
    "lgd_modelc_rhs_oriented": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P, P, P, P, P, P]),
    "lgd_modelc_amplitude": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P]),
    "lgd_heat_rhs": (C.c_int, [P, P, DBL, P, P]),
    "lgd_axpy": (C.c_int, [P, P, P, DBL]),
    "lgd_apply_wiener": (C.c_int, [P, P, P, P, U32, DBL, U64, U64, U64]),
    "lgd_apply_drivers": (C.c_int, [P, P, SZ, PP, C.POINTER(U32), SZ, C.POINTER(DriverC), DBL, U64, U64]),
    "lgd_modelc_step": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P, P, P, P, DBL, U64, U64, I32]),
    "lgd_heat_step": (C.c_int, [P, P, DBL, P, P, DBL]),
    "lgd_heat_step_bc": (C.c_int, [P, P, DBL, C.POINTER(FaceBcC), P, P, DBL]),
    "lgd_amr_create": (C.c_int, [P, C.POINTER(CartesianDesc), C.POINTER(U32), SZ, C.POINTER(I64), SZ, PP]),
    "lgd_amr_destroy": (C.c_int, [P]),
    "lgd_amr_num_blocks": (I64, [P]),
    "lgd_amr_num_levels": (I32, [P]),
    "lgd_amr_patch": (C.c_int, [P, I64, C.POINTER(I64)]),
    "lgd_amr_field_len": (I64, [P]),
    "lgd_amr_block_offset": (I64, [P, I64]),
    "lgd_amr_find_patch": (I64, [P, I32, I64, I64]),
    "lgd_amr_restrict_level": (C.c_int, [P, P, P, I32]),
    "lgd_amr_fill_level": (C.c_int, [P, P, P, P, DBL, I32]),
    "lgd_sim_create_amr": (C.c_int, [P, P, I32, C.POINTER(ModelCParamsC), DBL, I32, U64, I32,
                                     C.POINTER(RegridPolicyC), PP]),
    "lgd_sim_amr": (P, [P]),
    "lgd_ensemble_create": (C.c_int, [P, C.POINTER(RealizationC), SZ, C.POINTER(I64), DP, DBL, DBL, PP]),
    "lgd_ensemble_destroy": (C.c_int, [P]),
    "lgd_ensemble_num_blocks": (I64, [P]),
    "lgd_ensemble_block_start": (I64, [P, I64]),
    "lgd_ensemble_field": (C.c_int, [P, I32, PP]),
    "lgd_ensemble_initialize_grains": (C.c_int, [P, I64, DP, SZ, DBL]),
    "lgd_ensemble_step": (C.c_int, [P, DBL, I64, I32]),
    "lgd_ensemble_time": (DBL, [P]),
    "lgd_ensemble_solid_fraction": (C.c_int, [P, DP]),
    "lgd_sim_create_modelc": (C.c_int, [P, P, C.POINTER(ModelCParamsC), I32, U64, I32, PP]),
    "lgd_sim_create_heat": (C.c_int, [P, P, DBL, C.POINTER(FaceBcC), I32, PP]),
    "lgd_sim_destroy": (C.c_int, [P]),
    "lgd_sim_field": (C.c_int, [P, I32, PP]),
    "lgd_sim_step": (C.c_int, [P, DBL, I64]),
    "lgd_sim_set_graph_steps": (C.c_int, [P, I64]),
    "lgd_sim_regrid_stats": (C.c_int, [P, C.POINTER(C.c_int64), DP]),
    "lgd_sim_time": (DBL, [P]),
    "lgd_sim_step_index": (I64, [P]),
    "lgd_sim_set_tagger": (C.c_int, [P, P, P]),
    "lgd_sim_autotune_amr": (C.c_int, [P, DBL, I32, C.POINTER(AmrTuningResultC)]),
    "lgd_tag_sink_push": (C.c_int, [P, C.POINTER(C.c_int64), C.c_size_t]),
    "lgd_sim_initialize_grains": (C.c_int, [P, DP, SZ, DBL]),
    "lgd_sim_set_integrator": (C.c_int, [P, C.c_int32]),
    "lgd_sim_set_orientation": (C.c_int, [P, DP, SZ]),
    "lgd_sim_initialize_grains_oriented": (C.c_int, [P, DP, SZ, DBL]),
    "lgd_observer_create": (C.c_int, [P, C.c_char_p, U64, SZ, C.POINTER(ObservedFieldC), SZ,
                                      C.POINTER(ObservedFieldC), SZ, PP]),
    "lgd_observer_destroy": (C.c_int, [P]),
    "lgd_observer_observe": (C.c_int, [P, P, U64, DBL, U64, PP, C.POINTER(U32), PP, C.POINTER(U32)]),
    "lgd_observer_observe_amr": (C.c_int, [P, P, U64, DBL, U64, PP]),
    "lgd_observer_flush": (C.c_int, [P]),
    "lgd_observer_pause": (C.c_int, [P, C.c_int32]),
    "lgd_observer_set_blocking": (C.c_int, [P, C.c_int32]),
    "lgd_observer_stats": (C.c_int, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64), DP]),
    "lgd_observer_last_error": (C.c_int, [P, C.c_char_p, SZ]),
    "lgd_sim_attach_observer": (C.c_int, [P, P]),
    "lgd_sim_detach_observer": (C.c_int, [P, P]),
    "lgd_parquet_write_file": (C.c_int, [C.c_char_p, SZ, C.POINTER(C.c_char_p), C.POINTER(C.c_int32),
                                         PP, C.POINTER(U64), I64, C.POINTER(I64), SZ, C.c_char_p, SZ]),
    "lgd_adjoint_create": (C.c_int, [P, P, PP]),
    "lgd_adjoint_destroy": (C.c_int, [P]),
    "lgd_modelc_adjoint_step": (C.c_int, [P, C.POINTER(ModelCParamsC), P, P, DBL, P, P]),
    "lgd_modelc_adjoint_step_noalloc": (C.c_int, [P, P, C.POINTER(ModelCParamsC), P, P, DBL, P, P, P]),
    "lgd_adjoint_reset": (C.c_int, [P]),
    "lgd_adjoint_grad": (C.c_int, [P, DP]),
    "lgd_calibration_create": (C.c_int, [P, P, I64, DBL, I32, PP]),
    "lgd_calibration_destroy": (C.c_int, [P]),
    "lgd_calibration_forward": (C.c_int, [P, C.POINTER(ModelCParamsC), P, P, P, DP]),
    "lgd_calibration_backward": (C.c_int, [P, DP, P, P]),
    "lgd_calibration_counts": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64)]),
    "lgd_checkpoint_forward_steps": (I64, [I64, I32]),
}

_lib = None


class LegendreError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


def lib() -> C.CDLL:
    """Load the HIP backend (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP backend first "
                "(python -m legendre_amd.build or __graft_entry__.build())"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, ctx=None) -> None:
    if status != LGD_OK:
        msg = ""
        if ctx is not None:
            raw = lib().lgd_ctx_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise LegendreError(status, msg)
