"""ctypes binding of libwithin_amd.so (the C ABI declared in include/within_amd.h).

Loading fails loudly: there is no CPU fallback.  The library must have been built
(``within_amd.build.build()``) and, to solve, a GPU must be visible.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WITHIN_AMD_LIB", os.path.join(_HERE, "libwithin_amd.so"))

WX_OK = 0
WX_ERR_INVALID_INPUT = 1
WX_ERR_BUILD = 2
WX_ERR_RUNTIME = 3
WX_ERR_DEVICE = 4
WX_ERR_NOT_POSITIVE_DEF = 5
WX_ERR_INDEX = 6

STOP_REASONS = {
    0: "ZeroRhs",
    1: "InitialNormalEquationResidualZero",
    2: "ResidualTolerance",
    3: "NormalEquationTolerance",
    4: "WarmStartExact",
    5: "MaxIterations",
}


class PrecondConfig(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("schur", C.c_int32),
        ("approx_chol_seed", C.c_uint64),
        ("split_merge", C.c_int32),
        ("schur_seed", C.c_uint64),
        ("schur_split", C.c_int32),
        ("dense_threshold", C.c_int64),
        ("scaling_tolerance", C.c_double),
        ("scaling_max_sweeps", C.c_int64),
        ("scaling_error_on_failure", C.c_int32),
        ("reduction", C.c_int32),
        ("dense_policy", C.c_int32),
    ]


class LsmrOpts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("maxiter", C.c_int64),
        ("local_size", C.c_int64),
        ("precision", C.c_int32),
        ("check_every", C.c_int32),
        ("timing", C.c_int32),
        ("true_residual", C.c_int32),
    ]


class RhsStats(C.Structure):
    _fields_ = [
        ("converged", C.c_int32),
        ("stop_reason", C.c_int32),
        ("iterations", C.c_int64),
        ("residual", C.c_double),
        ("residual_norm", C.c_double),
        ("true_residual", C.c_double),
        ("time_solve", C.c_double),
    ]


EXPORTS = {
    "wx_last_error": ([], C.c_char_p),
    "wx_last_error_variant": ([], C.c_char_p),
    "wx_last_timing": ([C.c_void_p, C.c_uint64], C.c_uint64),
    "wx_device_info": ([C.c_char_p, C.c_uint64], C.c_int32),
    "wx_precond_config_default": ([C.POINTER(PrecondConfig)], None),
    "wx_lsmr_options_default": ([C.POINTER(LsmrOpts)], None),
    "wx_design_from_codes": ([C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)],
                             C.c_int),
    "wx_design_from_effects": ([C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_void_p)], C.c_int),
    "wx_design_release": ([C.c_void_p], None),
    "wx_solver_new": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "wx_solver_reweight": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "wx_solver_release": ([C.c_void_p], None),
    "wx_gram_device": ([C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p], C.c_int),
    "wx_gram": ([C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p], C.c_int),
    "wx_gram_device_noalloc": ([C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p], C.c_int),
    "wx_solver_n_obs": ([C.c_void_p], C.c_uint64),
    "wx_solver_n_dofs": ([C.c_void_p], C.c_uint64),
    "wx_solver_time_setup": ([C.c_void_p], C.c_double),
    "wx_solver_solve_batch": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(LsmrOpts), C.c_void_p,
                               C.c_void_p, C.POINTER(RhsStats)], C.c_int),
    "wx_solver_solve_batch_device": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(LsmrOpts),
                                      C.c_void_p, C.c_void_p, C.POINTER(RhsStats), C.c_void_p], C.c_int),
    "wx_solver_apply_operator": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64, C.c_void_p], C.c_int),
    "wx_solve_many_device": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(LsmrOpts),
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "wx_solve_many": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(LsmrOpts), C.c_void_p,
                       C.c_void_p, C.c_void_p], C.c_int),
    "wx_last_many_shape": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], None),
    "wx_set_many_shape": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32], None),
    "wx_set_many_tile": ([C.c_int32], None),
    "wx_last_many_tile": ([], C.c_int32),
    "wx_solver_projection_grad_device": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint64,
                                          C.POINTER(LsmrOpts), C.c_void_p, C.c_void_p, C.POINTER(RhsStats),
                                          C.c_void_p], C.c_int),
    "wx_solver_preconditioner": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "wx_solver_n_terms": ([C.c_void_p], C.c_uint64),
    "wx_solver_layout": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], None),
    "wx_solver_unidentified": ([C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)], C.c_int),
    "wx_solver_n_warnings": ([C.c_void_p], C.c_uint64),
    "wx_solver_warning": ([C.c_void_p, C.c_uint64], C.c_char_p),
    "wx_precond_build": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "wx_precond_release": ([C.c_void_p], None),
    "wx_precond_n": ([C.c_void_p], C.c_uint64),
    "wx_precond_variant": ([C.c_void_p], C.c_char_p),
    "wx_precond_apply": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p], C.c_int),
    "wx_precond_serialize": ([C.c_void_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint64)], C.c_int),
    "wx_precond_deserialize": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "wx_buffer_free": ([C.c_void_p], None),
    # include/within_amd_testing.h
    "wx_testing_precond_apply": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32], C.c_int),
    "wx_testing_set_fused_steps": ([C.c_int32], C.c_int32),
    "wx_testing_reparam": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                            C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p], C.c_int),
    "wx_testing_design_index": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p], C.c_int),
    "wx_testing_block_precond": ([C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
                                 C.c_int),
    "wx_testing_reduced_laplacian": ([C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.POINTER(C.c_uint32)], C.c_int),
}

_lib = None


def lib():
    """The loaded library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"within_amd native library not found at {LIB_PATH}; build it with "
                "`python -m within_amd.build` (hipcc, gfx950). There is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status: int) -> None:
    if status == WX_OK:
        return
    msg = lib().wx_last_error().decode("utf-8", "replace")
    variant = lib().wx_last_error_variant().decode("ascii", "replace")
    if status in (WX_ERR_INVALID_INPUT, WX_ERR_BUILD, WX_ERR_NOT_POSITIVE_DEF):
        err = ValueError(msg)
    elif status == WX_ERR_INDEX:
        err = IndexError(msg)
    else:
        err = RuntimeError(msg)
    # the reference error variant the status stands for ("BuildError::InvalidWeight", ...)
    err.variant = variant
    raise err


def device_info() -> tuple[int, str]:
    buf = C.create_string_buffer(256)
    n = lib().wx_device_info(buf, 256)
    return int(n), buf.value.decode()
