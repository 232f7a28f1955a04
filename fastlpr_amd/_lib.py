"""ctypes binding of the C-ABI library ``libfastlpr_amd.so`` (include/fastlpr_amd.h).

The library is built in-tree by ``fastlpr_amd._build.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_NAME = "libfastlpr_amd.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MAX_DIM = 3

EXPORTS = (
    "flpr_last_hip_error",
    "flpr_xstats",
    "flpr_xstats_noalloc",
    "flpr_points_prepare",
    "flpr_points_bin",
    "flpr_points_pack",
    "flpr_spread",
    "flpr_spread_wide",
    "flpr_resample_work",
    "flpr_resample",
    "flpr_grid_center",
    "flpr_grid_periodize",
    "flpr_taps",
    "flpr_conv",
    "flpr_conv_tiles",
    "flpr_conv_strip_len",
    "flpr_conv_strips",
    "flpr_max_moment",
    "flpr_key_to_double",
    "flpr_solve",
    "flpr_solve_refined",
    "flpr_adjoint",
    "flpr_quadform",
    "flpr_interp_eval",
    "flpr_fit_output",
    "flpr_select",
    "flpr_lcv",
    "flpr_spline_eval",
    "flpr_fill_zero",
    "flpr_batch_xstats",
    "flpr_batch_points_prepare",
    "flpr_batch_points_pack",
    "flpr_batch_spread",
    "flpr_batch_grid_center",
    "flpr_batch_taps",
    "flpr_batch_max_moment",
    "flpr_batch_solve",
    "flpr_batch_adjoint",
    "flpr_batch_quadform",
    "flpr_batch_select",
    "flpr_batch_fit_output",
    "flpr_batch_spline_coef",
    "flpr_batch_point_sums",
    "flpr_mlp_param_count",
    "flpr_mlp_gemm",
    "flpr_mlp_forward",
    "flpr_mlp_backward",
    "flpr_mlp_loss",
    "flpr_mlp_adam",
    "flpr_bracket_features",
)


class FlprGeom(C.Structure):
    """Mirror of ``FlprGeom`` in include/fastlpr_amd.h (field order matters)."""

    _fields_ = [
        ("d", C.c_int32),
        ("msp", C.c_int32),
        ("n_grid", C.c_int32 * MAX_DIM),
        ("L", C.c_int32 * MAX_DIM),
        ("E", C.c_int32 * MAX_DIM),
        ("G", C.c_int32 * MAX_DIM),
        ("tile", C.c_int32 * MAX_DIM),
        ("ntile", C.c_int32 * MAX_DIM),
        ("item_points", C.c_int32),
        ("spread_f32", C.c_int32),
        ("hx", C.c_double * MAX_DIM),
        ("inv4tau", C.c_double * MAX_DIM),
        ("mean", C.c_double * MAX_DIM),
        ("std_", C.c_double * MAX_DIM),
        ("centre", C.c_double * MAX_DIM),
        ("xmin_s", C.c_double * MAX_DIM),
        ("scale_s", C.c_double * MAX_DIM),
        ("xmax_s", C.c_double * MAX_DIM),
        ("ks", C.c_double * MAX_DIM),
        ("idx0", C.c_double * MAX_DIM),
        ("shift", C.c_double * MAX_DIM),
        ("graw0", C.c_double * MAX_DIM),
        ("graw1", C.c_double * MAX_DIM),
    ]


class FlprBatchItem(C.Structure):
    """Mirror of ``FlprBatchItem`` in include/fastlpr_amd.h."""

    _fields_ = [("g", FlprGeom), ("n", C.c_int64), ("pt", C.c_int64)]


class FlprError(RuntimeError):
    pass


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_GP = C.POINTER(FlprGeom)

_SIGS = {
    "flpr_last_hip_error": [C.c_char_p, _I32],
    "flpr_xstats": [_P, _I64, _I32, _P, _I32, _P, _P, _P],
    "flpr_xstats_noalloc": [_P, _I64, _I32, _P, _I32, _P, _P],
    "flpr_points_prepare": [_P, _I64, _GP, _P, _P, _P, _P],
    "flpr_points_bin": [_P, _I64, _I32, _P, _P, _P, _P, _P, _P],
    "flpr_points_pack": [_GP, _P, _P, _I64, _P, _I32, _P, _I32, _P, _P, _P, _P, _P],
    "flpr_spread": [_GP, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P, _P],
    "flpr_spread_wide": [_GP, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P],
    "flpr_resample_work": [_GP, _GP, _I32, _I32, _P],
    "flpr_grid_periodize": [_GP, _P, _I32, _P, _P],
    "flpr_conv_tiles": [_I32, _I32, _I32, _I32, _I32, _I32, _P, _P],
    "flpr_conv_strips": [_P, _P, _I32, _I32, _I32, _P, _P],
    "flpr_resample": [_GP, _GP, _P, _I32, _I32, _P, _P, _P, _P],
    "flpr_grid_center": [_P, _I32, _I64, _P, _P, _P, _P],
    "flpr_taps": [_GP, _P, _I32, _I32, _I32, _D, _P, _P, _P],
    "flpr_conv": [_I32, _I32, _P, _P, _P, _I32, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32,
                  _I32, _P, _I32, _P],
    "flpr_max_moment": [_I32, _P, _I64, _I32, C.POINTER(_I64), _I32, _I64, _I32, _P, _P],
    "flpr_solve": [_I32, _P, _I64, _I32, _I64, _I32, _I32, _I32, C.POINTER(_I64), C.POINTER(_I64), _P,
                   _D, _D, _P, _I32, _I32, _P, _P, _P, _P, _I64, _P],
    "flpr_solve_refined": [_P, _I64, _I32, _I64, _I32, _I32, C.POINTER(_I64), C.POINTER(_I64), _P,
                           _D, _D, _P, _I32, _I32, _P, _P, _P, _P, _I64, _P],
    "flpr_adjoint": [_GP, _P, _P, _P, _P, _I64, _I32, _P, _P, _P, _P],
    "flpr_quadform": [_GP, _P, _I32, _I32, _P, _P, _P, _I64, _P],
    "flpr_interp_eval": [_GP, _P, _I64, _P, _I32, _P, _P, _P, _P],
    "flpr_fit_output": [_GP, _P, _I64, _P, _I32, _P, _P, _I32, _P, _P, _P, _P, _P, _P],
    "flpr_spline_eval": [_GP, _P, _I64, _P, _I32, _P, _P],
    "flpr_lcv": [_GP, _P, _P, _I64, _P, _I32, _P, _P, _P, _P],
    "flpr_select": [_P, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _I32, _I32, _D, _P, _P, _P],
    "flpr_fill_zero": [_P, _I64, _P],
    "flpr_batch_xstats": [_P, _I32, _P, _I32, _P, _I32, _P, _P, _P],
    "flpr_batch_points_prepare": [_GP, _P, _I32, _I64, _P, _P, _P, _P, _P],
    "flpr_batch_points_pack": [_GP, _P, _I32, _I64, _P, _P, _P, _I32, _P, _I32, _P, _P, _P, _P,
                               _P],
    "flpr_batch_spread": [_GP, _P, _I32, _I64, _P, _P, _P, _I32, _I32, _P, _P, _I64, _P, _P],
    "flpr_batch_grid_center": [_I32, _P, _I64, _I32, _I64, _P, _P, _P, _P],
    "flpr_batch_taps": [_GP, _P, _I32, _P, _I32, _I32, _I32, _P, _P, _P],
    "flpr_batch_max_moment": [_I32, _I32, _P, _I64, _I64, _I32, C.POINTER(_I64), _I32, _I64, _I32,
                              _P, _P],
    "flpr_batch_solve": [_I32, _I32, _P, _I64, _I64, _I32, _I64, _I32, _I32, C.POINTER(_I64),
                         C.POINTER(_I64), _P, _D, _D, _P, _I32, _I32, _P, _I64, _P, _P, _I64, _P,
                         _I64, _P],
    "flpr_batch_adjoint": [_GP, _P, _I32, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P],
    "flpr_batch_quadform": [_GP, _I32, _P, _I64, _I32, _I32, _P, _P, _I64, _P, _I64, _P],
    "flpr_batch_select": [_P, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _I32, _I32, _D,
                          _P, _I64, _P, _P],
    "flpr_batch_fit_output": [_GP, _P, _I32, _I64, _P, _P, _I64, _I32, _P, _P, _I32, _P, _P, _P,
                              _P, _P, _I64, _P],
    "flpr_batch_spline_coef": [_GP, _I32, _P, _I64, _P, _I32, _P, _P, _P],
    "flpr_batch_point_sums": [_GP, _P, _I32, _I64, _P, _P, _I32, _P, _I32, _P, _P, _I32, _I32,
                              _I32, _P, _P, _I64, _P],
    "flpr_mlp_param_count": [],
    "flpr_mlp_gemm": [_I32, _I32, _I32, _P, _I32, _I32, _P, _I32, _I32, _P, _I32, _P, _P, _I32, _P],
    "flpr_mlp_forward": [_P, _I32, _P, _P, _P, _P, _P, _P, _P],
    "flpr_mlp_backward": [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "flpr_mlp_loss": [_I32, _P, _P, C.c_float, C.c_float, _P, _P, _P],
    "flpr_mlp_adam": [_I32, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float, _I32,
                      C.c_float, _P],
    "flpr_bracket_features": [_I32, _P, _I64, _I32, _I32, _P, _P, _P, _P, _P, _P, _P],
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the library (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    # Bind to the HIP runtime torch already carries: torch loads its own
    # libamdhip64 (same soname); loading this library first would pull a second
    # runtime from /opt/rocm into the process.
    import torch  # noqa: F401
    if not os.path.exists(path):
        raise FlprError(
            f"{LIB_NAME} not found at {path}; build it with fastlpr_amd._build.build() "
            "(hipcc --offload-arch=gfx950).  There is no CPU fallback.")
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.flpr_key_to_double.argtypes = [C.c_uint64]
    lib.flpr_key_to_double.restype = C.c_double
    lib.flpr_conv_strip_len.argtypes = [C.c_int32]
    lib.flpr_conv_strip_len.restype = C.c_int64
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point; raise FlprError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        kinds = {-1: "invalid argument", -2: "HIP launch/runtime error",
                 -3: "table or scratch too small"}
        detail = ""
        if rc == -2:
            buf = C.create_string_buffer(256)
            code = lib.flpr_last_hip_error(buf, 256)
            detail = f": hipError {code} {buf.value.decode(errors='replace')}"
        raise FlprError(f"{name} failed with status {rc} ({kinds.get(rc, 'unknown')}){detail}")


def i64_array(values):
    arr = (C.c_int64 * max(1, len(values)))()
    for i, v in enumerate(values):
        arr[i] = int(v)
    return arr
