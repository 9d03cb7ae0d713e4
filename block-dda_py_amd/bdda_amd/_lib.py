"""ctypes binding of libbdda.so (the C ABI declared in include/bdda.h).

The library is built in-tree by ``block-dda_py_amd/build_lib.py`` into
``block-dda_py_amd/lib/libbdda.so``.  There is no CPU fallback: if the
library is missing or no GPU is visible, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BDDA_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libbdda.so"))

OK, E_ARG, E_HIP, E_SINGULAR, E_OOM, E_STATE = range(6)
SOLVER = {"cocg_rq": 0, "cocg": 0, "bicgstab": 1}
PRECISION = {"fp64": 0, "mixed": 1}

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (restype int unless noted)
_SIGNATURES = {
    "bdda_last_error": ([], ctypes.c_char_p),
    "bdda_version": ([], _i),
    "bdda_set_fft_path": ([_i], _i),
    "bdda_set_operator_groups": ([_i], _i),
    "bdda_set_la_path": ([_i], _i),
    "bdda_plan_column_tile": ([_vp, _i], _i),
    "bdda_plan_set_column_tile": ([_vp, _i], _i),
    "bdda_plan_autotune": ([_vp, _i, _i, _pi, _vp], _i),
    "bdda_plan_set_x_kernel": ([_vp, _i], _i),
    "bdda_plan_x_kernel": ([_vp], _i),
    "bdda_plan_create": ([_i, _i, _i, ctypes.POINTER(_vp)], _i),
    "bdda_plan_destroy": ([_vp], _i),
    "bdda_plan_fft_dims": ([_vp, _pi], _i),
    "bdda_green_init": ([_vp, _d, _d], _i),
    "bdda_green_share": ([_vp, _vp], _i),
    "bdda_green_download": ([_vp, _vp], _i),
    "bdda_green_tensor": ([_vp, _d, _d, _vp], _i),
    "bdda_mvp_offdiag": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "bdda_set_target": ([_vp, _vp, _i64, _vp, _vp, _vp], _i),
    "bdda_block_mvp": ([_vp, _vp, _vp, _i, _vp], _i),
    "bdda_solve": ([_vp, _i, _i, _vp, _vp, _vp, _i, _d, _i, _i, _pi, _pd, _pd, _pd, _pi, _vp], _i),
    "bdda_project_guess": ([_vp, _vp, _vp, _i, _pd, _pd, _vp], _i),
    "bdda_true_residual": ([_vp, _vp, _vp, _i, _pd, _vp], _i),
    "bdda_solve_batch": ([_i, _vp, _i, _i, _vp, _vp, _vp, _vp, _d, _i, _i, _i, _pi, _pd, _pd, _pd, _pi], _i),
    "bdda_profile_enable": ([_vp, _i], _i),
    "bdda_profile_read": ([_vp, _pd, ctypes.POINTER(ctypes.c_int64)], _i),
    "bdda_rhs": ([_vp, _vp, _vp, _d, _i, _vp, _vp], _i),
    "bdda_observables": ([_vp, _vp, _i, _vp, _vp, _d, _vp, _vp], _i),
    "bdda_observables_noalloc": ([_vp, _vp, _i, _vp, _vp, _d, _vp, _vp], _i),
    "bdda_nearest_distance": ([_vp, _i64, _vp, _i64, _vp], _i),
    "bdda_inside_mask": ([_vp, _i, _i, _i, _d, _vp], _i),
    "bdda_gemm_bf16": ([_i, _i, _i, _i, ctypes.c_float, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                        ctypes.c_float, _vp, _i64, _i64, _i64, _vp], _i),
    "bdda_gemm_bf16_ex": ([_i, _i, _i, _i, ctypes.c_float, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                           ctypes.c_float, _vp, _i64, _i64, _i64, _vp, _i, _vp], _i),
    "bdda_fno_inputs": ([_vp, _vp, _vp, _vp, _i, _i64, _vp], _i),
    "bdda_cube_exchange": ([_vp, _vp, _vp, _vp, _i64, _i, _i, _i, _vp], _i),
    "bdda_masked_l2": ([_vp, _vp, _vp, _vp, _i, _i64, _vp, _vp, _i, _vp], _i),
    "bdda_gelu": ([_vp, _vp, _i64, _vp], _i),
    "bdda_gelu_bwd": ([_vp, _vp, _vp, _i64, _vp], _i),
    "bdda_bias_add": ([_vp, _vp, _i64, _i, _i64, _vp], _i),
    "bdda_fno_ycombine": ([_vp, _vp, _vp, _i64, _i, _vp], _i),
    "bdda_bias_grad": ([_vp, _vp, _i64, _i, _i64, _vp], _i),
    "bdda_adam": ([_vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _i,
                   _vp], _i),
}

_lib = None


class BddaError(RuntimeError):
    """A libbdda call failed (HIP runtime error, bad call order, ...)."""


def load():
    """Load libbdda.so once and declare every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BddaError(f"libbdda.so not found at {LIB_PATH}; build it with "
                        "`python block-dda_py_amd/build_lib.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc, what=""):
    """Map a libbdda status code to the reference's Python exception types."""
    if rc == OK:
        return
    msg = load().bdda_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == E_ARG:
        raise ValueError(text)
    if rc == E_SINGULAR:
        raise np.linalg.LinAlgError(text)
    if rc == E_OOM:
        raise MemoryError(text)
    raise BddaError(text)


def ptr(a):
    """Raw pointer of a contiguous numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return a.data_ptr()
