"""ctypes binding of libmrmath_gpu.so (the C ABI declared in include/mrmath_gpu.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmrmath_gpu.so"

(MRM_OK, MRM_SINGULAR, MRM_BAD_DIMS, MRM_BAD_ARG, MRM_HIP_ERROR, MRM_NO_MEMORY, MRM_NOT_INITIALISED,
 MRM_NOT_CONVERGED, MRM_TIMED_OUT) = range(9)
DO_NONE, DO_ADD, DO_SUB = 0, 1, 2
NN, TN, NT = 0, 1, 2
SOLVE_FP64, SOLVE_FP32_IR, SOLVE_BF16_IR, SOLVE_BF16X3_IR = 0, 1, 2, 3
OZAKI_INT8, OZAKI_BF16 = 0, 1

_i64 = C.c_int64
_i32 = C.c_int32
_pd = C.c_void_p  # double*: passed as raw addresses
_pv = C.c_void_p

class Stats(C.Structure):
    """mrm_stats: counters of the host-pointer GEMM entries."""
    _fields_ = [("host_calls", C.c_int64), ("h2d_bytes", C.c_int64), ("d2h_bytes", C.c_int64),
                ("device_gemms", C.c_int64)]


class SolveStats(C.Structure):
    """mrm_solve_stats: what a refined solve did."""
    _fields_ = [("iterations", C.c_int32), ("fell_back", C.c_int32), ("gmres", C.c_int32),
                ("krylov_steps", C.c_int32), ("backward_error", C.c_double), ("lowp_iterations", C.c_int32),
                ("lowp_krylov_steps", C.c_int32), ("lowp_backward_error", C.c_double)]


_thr = [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64, _i32]

# name -> (restype, argtypes)
SIGNATURES = {
    "mrm_get_stats": (None, [C.POINTER(Stats)]),
    "mrm_reset_stats": (None, []),
    "mrm_lu_panel_test_deadline": (_i64, [_i64]),
    "mrm_lu_lookahead_available": (_i32, []),
    "mrm_thr_mult": (None, _thr),
    "mrm_thr_mult_t1": (None, _thr),
    "mrm_thr_mult_t2": (None, _thr),
    "mrm_covariance": (_i32, [_pd, _i64, _pd, _i64, _i64, _i64, _i32]),
    "mrm_covariance_workspace_bytes": (_i64, [_i64, _i64]),
    "mrm_dev_covariance": (_i32, [_pd, _i64, _i64, _i64, _pd, _i64, _i32, _pv, _i64, _pv]),
    "mrm_init": (_i32, []),
    "mrm_shutdown": (None, []),
    "mrm_last_error": (_i32, []),
    "mrm_last_error_message": (C.c_char_p, []),
    "mrm_version": (_i32, []),
    "mrm_mult": (None, [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64]),
    "mrm_mult_t2": (None, [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64]),
    "mrm_blocked_mult": (None, [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i32, _pd]),
    "mrm_blocked_mult_t1": (None, [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i32, _pd]),
    "mrm_blocked_mult_t2": (None, [_pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i32, _pd]),
    "mrm_dev_dgemm": (_i32, [_i32, _i64, _i64, _i64, _pd, _i64, _pd, _i64, _pd, _i64, _i32, _pv]),
    "mrm_dev_dgemm_noalloc": (_i32, [_i32, _i64, _i64, _i64, _pd, _i64, _pd, _i64, _pd, _i64, _i32, _pv]),
    "mrm_ozaki_workspace_bytes": (_i64, [_i64, _i64, _i64, _i32]),
    "mrm_dev_dgemm_ozaki": (_i32, [_i32, _i64, _i64, _i64, _pd, _i64, _pd, _i64, _pd, _i64, _i32, _i32, _pv, _i64,
                                   _pv]),
    "mrm_mult_ozaki": (_i32, [_i32, _pd, _i64, _pd, _pd, _i64, _i64, _i64, _i64, _i64, _i64, _i32, _i32]),
    "mrm_ozaki_workspace_bytes_ex": (_i64, [_i64, _i64, _i64, _i32, _i32]),
    "mrm_dev_dgemm_ozaki_ex": (_i32, [_i32, _i64, _i64, _i64, _pd, _i64, _pd, _i64, _pd, _i64, _i32, _i32, _i32, _pv,
                                      _i64, _pv]),
    "mrm_lowp_workspace_bytes": (_i64, [_i64, _i64, _i64, _i32]),
    "mrm_dev_gemm_lowp": (_i32, [_i32, _i64, _i64, _i64, _pd, _i64, _pd, _i64, _pd, _i64, _pv, _i64, _pv]),
    "mrm_dev_lineq_solve_ir": (_i32, [_i32, _pd, _i64, _i64, _pd, _i64, _pd, _i64, _i64, _i32, C.c_double, _pv, _pv,
                                      _pv, _pv]),
    "mrm_dev_lineq_solve_ir_stats": (_i32, [_i32, _pd, _i64, _i64, _pd, _i64, _pd, _i64, _i64, _i32, C.c_double,
                                            C.POINTER(SolveStats), _pv]),
    "mrm_lineq_solve_ex": (_i32, [_pd, _i64, _i64, _pd, _i64, _pd, _i64, _i64, _i64, _i32]),
    "mrm_batched_dgemm_plan": (_i32, [_pv, _i64, _i32, _pv]),
    "mrm_batched_dgemm_plan_ex": (_i32, [_pv, _i64, _i32, _i32, _pv]),
    "mrm_batched_lu_plan": (_i32, [_pv, _i64, _i32, _pv, _pv]),
    "mrm_batched_run": (_i32, [_pv, _pv]),
    "mrm_batched_plan_tile": (_i32, [_pv]),
    "mrm_batched_destroy": (None, [_pv]),
    "mrm_inverse_inplace": (_i32, [_pd, _i64, _i64]),
    "mrm_determinant": (_i32, [_pd, _i64, _i64, C.POINTER(C.c_double)]),
    "mrm_lu_decomp_inplace": (_i32, [_pd, _i64, _i64, _pv]),
    "mrm_lu_backsubst": (_i32, [_pd, _i64, _i64, _pv, _pd, _i64]),
    "mrm_lineq_solve": (_i32, [_pd, _i64, _i64, _pd, _i64, _pd, _i64, _i64, _i64]),
    "mrm_dev_lu_factor": (_i32, [_pd, _i64, _i64, _pv, _pv, _pv]),
    "mrm_dev_lu_solve": (_i32, [_pd, _i64, _i64, _pv, _pd, _i64, _pd, _i64, _i64, _pv]),
}

_lib = None


class MrmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mrmath_gpu error {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch (when present) bundles a HIP runtime with the same soname as ROCm's;
    # loading it first makes the library bind to that one runtime instead of a
    # second copy, so device pointers from torch tensors are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m mrmath_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def stats() -> Stats:
    """The host-entry counters (mrm_get_stats)."""
    st = Stats()
    load().mrm_get_stats(C.byref(st))
    return st


def check(code: int | None = None) -> None:
    """Raises MrmError for a non-zero status (the given one or the thread's last)."""
    lib = load()
    if code is None:
        code = lib.mrm_last_error()
    if code != MRM_OK:
        msg = lib.mrm_last_error_message()
        raise MrmError(code, msg.decode() if msg else "")
