"""Batched small GEMMs and linear solves on the GPU (persistent launches
replayed from a hipGraph).  The reference has no batched entry; its callers
loop over MatrixMult / MatrixLinEQSolve per problem (src/MatrixASMStubSwitch.pas
:769-776, src/LinAlgLU.pas:481-592).  Matrices are device tensors (torch,
float64, row-major, unit column stride); descriptors hold their addresses, so
the tensors must outlive the plan.
"""
from __future__ import annotations

import ctypes as C

from . import _lib


class GemmDesc(C.Structure):
    _fields_ = [("M", C.c_int64), ("N", C.c_int64), ("K", C.c_int64),
                ("A", C.c_void_p), ("lda", C.c_int64), ("B", C.c_void_p), ("ldb", C.c_int64),
                ("C", C.c_void_p), ("ldc", C.c_int64)]


class LuDesc(C.Structure):
    _fields_ = [("n", C.c_int64), ("A", C.c_void_p), ("lda", C.c_int64), ("b", C.c_void_p), ("x", C.c_void_p)]


def _rowstride(t) -> int:
    if t.dim() != 2 or t.stride(1) != 1 or t.dtype.itemsize != 8:
        raise ValueError("matrices must be 2-D float64 tensors with unit column stride")
    return max(t.stride(0), t.shape[1], 1) * 8


class _Plan:
    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @property
    def tile(self) -> int:
        return _lib.load().mrm_batched_plan_tile(self._h)

    def run(self, stream=None) -> None:
        """Replays the plan on `stream` (a torch stream or raw handle; default:
        torch's current stream)."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        elif hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        _lib.check(_lib.load().mrm_batched_run(self._h, stream))

    def close(self) -> None:
        if self._h:
            _lib.load().mrm_batched_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dgemm_plan(problems, tile: int = 0, mode: int = _lib.SOLVE_FP64) -> _Plan:
    """problems: sequence of (A, B, C) float64 device tensors, C = A @ B.
    tile 0 lets the autotuner pick 32, 64 or 128 (runs the batch).  mode
    SOLVE_FP32_IR: FP32 pieces with error compensation on the FP32 matrix
    cores, FP64 sums per K chunk of 16 (tiles 64 / 128)."""
    descs = (GemmDesc * max(len(problems), 1))()
    for i, (A, B, Cm) in enumerate(problems):
        descs[i] = GemmDesc(A.shape[0], B.shape[1], A.shape[1], A.data_ptr(), _rowstride(A), B.data_ptr(),
                            _rowstride(B), Cm.data_ptr(), _rowstride(Cm))
    h = C.c_void_p()
    _lib.check(_lib.load().mrm_batched_dgemm_plan_ex(descs, len(problems), tile, mode, C.byref(h)))
    return _Plan(h)


def lu_plan(systems, mode: int = _lib.SOLVE_FP64, status=None) -> _Plan:
    """systems: sequence of (A, b, x) float64 device tensors (A n x n with
    n <= 128, b and x of n entries); status: optional int32 device tensor of
    len(systems) entries (MRM_OK / MRM_SINGULAR after each run)."""
    descs = (LuDesc * max(len(systems), 1))()
    for i, (A, b, x) in enumerate(systems):
        descs[i] = LuDesc(A.shape[0], A.data_ptr(), _rowstride(A), b.data_ptr(), x.data_ptr())
    h = C.c_void_p()
    st = status.data_ptr() if status is not None else None
    _lib.check(_lib.load().mrm_batched_lu_plan(descs, len(systems), mode, st, C.byref(h)))
    return _Plan(h)
