"""Host-side mirror of mrMath's LU entry points, bound to the GPU.

Same names, argument meaning and result codes as src/LinAlgLU.pas:26-62
(MatrixLUDecompInPlace, MatrixLUBackSubst, MatrixLinEQSolve); matrices are
numpy float64 arrays in the reference's layout.  Every call goes through
libmrmath_gpu.so.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .stubswitch import _mat

leOk, leSingular = _lib.MRM_OK, _lib.MRM_SINGULAR
smFP64, smFP32IR, smBF16IR = _lib.SOLVE_FP64, _lib.SOLVE_FP32_IR, _lib.SOLVE_BF16_IR


class ELinEQSingularException(Exception):
    """Raised by the object API on a singular system (src/DblMatrix.pas:1353-1356)."""


def _status(rc: int) -> int:
    if rc not in (leOk, leSingular):
        _lib.check(rc)
    return rc


def MatrixLUDecompInPlace(A: np.ndarray, indx: np.ndarray) -> int:
    """In-place LU with partial pivoting; indx (int32, width entries) receives the
    pivot rows.  Returns leOk or leSingular (src/LinAlgLU.pas:283-295)."""
    p, w, h, lw = _mat(A, "A")
    if w != h:
        raise ValueError("Dimension error: A must be square")
    if indx.dtype != np.int32 or indx.size < w:
        raise ValueError("indx must be an int32 array of width entries")
    return _status(_lib.load().mrm_lu_decomp_inplace(p, lw, w, indx.ctypes.data))


def MatrixLUBackSubst(LUDecomp: np.ndarray, indx: np.ndarray, B: np.ndarray) -> None:
    """Solves LU x = P b in place for the first column of B (src/LinAlgLU.pas:315-385)."""
    p, w, _, lw = _mat(LUDecomp, "LUDecomp")
    pb, _, hb, lwb = _mat(B, "B")
    if hb != w:
        raise ValueError("Dimension error")
    _lib.check(_lib.load().mrm_lu_backsubst(p, lw, w, indx.ctypes.data, pb, lwb))


def MatrixLinEQSolve(A: np.ndarray, B: np.ndarray, X: np.ndarray, NumRefinments: int = 0) -> int:
    """X = A^-1 B (width2 = B's width right-hand sides) plus NumRefinments FP64
    refinement steps (src/LinAlgLU.pas:481-592).  Returns leOk or leSingular."""
    pa, w, h, lwa = _mat(A, "A")
    pb, w2, hb, lwb = _mat(B, "B")
    px, wx, hx, lwx = _mat(X, "X")
    if w != h or hb != w or hx != w or wx != w2:
        raise ValueError("Dimension error")
    return _status(_lib.load().mrm_lineq_solve(pa, lwa, w, pb, lwb, px, lwx, w2, NumRefinments))


def MatrixLinEQSolveEx(A: np.ndarray, B: np.ndarray, X: np.ndarray, NumRefinments: int = 0,
                       mode: int = smFP64) -> int:
    """MatrixLinEQSolve with a factorisation precision: smFP64 is
    MatrixLinEQSolve; smFP32IR / smBF16IR run the LU's trailing updates on the
    FP32 / BF16 matrix cores and refine in FP64 to a normwise backward error of
    1e-14 (at most NumRefinments steps, 50 when 0).  Returns leOk or
    leSingular; raises MrmError when refinement does not converge."""
    pa, w, h, lwa = _mat(A, "A")
    pb, w2, hb, lwb = _mat(B, "B")
    px, wx, hx, lwx = _mat(X, "X")
    if w != h or hb != w or hx != w or wx != w2:
        raise ValueError("Dimension error")
    return _status(_lib.load().mrm_lineq_solve_ex(pa, lwa, w, pb, lwb, px, lwx, w2, NumRefinments, mode))


def MatrixInverseInPlace(A: np.ndarray) -> int:
    """A := A^-1 (src/LinAlgLU.pas:386-427). Returns leOk or leSingular (A unchanged)."""
    p, w, h, lw = _mat(A, "A")
    if w != h:
        raise ValueError("Dimension error: A must be square")
    return _status(_lib.load().mrm_inverse_inplace(p, lw, w))


def MatrixDeterminant(A: np.ndarray) -> float:
    """det(A) from the LU (src/LinAlgLU.pas:429-478); 0 for a singular matrix."""
    import ctypes as C
    p, w, h, lw = _mat(A, "A")
    if w != h:
        raise ValueError("Dimension error: A must be square")
    d = C.c_double()
    _status(_lib.load().mrm_determinant(p, lw, w, C.byref(d)))
    return d.value


def SolveLinEQ(A: np.ndarray, B: np.ndarray, NumRefinments: int = 0) -> np.ndarray:
    """TDoubleMatrix.SolveLinEQ (src/DblMatrix.pas:1343-1362): returns X, raises on singular."""
    X = np.zeros((A.shape[0], B.shape[1]))
    if MatrixLinEQSolve(A, B, X, NumRefinments) == leSingular:
        raise ELinEQSingularException("Matrix is singular")
    return X
