"""Host-side mirror of mrMath's stub-switch multiply surface, bound to the GPU.

Same names and argument meaning as the reference's MatrixASMStubSwitch
procedures (src/MatrixASMStubSwitch.pas:64-85, 769-863); matrices are numpy
float64 arrays in the reference's layout (row-major, row stride = LineWidth in
bytes, possibly padded).  Every call goes through the C ABI of
libmrmath_gpu.so; there is no CPU path.

Differences from the reference, by design:
  * dimension errors raise (the reference asserts, then misbehaves): a dest
    without the product's shape raises ValueError before any call, operands
    whose sizes do not fit together raise MrmError from the library;
  * blockSize / mem are accepted and ignored (the GPU tiles itself).
"""
from __future__ import annotations

import numpy as np

from . import _lib

doNone, doAdd, doSub = _lib.DO_NONE, _lib.DO_ADD, _lib.DO_SUB

# Tuning globals the reference keeps in src/BlockSizeSetup.pas:28-41; kept for
# the MatrixMult dispatch rule (blocked when every dimension >= BlockedMatrixMultSize).
BlockMatrixCacheSize = 128
BlockedMatrixMultSize = 512


def _mat(a: np.ndarray, name: str) -> tuple[int, int, int, int]:
    """(address, width, height, LineWidth) of a row-major float64 matrix view."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or a.ndim != 2:
        raise TypeError(f"{name} must be a 2-D float64 numpy array")
    if a.size == 0:
        return a.ctypes.data, a.shape[1], a.shape[0], max(8, 8 * a.shape[1])
    if a.strides[1] != 8 or a.strides[0] < 8 * a.shape[1]:
        raise ValueError(f"{name} must have unit column stride and row stride >= width*8")
    return a.ctypes.data, a.shape[1], a.shape[0], a.strides[0]


def _product_shape(trans: int, w1: int, h1: int, w2: int, h2: int) -> tuple[int, int]:
    """(height, width) of the product for the reference's dimension rules:
    NN h1 x w2, T1 (mt1^T mt2) w1 x w2, T2 (mt1 mt2^T) h1 x h2
    (src/MatrixASMStubSwitch.pas:744, 786; src/BlockedMult.pas:502, 622)."""
    if trans == _lib.NN:
        return h1, w2
    if trans == _lib.TN:
        return w1, w2
    return h1, h2


def _operands(trans, dest, mt1, mt2):
    """Addresses and sizes of the three operands; dest must have exactly the
    product's shape (the library writes height x width rows at dest's stride,
    so a smaller dest would be written past its end)."""
    d, dw, dh, dlw = _mat(dest, "dest")
    p1, w1, h1, lw1 = _mat(mt1, "mt1")
    p2, w2, h2, lw2 = _mat(mt2, "mt2")
    want = _product_shape(trans, w1, h1, w2, h2)
    if (dh, dw) != want:
        raise ValueError(f"Dimension error: dest is {dh} x {dw}, the product is {want[0]} x {want[1]}")
    return d, dlw, p1, p2, w1, h1, w2, h2, lw1, lw2


def _call_mult(fn, trans, dest, mt1, mt2, *extra):
    fn(*_operands(trans, dest, mt1, mt2), *extra)
    _lib.check()
    return dest


def MatrixMult(dest: np.ndarray, mt1: np.ndarray, mt2: np.ndarray) -> np.ndarray:
    """dest = mt1 * mt2 (src/MatrixASMStubSwitch.pas:769-776)."""
    lib = _lib.load()
    dims = (mt1.shape[1], mt1.shape[0], mt2.shape[1], mt2.shape[0])
    if min(dims) >= BlockedMatrixMultSize:
        return _call_mult(lib.mrm_blocked_mult, _lib.NN, dest, mt1, mt2, BlockMatrixCacheSize, doNone, None)
    return _call_mult(lib.mrm_mult, _lib.NN, dest, mt1, mt2)


def MatrixMultEx(dest, mt1, mt2, blockSize: int = BlockMatrixCacheSize, op: int = doNone, mem=None):
    """dest (op)= mt1 * mt2 via the blocked pointer (src/MatrixASMStubSwitch.pas:778-781)."""
    return _call_mult(_lib.load().mrm_blocked_mult, _lib.NN, dest, mt1, mt2, blockSize, op, None)


def MatrixMultT1(dest, mt1, mt2):
    """dest = mt1^T * mt2 (src/MatrixASMStubSwitch.pas:812-817)."""
    return _call_mult(_lib.load().mrm_blocked_mult_t1, _lib.TN, dest, mt1, mt2, BlockMatrixCacheSize, doNone, None)


def MatrixMultT1Ex(dest, mt1, mt2, blockSize: int = BlockMatrixCacheSize, op: int = doNone, mem=None):
    """dest (op)= mt1^T * mt2 (src/MatrixASMStubSwitch.pas:819-822)."""
    return _call_mult(_lib.load().mrm_blocked_mult_t1, _lib.TN, dest, mt1, mt2, blockSize, op, None)


def MatrixMultOzaki(dest, mt1, mt2, slices: int = 7, trans: int = _lib.NN, op: int = doNone):
    """dest (op)= op(mt1) * op(mt2) in FP64, computed from `slices` 7-bit INT8
    slices per operand on the INT8 matrix cores (trans: NN = MatrixMult,
    TN = MatrixMultT1, NT = MatrixMultT2 dimension rules).  Error per element
    <= 4 (slices+1) K 2^(-7 slices) max|row of op(mt1)| max|column of op(mt2)|
    plus FP64 rounding.  No reference counterpart (the reference multiplies in
    FP64 only)."""
    if trans not in (_lib.NN, _lib.TN, _lib.NT):
        raise ValueError(f"trans must be NN, TN or NT, not {trans}")
    _lib.check(_lib.load().mrm_mult_ozaki(trans, *_operands(trans, dest, mt1, mt2), op, slices))
    return dest


def MatrixMultT2(dest, mt1, mt2):
    """dest = mt1 * mt2^T via multT2Func (src/MatrixASMStubSwitch.pas:853-858)."""
    return _call_mult(_lib.load().mrm_mult_t2, _lib.NT, dest, mt1, mt2)


def MatrixMultT2Ex(dest, mt1, mt2, blockSize: int = BlockMatrixCacheSize, op: int = doNone, mem=None):
    """dest (op)= mt1 * mt2^T (src/MatrixASMStubSwitch.pas:860-863)."""
    return _call_mult(_lib.load().mrm_blocked_mult_t2, _lib.NT, dest, mt1, mt2, blockSize, op, None)


def ThrMatrixMult(dest, mt1, mt2, op: int = doNone):
    """dest (op)= mt1 * mt2 (src/ThreadedMatrixOperations.pas:729-735): one
    device product for the whole matrix instead of the thread pool's tiles."""
    return _call_mult(_lib.load().mrm_thr_mult, _lib.NN, dest, mt1, mt2, op)


def ThrMatrixMultT1(dest, mt1, mt2, op: int = doNone):
    """dest (op)= mt1^T * mt2 (src/ThreadedMatrixOperations.pas:34)."""
    return _call_mult(_lib.load().mrm_thr_mult_t1, _lib.TN, dest, mt1, mt2, op)


def ThrMatrixMultT2(dest, mt1, mt2, op: int = doNone):
    """dest (op)= mt1 * mt2^T (src/ThreadedMatrixOperations.pas:40)."""
    return _call_mult(_lib.load().mrm_thr_mult_t2, _lib.NT, dest, mt1, mt2, op)


def aligned_matrix(width: int, height: int) -> np.ndarray:
    """Zeroed matrix with the reference's storage rules (src/MatrixASMStubSwitch.pas:570-586):
    rows padded to a multiple of 4 doubles, 32-byte aligned base.  Returns the
    logical (height x width) view."""
    lw = (width + 3) // 4 * 4
    raw = np.zeros(height * lw + 4, dtype=np.float64)
    off = (-raw.ctypes.data // 8) % 4
    return raw[off:off + height * lw].reshape(height, lw)[:, :width]


class TDoubleMatrix:
    """Minimal mirror of the reference's TDoubleMatrix multiply entry points
    (src/DblMatrix.pas:2110-2198): each allocates the result and calls the
    stub-switch procedure."""

    def __init__(self, width: int, height: int, data: np.ndarray | None = None):
        self.data = aligned_matrix(width, height)
        if data is not None:
            self.data[...] = data

    @property
    def Width(self) -> int:
        return self.data.shape[1]

    @property
    def Height(self) -> int:
        return self.data.shape[0]

    def Mult(self, Value: "TDoubleMatrix") -> "TDoubleMatrix":
        res = TDoubleMatrix(Value.Width, self.Height)
        MatrixMult(res.data, self.data, Value.data)
        return res

    def MultT1(self, Value: "TDoubleMatrix") -> "TDoubleMatrix":
        res = TDoubleMatrix(Value.Width, self.Width)
        MatrixMultT1Ex(res.data, self.data, Value.data, BlockMatrixCacheSize, doNone)
        return res

    def MultT2(self, Value: "TDoubleMatrix") -> "TDoubleMatrix":
        res = TDoubleMatrix(Value.Height, self.Height)
        MatrixMultT2Ex(res.data, self.data, Value.data, BlockMatrixCacheSize, doNone)
        return res


class TThreadedMatrix(TDoubleMatrix):
    """Mirror of TThreadedMatrix.Mult (src/ThreadedMatrix.pas:347-357): the
    result of the whole product through ThrMatrixMult."""

    def Mult(self, Value: "TDoubleMatrix") -> "TThreadedMatrix":
        res = TThreadedMatrix(Value.Width, self.Height)
        ThrMatrixMult(res.data, self.data, Value.data)
        return res
