"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the GPU library and the CPU timing
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module.  Matrices are numpy float64 arrays in the reference's
layout (row-major; LineWidth = row stride in bytes).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_i64, _i32, _p = C.c_int64, C.c_int32, C.c_void_p


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int32)]


_SIG = {
    "orc_mt_init": (None, [C.POINTER(MT), _i32]),
    "orc_mt_init_by_array": (None, [C.POINTER(MT), C.POINTER(C.c_uint32), _i32]),
    "orc_mt_int32": (C.c_uint32, [C.POINTER(MT)]),
    "orc_mt_res53": (C.c_double, [C.POINTER(MT)]),
    "orc_mt_randint": (_i32, [C.POINTER(MT), _i32]),
    "orc_randgauss": (C.c_double, [C.POINTER(MT), C.POINTER(_i32), C.POINTER(C.c_double)]),
    "orc_fill_rand": (None, [_p, _i64, _i64, _i64, C.POINTER(_i32)]),
    "orc_fill_pm1": (None, [_p, _i64, _i64, _i64, _i32]),
    "orc_fill_pm1_wide": (None, [_p, _i64, _i64, _i64, _i32, _i32]),
    "orc_generic_mtx_mult": (None, [_p, _i64, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64]),
    "orc_generic_mtx_mult_transp": (None, [_p, _i64, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64]),
    "orc_generic_transp_mtx_mult": (None, [_p, _i64, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64]),
    "orc_blocked_mult": (None, [_i32, _p, _i64, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i32]),
    "orc_ozaki_gemm": (_i32, [_p, _i64, _p, _i64, _p, _i64, _i64, _i64, _i64, _i32, _i32, _i64]),
    "orc_lu_decomp_inplace": (_i32, [_p, _i64, _i64, _p]),
    "orc_lu_backsubst": (None, [_p, _i64, _i64, _p, _p, _i64]),
    "orc_lineq_solve": (_i32, [_p, _i64, _i64, _p, _i64, _p, _i64, _i64, _i64]),
    "orc_inverse_inplace": (_i32, [_p, _i64, _i64]),
    "orc_determinant": (C.c_double, [_p, _i64, _i64]),
    "orc_gemm_normwise_err": (C.c_double, [_p, _i64, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64, _i64, _i64]),
    "orc_backward_error": (C.c_double, [_p, _i64, _i64, _p, _i64, _p, _i64, _i64]),
    "orc_covariance": (None, [_p, _i64, _p, _i64, _i64, _i64, _i32]),
    "orc_thr_lu_decomp": (_i32, [_p, _i64, _i64, _p, C.c_int]),
    "orc_thr_lineq_solve": (_i32, [_p, _i64, _i64, _p, _i64, _p, _i64, _i64, _i64, C.c_int]),
    "orc_fast_mult": (C.c_int, [_p, _i64, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64, C.c_int, C.c_int]),
    "orc_num_cores_for_mult": (C.c_int, [_i64, _i64, _i64, _i64, C.c_int]),
}

_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        l = C.CDLL(str(LIB_PATH))
        for n, (res, args) in _SIG.items():
            f = getattr(l, n)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _m(a: np.ndarray):
    assert a.dtype == np.float64 and a.ndim == 2 and a.strides[1] == 8
    return a.ctypes.data, a.shape[1], a.shape[0], a.strides[0]


# ---- random inputs ----------------------------------------------------------
class MTStream:
    """MT19937 as the reference's TRandomGenerator(raMersenneTwister)."""

    def __init__(self, seed: int | None = None, key: list[int] | None = None):
        self.s = MT()
        if key is not None:
            arr = (C.c_uint32 * len(key))(*key)
            lib().orc_mt_init_by_array(C.byref(self.s), arr, len(key))
        else:
            lib().orc_mt_init(C.byref(self.s), int(seed))
        self._flag = _i32(0)
        self._g2 = C.c_double(0.0)

    def int32(self) -> int:
        return lib().orc_mt_int32(C.byref(self.s))

    def res53(self) -> float:
        return lib().orc_mt_res53(C.byref(self.s))

    def randint(self, rng: int) -> int:
        return lib().orc_mt_randint(C.byref(self.s), rng)

    def gauss(self) -> float:
        return lib().orc_randgauss(C.byref(self.s), C.byref(self._flag), C.byref(self._g2))


def rand_matrix(width: int, height: int, seed: int, lw_doubles: int | None = None):
    """TDoubleMatrix.CreateRand(width, height, raMersenneTwister, seed): returns (M, next seed)."""
    lw = lw_doubles or width
    buf = np.zeros((height, lw))
    s = _i32(seed)
    lib().orc_fill_rand(buf.ctypes.data, lw * 8, width, height, C.byref(s))
    return buf[:, :width], s.value


def pm1_matrix(width: int, height: int, seed: int) -> np.ndarray:
    """Entries 2*res53 - 1 from init_genrand(seed), row-major."""
    buf = np.empty((height, width))
    lib().orc_fill_pm1(buf.ctypes.data, width * 8, width, height, seed)
    return buf


def pm1_wide_matrix(width: int, height: int, seed: int, emax: int = 20) -> np.ndarray:
    buf = np.empty((height, width))
    lib().orc_fill_pm1_wide(buf.ctypes.data, width * 8, width, height, seed, emax)
    return buf


# ---- GEMM ---------------------------------------------------------------------
def generic_mult(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """GenericMtxMult: A (h1 x w1) * B (h2 x w2)."""
    pa, w1, h1, lw1 = _m(A)
    pb, w2, h2, lw2 = _m(B)
    assert w1 == h2
    C_ = np.zeros((h1, w2))
    lib().orc_generic_mtx_mult(C_.ctypes.data, w2 * 8, pa, pb, w1, h1, w2, h2, lw1, lw2)
    return C_


def generic_mult_transp(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """GenericMtxMultTransp: A * B^T."""
    pa, w1, h1, lw1 = _m(A)
    pb, w2, h2, lw2 = _m(B)
    assert w1 == w2
    C_ = np.zeros((h1, h2))
    lib().orc_generic_mtx_mult_transp(C_.ctypes.data, h2 * 8, pa, pb, w1, h1, w2, h2, lw1, lw2)
    return C_


def generic_transp_mult(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """GenericTranspMtxMult: A^T * B."""
    pa, w1, h1, lw1 = _m(A)
    pb, w2, h2, lw2 = _m(B)
    assert h1 == h2
    C_ = np.zeros((w1, w2))
    lib().orc_generic_transp_mtx_mult(C_.ctypes.data, w2 * 8, pa, pb, w1, h1, w2, h2, lw1, lw2)
    return C_


def blocked_mult(trans: int, dest: np.ndarray, A: np.ndarray, B: np.ndarray, block: int = 128, op: int = 0):
    pd, _, _, lwd = _m(dest)
    pa, w1, h1, lw1 = _m(A)
    pb, w2, h2, lw2 = _m(B)
    lib().orc_blocked_mult(trans, pd, lwd, pa, pb, w1, h1, w2, h2, lw1, lw2, block, op)
    return dest


def fast_mult(dest: np.ndarray, A: np.ndarray, B: np.ndarray, threads: int, block: int = 128, op: int = 0) -> int:
    """Restated threaded AVX2/FMA blocked path (CPU baseline). Returns threads used."""
    pd, _, _, lwd = _m(dest)
    pa, w1, h1, lw1 = _m(A)
    pb, w2, h2, lw2 = _m(B)
    return lib().orc_fast_mult(pd, lwd, pa, pb, w1, h1, w2, lw1, lw2, block, op, threads)


def gemm_err(Cm: np.ndarray, Cref: np.ndarray, A: np.ndarray, B: np.ndarray) -> float:
    pc, wc, hc, lwc = _m(Cm)
    pr, _, _, lwr = _m(Cref)
    pa, wa, ha, lwa = _m(A)
    pb, wb, hb, lwb = _m(B)
    return lib().orc_gemm_normwise_err(pc, lwc, pr, lwr, wc, hc, pa, lwa, wa, ha, pb, lwb, wb, hb)


OZAKI_KCHUNK = 16384


def ozaki_gemm(A: np.ndarray, B: np.ndarray, slices: int, C: np.ndarray | None = None, op: int = 0,
               kchunk: int = OZAKI_KCHUNK) -> np.ndarray:
    """C (op)= A @ B through `slices` 7-bit integer slices (the GPU's method,
    bit for bit).  Raises if an int32 chunk sum would overflow."""
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    if C is None:
        C = np.zeros((M, N))
    pc, _, _, lwc = _m(C)
    pa, _, _, lwa = _m(A)
    pb, _, _, lwb = _m(B)
    rc = lib().orc_ozaki_gemm(pc, lwc, pa, lwa, pb, lwb, M, N, K, op, slices, kchunk)
    if rc:
        raise ValueError(f"orc_ozaki_gemm failed ({rc})")
    return C


# ---- LU / solve ---------------------------------------------------------------
def lu_decomp_inplace(A: np.ndarray):
    pa, w, h, lw = _m(A)
    assert w == h
    indx = np.zeros(w, dtype=np.int32)
    rc = lib().orc_lu_decomp_inplace(pa, lw, w, indx.ctypes.data)
    return rc, indx


def lu_backsubst(LU: np.ndarray, indx: np.ndarray, B: np.ndarray, col: int = 0):
    pl, w, _, lwl = _m(LU)
    pb, _, _, lwb = _m(B)
    lib().orc_lu_backsubst(pl, lwl, w, indx.ctypes.data, pb + 8 * col, lwb)
    return B


def lineq_solve(A: np.ndarray, B: np.ndarray, refinements: int = 0):
    pa, w, _, lwa = _m(A)
    pb, w2, _, lwb = _m(B)
    X = np.zeros((w, w2))
    rc = lib().orc_lineq_solve(pa, lwa, w, pb, lwb, X.ctypes.data, w2 * 8, w2, refinements)
    return rc, X


def thr_lu_decomp(A: np.ndarray, threads: int):
    """The restated threaded LU (ThrMatrixLUDecomp) in place; (rc, indx)."""
    pa, w, h, lw = _m(A)
    assert w == h
    indx = np.zeros(w, dtype=np.int32)
    rc = lib().orc_thr_lu_decomp(pa, lw, w, indx.ctypes.data, threads)
    return rc, indx


def thr_lineq_solve(A: np.ndarray, B: np.ndarray, refinements: int = 0, threads: int = 1):
    """The restated ThrMatrixLinEQSolve (CPU baseline of the solve); (rc, X)."""
    pa, w, _, lwa = _m(A)
    pb, w2, _, lwb = _m(B)
    X = np.zeros((w, w2))
    rc = lib().orc_thr_lineq_solve(pa, lwa, w, pb, lwb, X.ctypes.data, w2 * 8, w2, refinements, threads)
    return rc, X


def inverse_inplace(A: np.ndarray) -> int:
    pa, w, _, lw = _m(A)
    return lib().orc_inverse_inplace(pa, lw, w)


def determinant(A: np.ndarray) -> float:
    pa, w, _, lw = _m(A)
    return lib().orc_determinant(pa, lw, w)


def backward_error(A: np.ndarray, X: np.ndarray, B: np.ndarray) -> float:
    pa, n, _, lwa = _m(A)
    px, nrhs, _, lwx = _m(X)
    pb, _, _, lwb = _m(B)
    return lib().orc_backward_error(pa, lwa, n, px, lwx, pb, lwb, nrhs)


# ---- statistics -----------------------------------------------------------------
def covariance(A: np.ndarray, unbiased: bool = True) -> np.ndarray:
    """Covariance (src/Statistics.pas:599-639): width x width, or the 1 x 1
    variance of a single row or column."""
    pa, w, h, lwa = _m(A)
    n = 1 if (w == 1 or h == 1) else w
    D = np.zeros((n, n))
    lib().orc_covariance(D.ctypes.data, n * 8, pa, lwa, w, h, 1 if unbiased else 0)
    return D
