"""Threaded callers of the multiply through the C ABI.

The reference cuts one product into up to 64 tiles for its thread pool
(ThrMatrixMultEx2, src/ThreadedMatrixOperations.pas:612-718) and calls the
blocked pointer once per tile from the pool threads.  Two things are
checked: ThrMatrixMult as one entry stages each operand once and runs one
device product; and tile calls from 16-64 threads at once (each its own
product, on its own pooled stream) reproduce the oracle's product.

Bar: ||C - Cref||_F / (||A||_F ||B||_F) <= 4*k*2^-53; integer data bit-exact.
"""
import threading

import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import stubswitch as ss

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def test_thr_mult_is_one_staged_product():
    lib = _lib.load()
    M, K, N = 1000, 700, 900
    A = orc.pm1_matrix(K, M, 4501)
    B = orc.pm1_matrix(N, K, 4502)
    D = np.zeros((M, N))
    lib.mrm_reset_stats()
    ss.ThrMatrixMult(D, A, B)
    st = _lib.stats()
    assert st.host_calls == 1 and st.device_gemms == 1
    assert st.h2d_bytes == (A.size + B.size) * 8
    assert st.d2h_bytes == D.size * 8
    assert orc.gemm_err(D, orc.generic_mult(A, B), A, B) <= 4 * K * U


@pytest.mark.parametrize("which", ["nn", "t1", "t2"])
@pytest.mark.parametrize("op", [ss.doNone, ss.doAdd, ss.doSub])
def test_thr_mult_layouts_and_ops_integer_exact(which, op):
    rng = np.random.default_rng(7 + op)
    M, K, N = 130, 257, 66
    if which == "nn":
        A = rng.integers(-8, 9, (M, K)).astype(np.float64)
        B = rng.integers(-8, 9, (K, N)).astype(np.float64)
        fn, ref = ss.ThrMatrixMult, A @ B
    elif which == "t1":
        A = rng.integers(-8, 9, (K, M)).astype(np.float64)
        B = rng.integers(-8, 9, (K, N)).astype(np.float64)
        fn, ref = ss.ThrMatrixMultT1, A.T @ B
    else:
        A = rng.integers(-8, 9, (M, K)).astype(np.float64)
        B = rng.integers(-8, 9, (N, K)).astype(np.float64)
        fn, ref = ss.ThrMatrixMultT2, A @ B.T
    D0 = rng.integers(-50, 51, (M, N)).astype(np.float64)
    D = D0.copy()
    # integer products and sums stay below 2^53: exact in every order
    fn(D, A, B, op)
    want = ref if op == ss.doNone else (D0 + ref if op == ss.doAdd else D0 - ref)
    assert np.array_equal(D, want)


def test_threaded_matrix_class():
    A = orc.pm1_matrix(300, 200, 11)
    B = orc.pm1_matrix(150, 300, 12)
    a = ss.TThreadedMatrix(300, 200, A)
    c = a.Mult(ss.TDoubleMatrix(150, 300, B))
    assert (c.Width, c.Height) == (150, 200)
    assert orc.gemm_err(np.ascontiguousarray(c.data), orc.generic_mult(A, B), A, B) <= 4 * 300 * U


def _ex2_tiles(height1: int, width2: int, cores: int):
    """The tile split of ThrMatrixMultEx2 (src/ThreadedMatrixOperations.pas:632-699):
    cores/2 row groups by 2 column groups, widths rounded up to even."""
    cpud2 = cores // 2
    heightCores = cpud2
    cpud2 = cores // heightCores
    thrWidth = width2 // cpud2 + (0 if width2 % cpud2 == 0 else 1)
    thrWidth += thrWidth & 1
    thrHeight = height1 // heightCores + (0 if height1 % heightCores == 0 else 1)
    tiles = []
    for j in range(cpud2):
        for i in range(heightCores):
            h = thrHeight if height1 > (i + 1) * thrHeight else height1 - i * thrHeight
            w = thrWidth if width2 > (j + 1) * thrWidth else width2 - j * thrWidth
            if h <= 0 or w <= 0:
                break
            tiles.append((i * thrHeight, h, j * thrWidth, w))
    return tiles


@pytest.mark.parametrize("cores", [16, 32, 64])
def test_pool_threads_tile_calls_at_once(cores):
    """cores threads call the blocked pointer together, each on its tile of
    dest with sub-matrix views of mt1 (rows) and mt2 (columns)."""
    lib = _lib.load()
    M, K, N = 1030, 517, 770
    A = orc.pm1_matrix(K, M, 4601 + cores)
    B = orc.pm1_matrix(N, K, 4701 + cores)
    D = np.zeros((M, N))
    tiles = _ex2_tiles(M, N, cores)
    assert len(tiles) == cores
    errs = []
    start = threading.Barrier(len(tiles))

    def work(t):
        r0, h, c0, w = t
        try:
            start.wait()
            ss.MatrixMultEx(D[r0:r0 + h, c0:c0 + w], A[r0:r0 + h], B[:, c0:c0 + w], 128, ss.doNone)
        except Exception as ex:  # pragma: no cover
            errs.append(ex)

    lib.mrm_reset_stats()
    th = [threading.Thread(target=work, args=(t,)) for t in tiles]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    st = _lib.stats()
    assert st.host_calls == len(tiles) and st.device_gemms == len(tiles)
    assert orc.gemm_err(D, orc.generic_mult(A, B), A, B) <= 4 * K * U
    # one ThrMatrixMult call instead: one product, each operand staged once
    D2 = np.zeros((M, N))
    lib.mrm_reset_stats()
    ss.ThrMatrixMult(D2, A, B)
    st = _lib.stats()
    assert st.device_gemms == 1 and st.h2d_bytes == (A.size + B.size) * 8
    assert orc.gemm_err(D2, orc.generic_mult(A, B), A, B) <= 4 * K * U


def test_threads_with_errors_do_not_disturb_others():
    """A tile call with mismatched sizes fails in its own thread (thread-local
    status) while the other threads' products complete."""
    lib = _lib.load()
    A = orc.pm1_matrix(64, 64, 1)
    B = orc.pm1_matrix(64, 64, 2)
    good = [np.zeros((64, 64)) for _ in range(8)]
    codes = {}

    def ok(i):
        ss.MatrixMultEx(good[i], A, B)
        codes[i] = lib.mrm_last_error()

    def bad():
        D = np.zeros((64, 64))
        lib.mrm_blocked_mult(D.ctypes.data, 512, A.ctypes.data, B.ctypes.data, 64, 64, 64, 63, 512, 512, 128, 0,
                             None)
        codes["bad"] = lib.mrm_last_error()

    th = [threading.Thread(target=ok, args=(i,)) for i in range(8)] + [threading.Thread(target=bad)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert codes["bad"] == _lib.MRM_BAD_DIMS
    ref = orc.generic_mult(A, B)
    for i in range(8):
        assert codes[i] == 0
        assert orc.gemm_err(good[i], ref, A, B) <= 4 * 64 * U
