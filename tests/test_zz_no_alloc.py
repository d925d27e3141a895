"""The allocation-free device GEMM (mrm_dev_dgemm_noalloc) against the
allocating entry it mirrors (mrm_dev_dgemm), through the C ABI.

Both run the same kernels with the same K split, so the results must be
identical bit for bit; two calls in a row (the second smaller) check that the
shared scratch array carries nothing from one call into the next.
"""
import pytest

from mrmath_amd import _lib

pytestmark = pytest.mark.gpu


def _pair(lib, trans, M, N, K, same, seed, op=0):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    rowsA, colsA = (K, M) if trans == _lib.TN else (M, K)
    rowsB, colsB = (N, K) if trans == _lib.NT else (K, N)
    A = torch.rand(rowsA, colsA, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = A if same else torch.rand(rowsB, colsB, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C0 = torch.rand(M, N, dtype=torch.float64, device="cuda", generator=g)
    C1, C2 = C0.clone(), C0.clone()
    s = torch.cuda.current_stream().cuda_stream
    args = (trans, M, N, K, A.data_ptr(), colsA * 8, B.data_ptr(), colsB * 8)
    assert lib.mrm_dev_dgemm(*args, C1.data_ptr(), N * 8, op, s) == 0
    rc = lib.mrm_dev_dgemm_noalloc(*args, C2.data_ptr(), N * 8, op, s)
    assert rc == 0, _lib.load().mrm_last_error_message()
    torch.cuda.synchronize()
    return C1, C2


def test_noalloc_equals_the_allocating_entry_twice():
    import torch
    lib = _lib.load()
    # 1: the covariance shape X^T X, X 200000 x 512: symmetric tiles, 51 K chunks
    C1, C2 = _pair(lib, _lib.TN, 512, 512, 200000, True, 2001)
    assert torch.equal(C1, C2)
    # 2: smaller, after the first: fewer chunks in the same scratch
    C1, C2 = _pair(lib, _lib.TN, 256, 256, 60000, True, 2002)
    assert torch.equal(C1, C2)


@pytest.mark.parametrize("trans,M,N,K,same,op", [
    (_lib.NN, 256, 192, 50000, False, 0),     # split K, general tiles
    (_lib.NT, 300, 300, 20000, True, 2),      # X X^T, doSub
    (_lib.TN, 130, 70, 9000, False, 1),       # ragged, doAdd
    (_lib.NN, 1000, 1000, 1000, False, 0),    # no split: no scratch used
])
def test_noalloc_shapes(trans, M, N, K, same, op):
    import torch
    lib = _lib.load()
    C1, C2 = _pair(lib, trans, M, N, K, same, M + N + K, op)
    assert torch.equal(C1, C2)


def test_noalloc_argument_errors():
    lib = _lib.load()
    assert lib.mrm_dev_dgemm_noalloc(_lib.NN, 4, 4, 4, None, 32, None, 32, None, 32, 0, None) == _lib.MRM_BAD_ARG
    assert lib.mrm_dev_dgemm_noalloc(5, 4, 4, 4, 8, 32, 8, 32, 8, 32, 0, None) == _lib.MRM_BAD_ARG
