"""GPU parity of the FP64 GEMM family against the oracle, through the C ABI.

Bar (stated once, used everywhere): ||C - Cref||_F / (||A||_F ||B||_F) <= 4*k*2^-53
against the oracle's GenericMtxMult restatement; integer-valued inputs and
the reference's known-answer tests must match bit for bit.
"""
import threading

import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import stubswitch as ss

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def tol(k):
    return 4 * k * U


def _strided(h, w, lw, fill, offset=0):
    """(h x w) view with row stride lw doubles starting `offset` doubles into a buffer."""
    raw = np.zeros(offset + h * lw + 8)
    v = raw[offset:offset + h * lw].reshape(h, lw)[:, :w]
    v[...] = fill
    return v


def test_kat_mult():
    """Test/TestSimpleMatrixOperations.pas:1712-1729 through MatrixMult (small path)."""
    mt1 = np.array([0, 1, 2, 3, 4, 5], dtype=np.float64)
    mt2 = np.array([1, 2, 3, 4, 5, 6], dtype=np.float64)
    A = mt1.reshape(2, 3)
    B = mt2.reshape(3, 2)
    D = np.zeros((2, 2))
    ss.MatrixMult(D, A, B)
    assert np.array_equal(D.ravel(), [13, 16, 40, 52])
    A2 = np.lib.stride_tricks.as_strided(mt1, (2, 2), (24, 8))
    B2 = np.lib.stride_tricks.as_strided(mt2[1:], (2, 2), (24, 8))
    D2 = np.zeros((2, 2))
    ss.MatrixMult(D2, A2, B2)
    assert np.array_equal(D2.ravel(), [5, 6, 26, 33])


@pytest.mark.parametrize("k,h,expect", [(610, 62, 1220), (611, 62, 1222), (611, 63, 1222), (401, 60, 802)])
def test_closed_form_mult(k, h, expect):
    """Test/TestMatrixClass.pas:473-523 via TDoubleMatrix.Mult."""
    a = ss.TDoubleMatrix(k, h, np.ones((h, k)))
    b = ss.TDoubleMatrix(1, k, 2 * np.ones((k, 1)))
    c = a.Mult(b)
    assert (c.Width, c.Height) == (1, h)
    assert np.allclose(c.data, expect, rtol=1e-12, atol=0)


def _ref_product(trans, A, B):
    if trans == 0:
        return orc.generic_mult(A, B)
    if trans == 1:
        return orc.generic_transp_mult(A, B)
    return orc.generic_mult_transp(A, B)


def _logical_ab(trans, A, B):
    Am = A if trans != 1 else A.T
    Bm = B if trans != 2 else B.T
    return Am, Bm


SHAPES = [(1, 1, 1), (2, 3, 5), (17, 33, 9), (128, 128, 16), (129, 127, 17), (255, 130, 300), (64, 300, 513)]


@pytest.mark.parametrize("trans", [0, 1, 2])
@pytest.mark.parametrize("M,N,K", SHAPES)
def test_gemm_random_vs_oracle(trans, M, N, K):
    seed = 1000 + M * 7 + N * 3 + K + trans
    Am = orc.pm1_matrix(K, M, seed)
    Bm = orc.pm1_matrix(N, K, seed + 1)
    A = Am if trans != 1 else np.ascontiguousarray(Am.T)
    B = Bm if trans != 2 else np.ascontiguousarray(Bm.T)
    D = np.zeros((M, N))
    fn = [ss.MatrixMultEx, ss.MatrixMultT1Ex, ss.MatrixMultT2Ex][trans]
    fn(D, A, B, 128, ss.doNone)
    ref = _ref_product(trans, A, B)
    assert orc.gemm_err(D, ref, Am, Bm) <= tol(K)


@pytest.mark.parametrize("trans", [0, 1, 2])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_gemm_ops_integer_exact(trans, op):
    """Integer inputs: every product and partial sum is exact, so doNone/doAdd/doSub
    must equal the oracle's blocked driver bit for bit (ragged 150x97x70)."""
    rng = np.random.default_rng(trans * 3 + op)
    M, N, K = 150, 97, 70
    Am = rng.integers(-64, 64, (M, K)).astype(np.float64)
    Bm = rng.integers(-64, 64, (K, N)).astype(np.float64)
    A = Am if trans != 1 else np.ascontiguousarray(Am.T)
    B = Bm if trans != 2 else np.ascontiguousarray(Bm.T)
    D0 = rng.integers(-1000, 1000, (M, N)).astype(np.float64)
    D = D0.copy()
    [ss.MatrixMultEx, ss.MatrixMultT1Ex, ss.MatrixMultT2Ex][trans](D, A, B, 128, op)
    E = D0.copy()
    orc.blocked_mult(trans, E, A, B, 128, op)
    assert np.array_equal(D, E)


@pytest.mark.parametrize("offset", [0, 1, 3])
def test_gemm_padded_and_unaligned_views(offset):
    """Sub-matrix views: padded LineWidth, odd element offsets (non-16-byte-aligned
    rows take the scalar-load path), destination padding untouched."""
    M, N, K = 70, 45, 133
    Am = orc.pm1_matrix(K, M, 71)
    Bm = orc.pm1_matrix(N, K, 72)
    A = _strided(M, K, K + 5, Am, offset)
    B = _strided(K, N, N + 3, Bm, offset)
    Draw = np.full(M * (N + 7) + 8, 7.0)
    D = Draw[:M * (N + 7)].reshape(M, N + 7)[:, :N]
    ss.MatrixMultEx(D, A, B, 48, ss.doNone)
    assert orc.gemm_err(np.ascontiguousarray(D), orc.generic_mult(Am, Bm), Am, Bm) <= tol(K)
    pad = Draw[:M * (N + 7)].reshape(M, N + 7)[:, N:]
    assert np.all(pad == 7.0)


def test_config1_1024_seed301():
    """The CPU configuration: 1024^3 with A, B from CreateRand(seed 301) chaining."""
    A, s = orc.rand_matrix(1024, 1024, 301)
    B, _ = orc.rand_matrix(1024, 1024, s)
    a = ss.TDoubleMatrix(1024, 1024, A)
    b = ss.TDoubleMatrix(1024, 1024, B)
    c = a.Mult(b)
    ref = orc.generic_mult(A, B)
    err = orc.gemm_err(np.ascontiguousarray(c.data), ref, A, B)
    assert err <= tol(1024)
    # a bit-identical restatement is impossible (different summation order), but the
    # element-wise agreement is far inside the reference tests' 1e-4 absolute tolerance
    assert np.max(np.abs(c.data - ref)) < 1e-10


def test_mult_t1_t2_classes():
    A = orc.pm1_matrix(40, 60, 5)
    B = orc.pm1_matrix(30, 60, 6)
    a = ss.TDoubleMatrix(40, 60, A)
    b = ss.TDoubleMatrix(30, 60, B)
    c = a.MultT1(b)
    assert (c.Width, c.Height) == (30, 40)
    assert orc.gemm_err(np.ascontiguousarray(c.data), orc.generic_transp_mult(A, B), A.T.copy(), B) <= tol(60)
    B2 = orc.pm1_matrix(40, 25, 7)
    b2 = ss.TDoubleMatrix(40, 25, B2)
    c2 = a.MultT2(b2)
    assert (c2.Width, c2.Height) == (25, 60)
    assert orc.gemm_err(np.ascontiguousarray(c2.data), orc.generic_mult_transp(A, B2), A, B2.T.copy()) <= tol(40)


def test_dimension_errors_and_empty():
    A = np.ones((3, 4))
    B = np.ones((5, 2))
    D = np.zeros((3, 2))
    with pytest.raises(_lib.MrmError) as e:
        ss.MatrixMultEx(D, A, B)
    assert e.value.code == _lib.MRM_BAD_DIMS
    # zero dimension: no-op, no error (src/BlockedMult.pas:374-375)
    Z = np.zeros((0, 4))
    ss.MatrixMultEx(np.zeros((0, 2)), Z, np.ones((4, 2)))
    assert _lib.load().mrm_last_error() == 0


def test_concurrent_callers_disjoint_tiles():
    """Pool threads call the blocked pointer at once on disjoint dest tiles
    (src/ThreadedMatrixOperations.pas:706-717)."""
    M, N, K = 256, 256, 200
    A = orc.pm1_matrix(K, M, 31)
    B = orc.pm1_matrix(N, K, 32)
    D = np.zeros((M, N))
    tiles = [(i, j) for i in range(0, M, 64) for j in range(0, N, 128)]
    errs = []

    def work(t):
        i, j = t
        try:
            ss.MatrixMultEx(D[i:i + 64, j:j + 128], A[i:i + 64], B[:, j:j + 128])
        except Exception as ex:  # pragma: no cover
            errs.append(ex)

    th = [threading.Thread(target=work, args=(t,)) for t in tiles]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs
    assert orc.gemm_err(D, orc.generic_mult(A, B), A, B) <= tol(K)


def test_device_entry_8192_linearity():
    """Full size (8192^3, device-resident, the bench workload): C*x must equal
    A*(B*x) to the normwise bound, and sampled rows must match the oracle."""
    import torch
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(8101)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C = torch.empty(n, n, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.mrm_dev_dgemm(_lib.NN, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, C.data_ptr(), n * 8,
                           _lib.DO_NONE, stream)
    assert rc == 0
    torch.cuda.synchronize()
    An, Bn, Cn = A.cpu().numpy(), B.cpu().numpy(), C.cpu().numpy()
    x = np.random.default_rng(1).standard_normal((n, 1))
    lhs = Cn @ x
    rhs = An @ (Bn @ x)
    scale = np.linalg.norm(An) * np.linalg.norm(Bn) * np.linalg.norm(x)
    assert np.linalg.norm(lhs - rhs) / scale <= tol(n)
    rows = [0, 1, 4095, 8191]
    ref = orc.generic_mult_transp(np.ascontiguousarray(An[rows]), np.ascontiguousarray(Bn.T))
    assert orc.gemm_err(np.ascontiguousarray(Cn[rows]), ref, np.ascontiguousarray(An[rows]), Bn) <= tol(n)


@pytest.mark.parametrize("trans", [1, 2])
def test_device_entry_t1_t2_4096_sampled(trans):
    import torch
    n = 4096
    g = torch.Generator(device="cuda").manual_seed(4101 + trans)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C = torch.empty(n, n, dtype=torch.float64, device="cuda")
    rc = _lib.load().mrm_dev_dgemm(trans, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, C.data_ptr(), n * 8,
                                   _lib.DO_NONE, torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    An, Bn, Cn = A.cpu().numpy(), B.cpu().numpy(), C.cpu().numpy()
    Am = An.T if trans == 1 else An
    Bm = Bn.T if trans == 2 else Bn
    rows = [0, 17, 2048, 4095]
    ref = orc.generic_mult(np.ascontiguousarray(Am[rows]), np.ascontiguousarray(Bm))
    assert orc.gemm_err(np.ascontiguousarray(Cn[rows]), ref, np.ascontiguousarray(Am[rows]), np.ascontiguousarray(Bm)) <= tol(n)


def test_covariance_shape_split_k():
    """X^T X with X tall (the covariance shape of Statistics.Covariance,
    src/Statistics.pas:599-639): few output tiles, long K, so the GEMM splits K
    and sums the chunks in a fixed order.  Checked against the oracle's
    GenericTranspMtxMult on a 20000 x 96 case and, at 200000 x 512, on sampled
    rows plus symmetry."""
    import torch
    lib = _lib.load()
    X = orc.pm1_matrix(96, 20000, 2001)
    C = np.zeros((96, 96))
    ss.MatrixMultT1(C, X, X)
    ref = orc.generic_transp_mult(X, X)
    assert orc.gemm_err(C, ref, np.ascontiguousarray(X.T), X) <= 4 * 20000 * U
    # full size on device
    n, k = 512, 200000
    g = torch.Generator(device="cuda").manual_seed(2001)
    Xd = torch.rand(k, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Cd = torch.empty(n, n, dtype=torch.float64, device="cuda")
    s_ = torch.cuda.current_stream().cuda_stream
    assert lib.mrm_dev_dgemm(_lib.TN, n, n, k, Xd.data_ptr(), n * 8, Xd.data_ptr(), n * 8, Cd.data_ptr(), n * 8, 0,
                             s_) == 0
    torch.cuda.synchronize()
    Cn = Cd.cpu().numpy()
    Xn = Xd.cpu().numpy()
    assert np.max(np.abs(Cn - Cn.T)) <= 4 * k * U * np.max(np.abs(np.diag(Cn)))
    cols = np.array([0, 1, 255, 511])
    ref_rows = orc.generic_transp_mult(np.ascontiguousarray(Xn[:, cols]), Xn)
    err = np.linalg.norm(Cn[cols] - ref_rows) / (np.linalg.norm(Xn[:, cols]) * np.linalg.norm(Xn))
    assert err <= 4 * k * U, err


@pytest.mark.parametrize("trans", [_lib.TN, _lib.NT])
@pytest.mark.parametrize("n,k,op", [(300, 1000, 0), (300, 1000, 2), (200, 50000, 1), (129, 40000, 0)])
def test_symmetric_product_equals_general_path(trans, n, k, op):
    """A^T A (TN) and A A^T (NT) with the same operand for both sides compute
    only the tiles on and above the diagonal and mirror them.  The mirrored
    sums are the same products in the same order, so the result must equal,
    bit for bit, the general path run on a copy of the operand (a different
    pointer), for every destination operation, with and without split K."""
    import torch
    lib = _lib.load()
    g = torch.Generator(device="cuda").manual_seed(n + k + op)
    shape = (k, n) if trans == _lib.TN else (n, k)
    X = torch.rand(*shape, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Y = X.clone()
    C0 = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g)
    C1, C2 = C0.clone(), C0.clone()
    ld = shape[1] * 8
    s_ = torch.cuda.current_stream().cuda_stream
    assert lib.mrm_dev_dgemm(trans, n, n, k, X.data_ptr(), ld, X.data_ptr(), ld, C1.data_ptr(), n * 8, op, s_) == 0
    assert lib.mrm_dev_dgemm(trans, n, n, k, X.data_ptr(), ld, Y.data_ptr(), ld, C2.data_ptr(), n * 8, op, s_) == 0
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    if op == 0:
        assert torch.equal(C1, C1.T)


def test_symmetric_host_entry_stages_once():
    """The host MultT1 / MultT2 entries with the same matrix twice take the
    symmetric device path; the result equals the one for a copy of the
    operand bit for bit."""
    X = orc.pm1_matrix(300, 2000, 77)
    for mult in (ss.MatrixMultT1, ss.MatrixMultT2):
        Y = X if mult is ss.MatrixMultT1 else np.ascontiguousarray(X.T)
        n = Y.shape[1] if mult is ss.MatrixMultT1 else Y.shape[0]
        C1, C2 = np.zeros((n, n)), np.zeros((n, n))
        mult(C1, Y, Y)
        mult(C2, Y, Y.copy())
        assert np.array_equal(C1, C2)
        assert np.array_equal(C1, C1.T)


def test_reference_fixture_products():
    """TestMult / TestMultT1 / TestMultT2 (Test/TestMatrixClass.pas:456-559) on
    the reference's fixtures MatrixData1/2 (read with the RMCLS reader): the
    dimension checks of the reference tests, and the values, which are
    integer-valued and therefore exact."""
    from pathlib import Path
    from mrmath_amd import persistence as ps
    golden = Path(__file__).resolve().parent / "golden"
    m1 = ps.read_matrix(golden / "MatrixData1.txt")
    m2 = ps.read_matrix(golden / "MatrixData2.txt")
    r1 = ss.TDoubleMatrix(m1.Width, m1.Height, m1.SubMatrix())
    r2t = ss.TDoubleMatrix(m2.Height, m2.Width, m2.SubMatrix().T)   # fRefMatrix2.Transpose
    mtx = r1.Mult(r2t)
    assert mtx.Width == mtx.Height == 10
    assert np.array_equal(mtx.data, m1.data @ m2.data.T)
    mtx = r1.MultT1(ss.TDoubleMatrix(m1.Width, m1.Height, m1.SubMatrix()))
    assert mtx.Width == mtx.Height == 5
    assert np.array_equal(mtx.data, m1.data.T @ m1.data)
    mtx = r1.MultT2(ss.TDoubleMatrix(m1.Width, m1.Height, m1.SubMatrix()))
    assert mtx.Width == mtx.Height == 10
    assert np.array_equal(mtx.data, m1.data @ m1.data.T)
