"""GPU parity of the INT8- and BF16-slice FP64 GEMM (mrm_dev_dgemm_ozaki[_ex] /
mrm_mult_ozaki).

Bars: bit-exact against the oracle's restatement of the same method
(oracle.ozaki_gemm) -- every step before the FP64 accumulation is exact and the
accumulation order is fixed; and against the FP64 oracle GEMM, the error per
slice count is reported and, from 6 slices on, within the GEMM bar
||C - Cref||_F / (||A||_F ||B||_F) <= 4 k 2^-53 on the wide-range inputs
(2 res53 - 1) 2^e, e in [-20, 20].
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import stubswitch as ss

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _operands(trans, A, B):
    """Stored operands for the GPU given the logical A (M x K), B (K x N)."""
    a = np.ascontiguousarray(A.T) if trans == _lib.TN else A
    b = np.ascontiguousarray(B.T) if trans == _lib.NT else B
    return a, b


@pytest.mark.parametrize("trans", [_lib.NN, _lib.TN, _lib.NT])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (5, 7, 3), (128, 128, 64), (130, 257, 100), (256, 192, 320)])
@pytest.mark.parametrize("s", [1, 3, 7])
def test_bit_exact_vs_oracle(trans, M, N, K, s):
    A = orc.pm1_wide_matrix(K, M, 100 + M + s)
    B = orc.pm1_wide_matrix(N, K, 200 + N + s)
    a, b = _operands(trans, A, B)
    Cg = np.full((M, N), np.nan)
    ss.MatrixMultOzaki(Cg, a, b, slices=s, trans=trans)
    Co = orc.ozaki_gemm(A, B, s)
    assert np.array_equal(Cg, Co)


@pytest.mark.parametrize("s", range(1, 9))
@pytest.mark.parametrize("op", [ss.doNone, ss.doSub])
def test_every_slice_count_bit_exact(s, op):
    """Every slice count runs its own launches (groups of up to four slice
    diagonals per launch); each must equal the oracle, with none and with the
    signed accumulation of doSub across the groups."""
    M, N, K = 200, 150, 300
    A = orc.pm1_wide_matrix(K, M, 300 + s)
    B = orc.pm1_wide_matrix(N, K, 400 + s)
    C0 = orc.pm1_matrix(N, M, 500 + s)
    Cg = C0.copy()
    ss.MatrixMultOzaki(Cg, A, B, slices=s, op=op)
    Co = orc.ozaki_gemm(A, B, s, C0.copy(), op=op)
    assert np.array_equal(Cg, Co)


@pytest.mark.parametrize("op", [ss.doAdd, ss.doSub])
def test_ops_bit_exact(op):
    M, N, K = 70, 90, 130
    A = orc.pm1_matrix(K, M, 1)
    B = orc.pm1_matrix(N, K, 2)
    C0 = orc.pm1_matrix(N, M, 3)
    Cg = C0.copy()
    ss.MatrixMultOzaki(Cg, A, B, slices=5, op=op)
    Co = orc.ozaki_gemm(A, B, 5, C0.copy(), op=op)
    assert np.array_equal(Cg, Co)


def test_padded_views_and_integer_exactness():
    rng = np.random.default_rng(9)
    rawA = np.full((60, 80), 7.0)
    rawB = np.full((50, 45), 7.0)
    rawC = np.full((60, 45), 3.0)
    A, B, Cv = rawA[:, :50], rawB[:, :40], rawC[:, :40]
    A[...] = rng.integers(-127, 128, A.shape)
    B[...] = rng.integers(-(2 ** 20), 2 ** 20, B.shape)
    ss.MatrixMultOzaki(Cv, A, B, slices=3)
    assert np.array_equal(Cv, orc.generic_mult(np.ascontiguousarray(A), np.ascontiguousarray(B)))
    assert np.all(rawC[:, 40:] == 3.0)


def test_k_chunks():
    """K above one int32 chunk (16384) takes two chunks; still bit-exact."""
    M, N, K = 64, 48, 16384 + 700
    A = orc.pm1_wide_matrix(K, M, 41)
    B = orc.pm1_wide_matrix(N, K, 42)
    Cg = np.zeros((M, N))
    ss.MatrixMultOzaki(Cg, A, B, slices=4)
    assert np.array_equal(Cg, orc.ozaki_gemm(A, B, 4))


def test_error_per_slice_count():
    """Wide-range inputs at 512^3: error against the FP64 oracle for s = 3..7."""
    n = 512
    A = orc.pm1_wide_matrix(n, n, 3101)
    B = orc.pm1_wide_matrix(n, n, 3102)
    ref = orc.generic_mult(A, B)
    errs = {}
    for s in range(3, 8):
        Cg = np.zeros((n, n))
        ss.MatrixMultOzaki(Cg, A, B, slices=s)
        errs[s] = orc.gemm_err(Cg, ref, A, B)
        if s == 7:
            assert np.array_equal(Cg, orc.ozaki_gemm(A, B, 7))
    print("normwise error per slice count:", {k: f"{v:.3e}" for k, v in errs.items()})
    for s in range(3, 7):
        assert errs[s + 1] < errs[s] / 32  # about 2^-7 per slice
    assert errs[7] <= 4 * n * U


def test_errors_and_empty():
    A = np.ones((4, 5))
    B = np.ones((6, 3))
    with pytest.raises(_lib.MrmError):
        ss.MatrixMultOzaki(np.zeros((4, 3)), A, B, slices=3)
    with pytest.raises(_lib.MrmError):
        ss.MatrixMultOzaki(np.zeros((4, 3)), A, np.ones((5, 3)), slices=0)
    with pytest.raises(_lib.MrmError):
        ss.MatrixMultOzaki(np.zeros((4, 3)), A, np.ones((5, 3)), slices=9)
    ss.MatrixMultOzaki(np.zeros((0, 3)), np.zeros((0, 5)), np.ones((5, 3)), slices=3)
    lib = _lib.load()
    assert lib.mrm_ozaki_workspace_bytes(8192, 8192, 8192, 7) == 2 * 8192 * 4 + 7 * 2 * 8192 * 8192
    assert lib.mrm_ozaki_workspace_bytes(10, 10, 10, 0) == 0
    # too small a workspace is refused before anything is enqueued
    rc = lib.mrm_dev_dgemm_ozaki(_lib.NN, 128, 128, 128, 1, 1024, 1, 1024, 1, 1024, 0, 3, 1, 100, None)
    assert rc == _lib.MRM_BAD_ARG


@pytest.mark.parametrize("trans", [_lib.NN, _lib.TN, _lib.NT])
def test_device_8192_error_per_slice(trans):
    """Full size on device tensors: at s = 6 (the FP64-accurate operating
    point of the bench) and s = 7 the emulated product agrees with the native
    FP64 MFMA GEMM (itself checked against the oracle) to the GEMM bar, and
    sampled rows agree with the oracle GEMM directly."""
    import torch
    lib = _lib.load()
    n = 8192 if trans == _lib.NN else 4096
    g = torch.Generator(device="cuda").manual_seed(trans)
    e1 = torch.randint(-20, 21, (n, n), device="cuda", generator=g).double()
    e2 = torch.randint(-20, 21, (n, n), device="cuda", generator=g).double()
    A = (torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * torch.exp2(e1)
    B = (torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * torch.exp2(e2)
    del e1, e2
    s_ = torch.cuda.current_stream().cuda_stream
    Cd = torch.empty(n, n, dtype=torch.float64, device="cuda")
    assert lib.mrm_dev_dgemm(trans, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, Cd.data_ptr(), n * 8, 0, s_) == 0
    wb = lib.mrm_ozaki_workspace_bytes(n, n, n, 7)
    W = torch.empty(wb, dtype=torch.uint8, device="cuda")
    An = (A.t() if trans == _lib.TN else A).contiguous().cpu().numpy()
    Bn = (B.t() if trans == _lib.NT else B).contiguous().cpu().numpy()
    na = float(np.linalg.norm(An)) * float(np.linalg.norm(Bn))
    rows = np.arange(0, n, n // 8)
    ref_rows = orc.generic_mult(np.ascontiguousarray(An[rows]), Bn)
    for s in (6, 7):
        Co = torch.empty_like(Cd)
        assert lib.mrm_dev_dgemm_ozaki(trans, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, Co.data_ptr(), n * 8,
                                       0, s, W.data_ptr(), wb, s_) == 0
        torch.cuda.synchronize()
        err = float(torch.linalg.norm(Co - Cd)) / na
        Cr = Co.cpu().numpy()[rows]
        err_rows = float(np.linalg.norm(Cr - ref_rows)) / (float(np.linalg.norm(An[rows])) * float(np.linalg.norm(Bn)))
        print(f"trans={trans} n={n} s={s} normwise error vs FP64 GEMM {err:.3e}, sampled rows vs oracle "
              f"{err_rows:.3e}, bar {4 * n * U:.3e}")
        assert err <= 4 * n * U
        assert err_rows <= 4 * n * U


def _dev_ozaki(A, B, s, kind, trans=_lib.NN, C0=None, op=0):
    """C (op)= A B through mrm_dev_dgemm_ozaki_ex on device copies of the
    logical A (M x K) and B (K x N); returns the result on the host."""
    import torch
    lib = _lib.load()
    M, K = A.shape
    N = B.shape[1]
    a, b = _operands(trans, A, B)
    Ad = torch.from_numpy(a).cuda()
    Bd = torch.from_numpy(b).cuda()
    Cd = torch.from_numpy(C0.copy() if C0 is not None else np.zeros((M, N))).cuda()
    wb = lib.mrm_ozaki_workspace_bytes_ex(M, N, K, s, kind)
    W = torch.empty(max(wb, 1), dtype=torch.uint8, device="cuda")
    rc = lib.mrm_dev_dgemm_ozaki_ex(trans, M, N, K, Ad.data_ptr(), a.shape[1] * 8, Bd.data_ptr(), b.shape[1] * 8,
                                    Cd.data_ptr(), N * 8, op, s, kind, W.data_ptr(), wb,
                                    torch.cuda.current_stream().cuda_stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return Cd.cpu().numpy()


@pytest.mark.parametrize("trans", [_lib.NN, _lib.TN, _lib.NT])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (130, 257, 100), (256, 192, 320)])
def test_bf16_slices_bit_exact_vs_oracle(trans, M, N, K):
    """BF16 digits give the INT8 digits' exact diagonal sums, so the result
    equals the oracle's bit for bit."""
    A = orc.pm1_wide_matrix(K, M, 700 + M)
    B = orc.pm1_wide_matrix(N, K, 800 + N)
    for s in (1, 4, 6):
        assert np.array_equal(_dev_ozaki(A, B, s, _lib.OZAKI_BF16, trans), orc.ozaki_gemm(A, B, s))


@pytest.mark.parametrize("s", range(1, 9))
@pytest.mark.parametrize("op", [ss.doNone, ss.doSub])
def test_bf16_every_slice_count_bit_exact(s, op):
    """Every BF16 group table row (up to five launches of one or two
    diagonals), with none and with the signed accumulation of doSub."""
    M, N, K = 200, 150, 300
    A = orc.pm1_wide_matrix(K, M, 900 + s)
    B = orc.pm1_wide_matrix(N, K, 1000 + s)
    C0 = orc.pm1_matrix(N, M, 1100 + s)
    Cg = _dev_ozaki(A, B, s, _lib.OZAKI_BF16, C0=C0, op=op)
    assert np.array_equal(Cg, orc.ozaki_gemm(A, B, s, C0.copy(), op=op))


def test_bf16_fp32_sums_at_their_bound():
    """All digits 127 (entries just below the row / column scale): every FP32
    partial sum of 8 K steps reaches (d+1) 128 127^2, up to 8 128 127^2 =
    16516096 for the last diagonal at s = 8, just under 2^24 = 16777216; the
    sums must still be exact (result equal to the INT8 slices' and the
    oracle's), also over K beyond one int32 chunk (two chunks of 16384)."""
    x = 1.0 - 2.0 ** -60
    for M, N, K in ((64, 96, 1024), (40, 33, 16384 + 300)):
        A = np.full((M, K), x)
        A[1::2] *= -1.0
        B = np.full((K, N), x)
        B[:, ::3] *= -1.0
        Cb = _dev_ozaki(A, B, 8, _lib.OZAKI_BF16)
        Ci = _dev_ozaki(A, B, 8, _lib.OZAKI_INT8)
        assert np.array_equal(Cb, Ci)
        assert np.array_equal(Cb, orc.ozaki_gemm(A, B, 8))


def test_bf16_device_2048_equals_int8():
    """Device tensors at 2048^3 on the wide-range inputs: BF16 and INT8 slices
    agree bit for bit at s = 6, and the result meets the GEMM bar."""
    import torch
    lib = _lib.load()
    n = 2048
    g = torch.Generator(device="cuda").manual_seed(77)
    e1 = torch.randint(-20, 21, (n, n), device="cuda", generator=g).double()
    A = (torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * torch.exp2(e1)
    B = (torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * torch.exp2(e1.t().contiguous())
    s_ = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind in (_lib.OZAKI_INT8, _lib.OZAKI_BF16):
        wb = lib.mrm_ozaki_workspace_bytes_ex(n, n, n, 6, kind)
        W = torch.empty(wb, dtype=torch.uint8, device="cuda")
        Co = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
        assert lib.mrm_dev_dgemm_ozaki_ex(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, Co.data_ptr(), n * 8,
                                          0, 6, kind, W.data_ptr(), wb, s_) == 0
        out[kind] = Co
    torch.cuda.synchronize()
    assert torch.equal(out[_lib.OZAKI_INT8], out[_lib.OZAKI_BF16])
    Cd = torch.empty(n, n, dtype=torch.float64, device="cuda")
    assert lib.mrm_dev_dgemm(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, Cd.data_ptr(), n * 8, 0, s_) == 0
    torch.cuda.synchronize()
    err = float(torch.linalg.norm(out[_lib.OZAKI_BF16] - Cd)) / (float(torch.linalg.norm(A)) * float(torch.linalg.norm(B)))
    assert err <= 4 * n * U


def test_bf16_kind_and_workspace_checks():
    lib = _lib.load()
    assert lib.mrm_ozaki_workspace_bytes_ex(8192, 8192, 8192, 7, _lib.OZAKI_BF16) == \
        2 * 8192 * 4 + 2 * 7 * 2 * 8192 * 8192
    assert lib.mrm_ozaki_workspace_bytes_ex(8192, 8192, 8192, 7, _lib.OZAKI_INT8) == \
        lib.mrm_ozaki_workspace_bytes(8192, 8192, 8192, 7)
    assert lib.mrm_ozaki_workspace_bytes_ex(10, 10, 10, 3, 2) == 0
    rc = lib.mrm_dev_dgemm_ozaki_ex(_lib.NN, 128, 128, 128, 1, 1024, 1, 1024, 1, 1024, 0, 3, 2, 1, 1 << 30, None)
    assert rc == _lib.MRM_BAD_ARG
    # the INT8-sized workspace is too small for BF16 slices
    wb = lib.mrm_ozaki_workspace_bytes_ex(128, 128, 128, 3, _lib.OZAKI_INT8)
    rc = lib.mrm_dev_dgemm_ozaki_ex(_lib.NN, 128, 128, 128, 1, 1024, 1, 1024, 1, 1024, 0, 3, _lib.OZAKI_BF16, 1, wb,
                                    None)
    assert rc == _lib.MRM_BAD_ARG
