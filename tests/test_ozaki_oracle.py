"""Pins of the INT8-slice GEMM restatement (oracle.ozaki_gemm) on the CPU.

The reference multiplies in FP64 only, so the split-precision oracle is pinned
by what the mathematics fixes: exact products where the slices hold every
digit, the a-priori error bound, the 2^-7 decay per slice, and agreement with
the FP64 oracle GEMM at full slice count.
"""
import numpy as np
import pytest

import oracle as orc


def _bound(A, B, s):
    """4 (s+1) K 2^(-7 s) max|A_i.| max|B_.j| plus FP64 rounding of s terms."""
    K = A.shape[1]
    ra = np.abs(A).max(axis=1)[:, None]
    cb = np.abs(B).max(axis=0)[None, :]
    return 4 * (s + 1) * K * 2.0 ** (-7 * s) * ra * cb + 2 * s * 2.0 ** -53 * (np.abs(A) @ np.abs(B))


@pytest.mark.parametrize("s", [1, 2, 3, 5, 8])
def test_exact_when_slices_hold_every_digit(s):
    """A's rows fit in the first 7-bit digit, B's columns in s digits: every
    nonzero slice pair has t + u <= s - 1, so nothing is dropped and the
    result is the exact integer product."""
    rng = np.random.default_rng(s)
    M, K, N = 9, 33, 11
    A = rng.integers(-127, 128, (M, K)).astype(np.float64)
    B = rng.integers(-(2 ** min(7 * s - 1, 40)), 2 ** min(7 * s - 1, 40), (K, N)).astype(np.float64)
    C = orc.ozaki_gemm(A, B, s)
    exact = orc.generic_mult(A, B)
    assert np.array_equal(C, exact)
    assert np.array_equal(C, (A.astype(object) @ B.astype(object)).astype(np.float64))


def test_dropped_pairs_change_the_result():
    """With both operands needing two digits, s = 1 must lose the cross terms
    (a plausible bug that keeps them, or scales by the wrong power, fails)."""
    A = np.array([[127.0 + 1.0 / 128]])
    B = np.array([[127.0 + 1.0 / 128]])
    exact = A[0, 0] * B[0, 0]
    c1 = orc.ozaki_gemm(A, B, 1)[0, 0]
    c2 = orc.ozaki_gemm(A, B, 2)[0, 0]
    c3 = orc.ozaki_gemm(A, B, 3)[0, 0]
    assert c1 == 127.0 * 127.0
    assert c2 == 127.0 * 127.0 + 2 * 127.0 / 128
    assert c3 == exact


@pytest.mark.parametrize("s", [2, 3, 4, 5, 6, 7, 8])
def test_error_bound_uniform_and_wide(s):
    for seed, wide in [(11, False), (12, True)]:
        M, K, N = 40, 300, 37
        if wide:
            A = orc.pm1_wide_matrix(K, M, seed)
            B = orc.pm1_wide_matrix(N, K, seed + 100)
        else:
            A = orc.pm1_matrix(K, M, seed)
            B = orc.pm1_matrix(N, K, seed + 100)
        C = orc.ozaki_gemm(A, B, s)
        exact = (A.astype(np.longdouble) @ B.astype(np.longdouble))
        err = np.abs(C - exact).astype(np.float64)
        assert np.all(err <= _bound(A, B, s)), (s, wide, float(np.max(err / _bound(A, B, s))))


def test_error_decays_by_2_pow_7_per_slice():
    A = orc.pm1_matrix(256, 64, 21)
    B = orc.pm1_matrix(64, 256, 22)
    exact = A.astype(np.longdouble) @ B.astype(np.longdouble)
    errs = [float(np.abs(orc.ozaki_gemm(A, B, s) - exact).max()) for s in (2, 3, 4, 5)]
    for a, b in zip(errs, errs[1:]):
        assert 16 < a / b < 1024, errs


def test_full_slices_match_fp64_gemm():
    """At 8 slices (56 bits) the result agrees with the FP64 oracle GEMM to
    the normwise bar used for the GPU GEMM, 4 K 2^-53."""
    A = orc.pm1_wide_matrix(200, 60, 31)
    B = orc.pm1_wide_matrix(50, 200, 32)
    C = orc.ozaki_gemm(A, B, 8)
    ref = orc.generic_mult(A, B)
    assert orc.gemm_err(C, ref, A, B) <= 4 * 200 * 2.0 ** -53


def test_k_chunks_sum_the_same_exact_product():
    rng = np.random.default_rng(4)
    A = rng.integers(-127, 128, (6, 50)).astype(np.float64)
    B = rng.integers(-127, 128, (50, 5)).astype(np.float64)
    ref = orc.generic_mult(A, B)
    for kc in (1, 7, 16, 50, 1000):
        assert np.array_equal(orc.ozaki_gemm(A, B, 1, kchunk=kc), ref)


def test_ops_and_zero_rows():
    rng = np.random.default_rng(8)
    A = rng.integers(-127, 128, (5, 8)).astype(np.float64)
    A[2] = 0.0
    B = rng.integers(-127, 128, (8, 4)).astype(np.float64)
    C0 = rng.integers(-50, 50, (5, 4)).astype(np.float64)
    P = orc.generic_mult(A, B)
    assert np.array_equal(orc.ozaki_gemm(A, B, 1, C0.copy(), op=1), C0 + P)
    assert np.array_equal(orc.ozaki_gemm(A, B, 1, C0.copy(), op=2), C0 - P)
    assert np.all(orc.ozaki_gemm(A, B, 3)[2] == 0.0)
