"""Pins the CPU oracle against the reference's golden vectors, known-answer
tests and closed forms (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a swapped index, a transposed operand, a wrong pivot rule) fails it.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import _pins
import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"
KAT = json.loads((GOLDEN / "kat.json").read_text())


def test_mersenne_twister_golden_file():
    """Reproduces Test/marsenneTwistOrig.txt byte for byte, the way
    Test/TestRandom.pas:249-311 writes it (init_by_array(0x123,0x234,0x345,0x456),
    1000 genrand_int32 then 1000 genrand_res53 at %10.8f, CRLF lines)."""
    mt = orc.MTStream(key=[0x123, 0x234, 0x345, 0x456])
    out = ["1000 outputs of genrand_int32()\r\n"]
    for _ in range(200):
        out.append("".join("%10u " % mt.int32() for _ in range(5)) + "\r\n")
    out.append("\r\n1000 outputs of genrand_real2()\r\n")
    for _ in range(200):
        out.append("".join("%10.8f " % mt.res53() for _ in range(5)) + "\r\n")
    assert "".join(out).encode() == (GOLDEN / "marsenneTwistOrig.txt").read_bytes()


def test_mt_init_genrand_known_value():
    # mt19937ar reference: init_genrand(5489) -> first output 3499211612
    assert orc.MTStream(seed=5489).int32() == 3499211612


def test_rand_matrix_seed_chaining():
    """InitRandom refills seed := RandInt after the fill (src/DblMatrix.pas:2751-2759)."""
    a, s1 = orc.rand_matrix(7, 5, 301)
    mt = orc.MTStream(seed=301)
    expect = np.array([[mt.res53() for _ in range(7)] for _ in range(5)])
    assert np.array_equal(a, expect)
    assert s1 == (mt.int32() & 0x7FFFFFFF) % 2147483647
    assert np.all((a >= 0) & (a < 1))
    # padded LineWidth gives the same logical content
    b, s2 = orc.rand_matrix(7, 5, 301, lw_doubles=8)
    assert np.array_equal(a, b) and s1 == s2


def test_gauss_polar_transform():
    """RandGauss returns v2*f then v1*f with f = sqrt(-2 ln r / r) from the MT stream."""
    g = orc.MTStream(seed=77)
    mt = orc.MTStream(seed=77)
    while True:
        v1 = 2 * mt.res53() - 1
        v2 = 2 * mt.res53() - 1
        r = v1 * v1 + v2 * v2
        if r < 1 and r != 0:
            break
    f = np.sqrt(-2 * np.log(r) / r)
    assert g.gauss() == pytest.approx(v2 * f, rel=1e-15)
    assert g.gauss() == pytest.approx(v1 * f, rel=1e-15)


def _flat(vals, rows, lw):
    return np.array(vals, dtype=np.float64)


def test_generic_mult_kat():
    k = KAT["gemm_mult"]
    m1 = np.array(k["mt1_flat"], dtype=np.float64)
    m2 = np.array(k["mt2_flat"], dtype=np.float64)
    c = k["case1"]
    A = np.lib.stride_tricks.as_strided(m1, (c["height1"], c["width1"]), (c["lw1"] * 8, 8))
    B = np.lib.stride_tricks.as_strided(m2, (c["height2"], c["width2"]), (c["lw2"] * 8, 8))
    assert np.array_equal(orc.generic_mult(A, B).ravel(), c["expect"])
    c = k["case2"]
    A = np.lib.stride_tricks.as_strided(m1, (c["height1"], c["width1"]), (c["lw1"] * 8, 8))
    B = np.lib.stride_tricks.as_strided(m2[c["mt2_offset"]:], (c["height2"], c["width2"]), (c["lw2"] * 8, 8))
    assert np.array_equal(orc.generic_mult(A, B).ravel(), c["expect"])


@pytest.mark.parametrize("case", KAT["gemm_closed_form"]["cases"])
def test_generic_mult_closed_form(case):
    A = np.ones((case["h"], case["k"]))
    B = 2 * np.ones((case["k"], 1))
    C = orc.generic_mult(A, B)
    assert C.shape == (case["h"], 1)
    assert np.allclose(C, case["expect"], rtol=KAT["gemm_closed_form"]["rel_tol"], atol=0)


def test_generic_mult_transposes_integer_exact():
    """Integer-valued inputs make every product exact: A*B, A*B^T, A^T*B must equal
    numpy bit for bit (pins operand orientation of all three loops)."""
    rng = np.random.default_rng(3)
    A = rng.integers(-50, 50, (37, 23)).astype(np.float64)
    B = rng.integers(-50, 50, (23, 29)).astype(np.float64)
    Bt = np.ascontiguousarray(B.T)
    At = np.ascontiguousarray(A.T)
    assert np.array_equal(orc.generic_mult(A, B), A @ B)
    assert np.array_equal(orc.generic_mult_transp(A, Bt), A @ B)
    assert np.array_equal(orc.generic_transp_mult(At, B), A @ B)


def test_generic_mult_sequential_order():
    """GenericMtxMult sums k in order without FMA: one row times one column must
    equal a left-to-right Python float sum (catches reordered or fused sums)."""
    rng = np.random.default_rng(11)
    a = rng.standard_normal((1, 301))
    b = rng.standard_normal((301, 1))
    s = 0.0
    for i in range(301):
        s = s + float(a[0, i]) * float(b[i, 0])
    assert orc.generic_mult(a, b)[0, 0] == s


@pytest.mark.parametrize("trans", [0, 1, 2])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_blocked_mult_ops(trans, op):
    """dest (op)= product with ragged tiles (block 16 over 50x37x41)."""
    rng = np.random.default_rng(5 + trans * 3 + op)
    M, N, K = 50, 37, 41
    Am = rng.integers(-9, 9, (M, K)).astype(np.float64)
    Bm = rng.integers(-9, 9, (K, N)).astype(np.float64)
    A = Am if trans != 1 else np.ascontiguousarray(Am.T)
    B = Bm if trans != 2 else np.ascontiguousarray(Bm.T)
    D0 = rng.integers(-9, 9, (M, N)).astype(np.float64)
    D = D0.copy()
    orc.blocked_mult(trans, D, A, B, 16, op)
    P = Am @ Bm
    expect = P if op == 0 else (D0 + P if op == 1 else D0 - P)
    assert np.array_equal(D, expect)


def test_fast_mult_matches_generic():
    A, s = orc.rand_matrix(300, 140, 301)
    B, _ = orc.rand_matrix(170, 300, s)
    C = np.zeros((140, 170))
    used = orc.fast_mult(C, A, B, threads=4, block=128)
    assert used >= 1
    assert orc.gemm_err(C, orc.generic_mult(A, B), A, B) < 4 * 300 * 2.0 ** -53


def test_num_cores_heuristic():
    # src/ThreadedMatrixOperations.pas:486-503: 1024^3 with bs 128 -> 8 + 8 = 16
    assert orc.lib().orc_num_cores_for_mult(1024, 1024, 1024, 128, 64) == 16
    assert orc.lib().orc_num_cores_for_mult(100, 1000, 100, 128, 64) == 8
    assert orc.lib().orc_num_cores_for_mult(8192, 8192, 8192, 128, 12) == 12


# ---- LU / solve ---------------------------------------------------------------

def test_lu_kat():
    k = KAT["lu"]
    A = np.array(k["A"], dtype=np.float64).reshape(2, 2)
    rc, indx = orc.lu_decomp_inplace(A)
    assert rc == 0
    assert np.array_equal(A.ravel(), k["LU"])
    assert list(indx) == k["indx"]
    b = np.array(k["b"], dtype=np.float64).reshape(2, 1)
    orc.lu_backsubst(A, indx, b)
    assert np.array_equal(b.ravel(), k["x"])


def test_lu_kat_strided():
    s = KAT["lu"]["strided"]
    raw = np.array(s["A_lw3"], dtype=np.float64)
    A = np.lib.stride_tricks.as_strided(raw, (2, 2), (24, 8))
    rc, indx = orc.lu_decomp_inplace(A)
    assert rc == 0 and np.array_equal(raw, s["LU_lw3"])
    Braw = np.array(s["B_lw2"], dtype=np.float64)
    B = Braw.reshape(2, 2)
    orc.lu_backsubst(A, indx, B, col=0)
    assert np.array_equal(Braw, s["X_lw2"])


def test_det_and_inverse_kat():
    A = np.array(KAT["det"]["A"], dtype=np.float64).reshape(2, 2)
    assert orc.determinant(A) == KAT["det"]["det"]
    inv = np.array(KAT["inverse"]["A"], dtype=np.float64).reshape(2, 2)
    assert orc.inverse_inplace(inv) == 0
    assert np.array_equal(inv.ravel(), KAT["inverse"]["inv"])


def test_singular_detected():
    A = np.array(KAT["singular"]["A"], dtype=np.float64).reshape(3, 3)
    assert orc.inverse_inplace(A.copy()) == 1
    assert orc.lu_decomp_inplace(A.copy())[0] == 1
    assert orc.determinant(A) == 0.0


def _perm_from_indx(indx):
    p = np.arange(len(indx))
    for i, r in enumerate(indx):
        p[[i, r]] = p[[r, i]]
    return p


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 33, 64, 97])
def test_lu_reconstructs_and_pivots_partially(n):
    """P*A = L*U, |L| <= 1 (partial pivoting), pivot choice = first max of the
    remaining column (checked by replaying textbook GEPP in numpy)."""
    rng = np.random.default_rng(n)
    A0 = rng.standard_normal((n, n))
    A = A0.copy()
    rc, indx = orc.lu_decomp_inplace(A)
    assert rc == 0
    L = np.tril(A, -1) + np.eye(n)
    U = np.triu(A)
    p = _perm_from_indx(indx)
    assert np.allclose(A0[p], L @ U, atol=1e-12 * n)
    assert np.all(np.abs(L) <= 1 + 1e-15)
    # textbook GEPP with first-max tie rule
    W = A0.copy()
    piv = []
    for k in range(n):
        r = k + int(np.argmax(np.abs(W[k:, k])))
        piv.append(r)
        W[[k, r]] = W[[r, k]]
        W[k + 1:, k] /= W[k, k]
        W[k + 1:, k + 1:] -= np.outer(W[k + 1:, k], W[k, k + 1:])
    assert list(indx) == piv
    assert np.allclose(A, W, atol=1e-11)


@pytest.mark.parametrize("n,nrhs", [(5, 1), (48, 3), (120, 4)])
def test_lineq_solve_matches_numpy(n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    rc, X = orc.lineq_solve(A, B)
    assert rc == 0
    assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-10, atol=1e-12)
    assert orc.backward_error(A, X, B) < 1e-15


@pytest.fixture(scope="module")
def wilkinson_system():
    """n = 50 Wilkinson matrix, 3 right-hand sides, and the exact solution."""
    n = 50
    A = _pins.wilkinson(n)
    B = orc.pm1_matrix(3, n, 7)
    Xe = np.stack([_pins.exact_solve(A, B[:, j]) for j in range(3)], axis=1)
    return A, B, Xe


def test_lineq_solve_refinement_pinned_by_the_exact_solution(wilkinson_system):
    """MatrixLinEQSolve's refinement loop (src/LinAlgLU.pas:542-586) against
    the exact rational solution: without refinement the pivot growth 2^49
    leaves a forward error ~1e-3; one step r = A x - b, x -= (LU)^-1 r
    restores full accuracy, and further steps keep it.  A loop that does
    nothing, adds the correction, or reuses the first residual fails."""
    A, B, Xe = wilkinson_system
    errs = [_pins.forward_error(orc.lineq_solve(A, B, k)[1], Xe) for k in range(4)]
    assert errs[0] > 1e-5
    assert all(e <= 2e-15 for e in errs[1:]), errs
    # NumRefinments 0 and 1 differ by more than the remaining error
    _, X0 = orc.lineq_solve(A, B, 0)
    _, X1 = orc.lineq_solve(A, B, 1)
    assert np.abs(X1 - X0).max() > 1e3 * np.abs(X1 - Xe).max()
    # backward error: large before, at the FP64 level after
    assert orc.backward_error(A, X0, B) > 1e-7
    assert orc.backward_error(A, X1, B) < 1e-16


def test_refinement_pin_rejects_planted_mistakes(wilkinson_system):
    """The pin above fails a wrong-sign correction and a stale residual
    (both planted here, on the CPU, from the oracle's own LU and residual)."""
    A, B, Xe = wilkinson_system
    n = A.shape[0]
    LU = A.copy()
    rc, indx = orc.lu_decomp_inplace(LU)
    assert rc == 0
    _, X0 = orc.lineq_solve(A, B, 0)
    R = A @ X0 - B
    D = R.copy()
    for j in range(B.shape[1]):
        orc.lu_backsubst(LU, indx, D, j)
    assert _pins.forward_error(X0 - D, Xe) <= 2e-15      # the correct step
    assert _pins.forward_error(X0 + D, Xe) > 1e-5        # wrong sign
    assert _pins.forward_error(X0 - 2 * D, Xe) > 1e-5    # two steps on the first residual
    assert n == 50


def test_lineq_solve_refinement_per_column(wilkinson_system):
    """Every right-hand side is refined on its own (the reference loops over
    the columns): a 3-column solve equals three 1-column solves."""
    A, B, _ = wilkinson_system
    _, X = orc.lineq_solve(A, B, 2)
    for j in range(B.shape[1]):
        _, xj = orc.lineq_solve(A, np.ascontiguousarray(B[:, j:j + 1]), 2)
        assert np.array_equal(X[:, j:j + 1], xj)


def test_determinant_and_inverse_random():
    rng = np.random.default_rng(21)
    A = rng.standard_normal((31, 31))
    assert orc.determinant(A) == pytest.approx(np.linalg.det(A), rel=1e-10)
    I = A.copy()
    assert orc.inverse_inplace(I) == 0
    assert np.allclose(I, np.linalg.inv(A), rtol=1e-9, atol=1e-11)


def test_backward_error_metric():
    A = np.eye(3)
    B = np.ones((3, 1))
    X = np.ones((3, 1))
    assert orc.backward_error(A, X, B) == 0.0
    X[1, 0] = 1.5  # r = 0.5, ||A||=1, ||x||=1.5, ||b||=1
    assert orc.backward_error(A, X, B) == pytest.approx(0.5 / 2.5)


def test_covariance_matches_the_textbook_definition():
    """Covariance (src/Statistics.pas:599-639): unbiased = numpy.cov, biased
    = numpy.cov(bias=True), both to rounding."""
    X = orc.pm1_matrix(9, 200, 31)
    C = orc.covariance(X, True)
    assert np.abs(C - np.cov(X, rowvar=False)).max() < 1e-15
    Cb = orc.covariance(X, False)
    assert np.abs(Cb - np.cov(X, rowvar=False, bias=True)).max() < 1e-15
    assert np.array_equal(C, C.T)


def test_covariance_closed_forms():
    """Columns that are exact affine functions of one another: a column plus
    a constant has the same (co)variance (the mean is removed), 2x a column
    has 4x the variance and 2x the covariance; integer data gives exact values."""
    t = np.arange(10.0)
    X = np.stack([t, t + 100.0, 2 * t, -t], axis=1)
    C = orc.covariance(X, True)
    v = float(np.sum((t - t.mean()) ** 2) / 9)  # 9.1666...
    want = v * np.array([[1, 1, 2, -1], [1, 1, 2, -1], [2, 2, 4, -2], [-1, -1, -2, 1]])
    assert np.allclose(C, want, rtol=1e-15, atol=0)


def test_covariance_single_row_or_column_is_the_variance():
    """height = 1 or width = 1 gives MatrixVar of the vector (src/Statistics.pas:606-610)."""
    x = orc.pm1_matrix(1, 17, 3)
    assert orc.covariance(x, True)[0, 0] == pytest.approx(np.var(x, ddof=1), rel=1e-14)
    assert orc.covariance(x.T.copy(), False)[0, 0] == pytest.approx(np.var(x), rel=1e-14)


@pytest.mark.parametrize("n,threads", [(200, 1), (500, 4), (700, 8)])
def test_threaded_lu_matches_the_recursive_lu(n, threads):
    """The restated ThrMatrixLUDecomp (src/LinAlgLU.pas:674-803) is the
    recursive LU with threaded parts: same pivots as the oracle's LU, factors
    equal to the rounding of the different multiply kernel."""
    A = orc.pm1_matrix(n, n, 300 + n)
    A1, A2 = A.copy(), A.copy()
    rc1, i1 = orc.lu_decomp_inplace(A1)
    rc2, i2 = orc.thr_lu_decomp(A2, threads)
    assert rc1 == rc2 == 0
    assert np.array_equal(i1, i2)
    assert np.abs(A1 - A2).max() <= 64 * n * 2.0 ** -53 * np.abs(A).max()


def test_threaded_lineq_solve_and_refinement(wilkinson_system):
    """ThrMatrixLinEQSolve (src/LinAlgLU.pas:986-1128): backward-stable on a
    random system with threads, and its refinement passes the exact-solution
    pin of the Wilkinson system."""
    n = 600
    A = orc.pm1_matrix(n, n, 61)
    B = orc.pm1_matrix(7, n, 62)
    rc, X = orc.thr_lineq_solve(A, B, 0, 6)
    assert rc == 0 and orc.backward_error(A, X, B) <= 1e-14
    Aw, Bw, Xe = wilkinson_system
    errs = [_pins.forward_error(orc.thr_lineq_solve(Aw, Bw, k, 3)[1], Xe) for k in range(3)]
    assert errs[0] > 1e-5 and errs[1] <= 2e-15 and errs[2] <= 2e-15
    S = np.array([[1.0, 2.0], [2.0, 4.0]])
    assert orc.thr_lineq_solve(S, np.ones((2, 1)), 0, 2)[0] == 1  # leSingular
