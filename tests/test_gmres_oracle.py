"""Pins of the oracle's restatement of the mixed-precision solve
(oracle/ir_gmres.py): the low-precision roundings, the factorisation, GMRES
and the refinement driver, each against what the mathematics fixes, plus
planted mistakes that the pins must catch.  CPU only.
"""
import numpy as np
import pytest

import oracle as orc
from oracle import ir_gmres as ig


def _system(n, kappa, nrhs, seed):
    rng = np.random.default_rng(seed)
    Uq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Vq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(kappa), n)
    return np.ascontiguousarray((Uq * s) @ Vq.T), rng.standard_normal((n, nrhs))


def test_bf16_round_to_nearest_even():
    """Halfway cases go to the even neighbour, others to the nearer one; the
    result has 8 significant bits; brute force over random FP32 values."""
    one = 1.0
    assert ig.to_bf16(np.array([one + 2.0 ** -8]))[0] == 1.0                 # tie, even below
    assert ig.to_bf16(np.array([one + 3 * 2.0 ** -8]))[0] == one + 2 ** -6    # tie, even above
    assert ig.to_bf16(np.array([one + 2.0 ** -8 + 2.0 ** -20]))[0] == one + 2 ** -7
    assert ig.to_bf16(np.array([-(one + 2.0 ** -8)]))[0] == -1.0
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(4000) * 2.0 ** rng.integers(-30, 30, 4000)).astype(np.float32)
    r = ig.to_bf16(x.astype(np.float64))
    u = x.view(np.uint32).astype(np.uint64)
    lo = ((u >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)
    hi = (((u >> 16) + 1) << 16).astype(np.uint32).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - lo), np.abs(hi - xd)
    even_lo = ((u >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(even_lo, lo, hi)))
    assert np.array_equal(r.astype(np.float64), want)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


@pytest.mark.parametrize("mode", [ig.SOLVE_FP32_IR, ig.SOLVE_BF16_IR, ig.SOLVE_BF16X3_IR])
def test_lowp_product_exact_on_small_integers_and_rounding_bound(mode):
    """Integers in [-8, 8] are exact in every mode and their sums stay far
    below 2^24: the product is exact.  Otherwise each entry of the product is
    within (2u + K 2^-24) (|L||U|) of the FP64 one, u the operand rounding."""
    rng = np.random.default_rng(mode)
    L = rng.integers(-8, 9, (40, 70)).astype(np.float64)
    U = rng.integers(-8, 9, (70, 30)).astype(np.float64)
    assert np.array_equal(ig.lowp_product(L, U, mode), L @ U)
    L = rng.standard_normal((50, 300))
    U = rng.standard_normal((300, 40))
    u = {ig.SOLVE_FP32_IR: 2.0 ** -24, ig.SOLVE_BF16_IR: 2.0 ** -8, ig.SOLVE_BF16X3_IR: 2.0 ** -16}[mode]
    k = 300 * (3 if mode == ig.SOLVE_BF16X3_IR else 1)
    bound = (2 * u + k * 2.0 ** -24) * (np.abs(L) @ np.abs(U))
    err = np.abs(ig.lowp_product(L, U, mode) - L @ U)
    assert np.all(err <= bound)
    # and the rounding is really there (not an FP64 product in disguise)
    assert np.max(err / bound) > 1e-3


@pytest.mark.parametrize("n", [300, 700])
def test_lu_factor_fp64_pivots_match_the_reference_lu(n):
    """FP64 factors: P A = L U to rounding, |L| <= 1, and the pivot sequence of
    the oracle's restatement of the reference's recursive LU (300 rows: two
    outer blocks, 700 rows: three)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    LU, perm = ig.lu_factor(A, ig.SOLVE_FP64)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.max(np.abs(L)) <= 1.0
    assert np.linalg.norm(A[perm] - L @ U) <= 1e-13 * np.linalg.norm(A)
    A2 = A.copy()
    rc, indx = orc.lu_decomp_inplace(A2)
    assert rc == 0
    p = np.arange(n)
    for j, i in enumerate(indx):
        p[[j, i]] = p[[i, j]]
    assert np.array_equal(p, perm)


def test_lu_factor_low_precision_update_is_the_only_difference():
    """With FP32 updates the factors of a 700-row matrix differ from the FP64
    ones by the order of 2^-24 (the first 256 columns not at all: their block
    is factored in FP64), and reproduce A to that order."""
    rng = np.random.default_rng(7)
    n = 700
    A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    LU64, p64 = ig.lu_factor(A, ig.SOLVE_FP64)
    LU32, p32 = ig.lu_factor(A, ig.SOLVE_FP32_IR)
    assert np.array_equal(p64, p32)
    assert np.array_equal(LU64[:, :ig.NB_LOWP], LU32[:, :ig.NB_LOWP])
    assert not np.array_equal(LU64[:, ig.NB_LOWP:], LU32[:, ig.NB_LOWP:])
    d = np.max(np.abs(LU64 - LU32)) / np.max(np.abs(LU64))
    assert 2.0 ** -40 < d < 2.0 ** -18


def _precond_op(A, LU, perm):
    return lambda v: ig.lu_solve(LU, perm, (A @ v.reshape(-1, 1))).ravel()


def test_gmres_exact_preconditioner_one_step():
    """With the FP64 factors of A itself the preconditioned matrix is the
    identity: GMRES stops after one step with D = A^-1 R."""
    rng = np.random.default_rng(3)
    n = 80
    A = rng.standard_normal((n, n))
    R = rng.standard_normal((n, 3))
    LU, perm = ig.lu_factor(A, ig.SOLVE_FP64)
    D, steps = ig.gmres_correction(A, LU, perm, R)
    assert steps == 1
    assert np.linalg.norm(D - np.linalg.solve(A, R)) <= 1e-12 * np.linalg.norm(D)


def _minres_bruteforce(A, LU, perm, r, k):
    """argmin over d in K_k(M A, M r) of ||M r - M A d||_2, M = (LU)^-1 P, from
    an explicit orthonormal Krylov basis and a dense least-squares solve."""
    op = _precond_op(A, LU, perm)
    z = ig.lu_solve(LU, perm, r.reshape(-1, 1)).ravel()
    K = [z / np.linalg.norm(z)]
    for _ in range(k - 1):
        w = op(K[-1])
        for q in K:
            w = w - np.dot(q, w) * q
        for q in K:
            w = w - np.dot(q, w) * q
        K.append(w / np.linalg.norm(w))
    Q = np.stack(K, axis=1)
    BQ = np.stack([op(Q[:, i]) for i in range(k)], axis=1)
    y, *_ = np.linalg.lstsq(BQ, z, rcond=None)
    return Q @ y


def _poor_factors(A, seed):
    rng = np.random.default_rng(seed)
    E = rng.standard_normal(A.shape)
    return ig.lu_factor(A + 0.3 * np.linalg.norm(A, 2) / np.linalg.norm(E, 2) * E, ig.SOLVE_FP64)


@pytest.mark.parametrize("k", [1, 2, 4, 6])
def test_gmres_minimises_the_preconditioned_residual(k):
    """After k steps the correction is the minimiser of the preconditioned
    residual over the k-dimensional Krylov space (dense brute force)."""
    rng = np.random.default_rng(11)
    n = 40
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    LU, perm = _poor_factors(A, 12)
    R = rng.standard_normal((n, 2))
    D, steps = ig.gmres_correction(A, LU, perm, R, m=k, tol=0.0, probe=100)
    assert steps == k
    for c in range(2):
        want = _minres_bruteforce(A, LU, perm, R[:, c], k)
        assert np.linalg.norm(D[:, c] - want) <= 1e-9 * np.linalg.norm(want), (k, c)


def test_gmres_converges_and_gives_up_as_stated():
    """With poor but usable factors GMRES reaches its tolerance within the
    dimension; with a useless preconditioner (factors of an unrelated matrix)
    it stops after the 10 probe steps."""
    rng = np.random.default_rng(5)
    n = 60
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    LU, perm = _poor_factors(A, 6)
    R = rng.standard_normal((n, 1))
    D, steps = ig.gmres_correction(A, LU, perm, R, m=40, tol=1e-8)
    assert 1 < steps < 40
    z = ig.lu_solve(LU, perm, R)
    res = ig.lu_solve(LU, perm, R - A @ D)
    assert np.linalg.norm(res) <= 1.0001e-8 * np.linalg.norm(z)
    n = 300
    A = rng.standard_normal((n, n))
    R = rng.standard_normal((n, 1))
    LUb, pb = ig.lu_factor(rng.standard_normal((n, n)), ig.SOLVE_FP64)
    _, steps = ig.gmres_correction(A, LUb, pb, R, m=40, tol=1e-8)
    assert steps == ig.GMRES_PROBE


def test_planted_mistakes_fail_the_pins():
    """A wrong-sign Givens rotation breaks the minimal-residual pin; dropping
    the preconditioner breaks the one-step pin."""
    rng = np.random.default_rng(11)
    n = 40
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    LU, perm = _poor_factors(A, 12)
    R = rng.standard_normal((n, 1))
    D, _ = ig.gmres_correction(A, LU, perm, R, m=4, tol=0.0, probe=100, _plant="givens_sign")
    want = _minres_bruteforce(A, LU, perm, R[:, 0], 4)
    assert np.linalg.norm(D[:, 0] - want) > 1e-3 * np.linalg.norm(want)
    LUe, pe = ig.lu_factor(A, ig.SOLVE_FP64)
    _, steps = ig.gmres_correction(A, LUe, pe, R, _plant="no_preconditioner")
    assert steps > 1


def test_driver_paths():
    """FP32 factors at kappa 1e2: plain steps only.  BF16 factors at n = 700,
    kappa 1e3 (kappa u_bf16 above 1): the plain step stalls and GMRES finishes.
    kappa 1e13 with BF16 factors: the low-precision attempt fails and the FP64
    factorisation answers.  Every result meets the 1e-14 backward error (the
    C oracle's long-double residual)."""
    A, B = _system(700, 1e2, 3, 1)
    rc, X, st = ig.solve_ir(A, B, ig.SOLVE_FP32_IR)
    assert rc == 0 and st["gmres"] == 0 and st["fell_back"] == 0 and 1 <= st["iterations"] <= 4
    assert orc.backward_error(A, np.ascontiguousarray(X), B) <= 1e-14
    A, B = _system(700, 1e3, 3, 2)
    rc, X, st = ig.solve_ir(A, B, ig.SOLVE_BF16_IR)
    assert rc == 0 and st["gmres"] == 1 and st["krylov_steps"] > 0 and st["fell_back"] == 0
    assert orc.backward_error(A, np.ascontiguousarray(X), B) <= 1e-14
    A, B = _system(700, 1e13, 2, 3)
    rc, X, st = ig.solve_ir(A, B, ig.SOLVE_BF16_IR)
    assert rc == 0 and st["fell_back"] == 1
    assert orc.backward_error(A, np.ascontiguousarray(X), B) <= 1e-14


def test_gmres_step_gaining_less_than_tenfold_ends_the_attempt(monkeypatch):
    """A GMRES step must gain tenfold like a plain step; one that gains about
    3x (a correction of 0.7 A^-1 r leaves 0.3 r) ends the low-precision
    attempt after that single step, and the FP64 factors answer.  (A rule
    that only asks each GMRES step to halve the error takes 20 such steps
    here.)"""
    A, B = _system(400, 1e2, 2, 9)
    calls = []

    def weak(A_, LU, perm, R, **kw):
        calls.append(1)
        return 0.7 * np.linalg.solve(A_, R), 5

    monkeypatch.setattr(ig, "gmres_correction", weak)
    rc, X, st = ig.solve_ir(A, B, ig.SOLVE_BF16_IR)
    assert rc == 0 and st["fell_back"] == 1
    assert len(calls) == 1 and st["lowp_krylov_steps"] == 5
    assert orc.backward_error(A, np.ascontiguousarray(X), B) <= 1e-14
