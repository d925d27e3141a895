"""CPU restatement of the mixed-precision solve with FP64 refinement and GMRES
corrections (the path of mrm_dev_lineq_solve_ir_stats).

TEST INFRASTRUCTURE ONLY: a checker for the GPU library; only tests/ may
import it.  numpy / scipy, FP64 unless stated.

The steps, as the GPU runs them:
  factorisation  right-looking LU with partial pivoting in outer blocks of 256
                 columns: each block's columns factored in FP64 with the
                 reference's pivot rule (first maximum of |a| in the current
                 row order, singular at |pivot| <= 10/MaxDouble, reciprocal
                 scaling; src/LinAlgLU.pas:239-280), the rows interchanged
                 across the matrix, U12 = L11^-1 A12 in FP64, and the trailing
                 update A22 -= L21 U12 (src/LinAlgLU.pas:211-227) with both
                 operands rounded to FP32, to BF16, or to two BF16 pieces (the
                 product Ah Bh + Ah Bl + Al Bh as one sum over 3K), multiplied
                 with FP32 sums;
  refinement     x = (LU)^-1 P b, then per step r = b - A x in FP64 and the
                 backward error max_j ||r_j||inf / (||A||inf ||x_j||inf +
                 ||b_j||inf); plain steps x += (LU)^-1 P r (the reference's
                 MatrixLinEQSolve loop, src/LinAlgLU.pas:542-586) while each
                 gains at least tenfold, GMRES corrections after the first
                 that does not; a GMRES step that also gains less than tenfold
                 ends the attempt, which is then redone with FP64 factors;
  GMRES          left-preconditioned by the factors (Carson and Higham, SISC
                 2017 and 2018): V0 = (LU)^-1 P r / beta, Arnoldi on
                 (LU)^-1 P A with classical Gram-Schmidt twice, Givens rotations
                 on each Hessenberg column, stop when |g_(k+1)| <= tol beta, at
                 breakdown, at the dimension, or after 10 steps that have not
                 reduced every column tenfold; one Krylov space per right-hand
                 side.
The in-block FP64 factorisation here is unblocked (the GPU factors 32-column
panels and updates the rest of the block by GEMM): the same operations in
another summation order, so the factors agree to rounding, not bit for bit.
The GMRES restatement takes a `_plant` argument through which the tests plant
known mistakes, to show that their pins catch them.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular

SOLVE_FP64, SOLVE_FP32_IR, SOLVE_BF16_IR, SOLVE_BF16X3_IR = 0, 1, 2, 3
OK, SINGULAR, NOT_CONVERGED = 0, 1, 2
NB_LOWP = 256          # outer block of the low-precision factorisation
GMRES_DIM = 40         # Krylov dimension per correction
GMRES_TOL = 1e-8       # relative tolerance on the preconditioned residual
GMRES_PROBE = 10       # steps before a correction that gains less than 10x is given up
PLAIN_RATE = 0.1       # a plain step must gain this factor, else GMRES takes over
TINY = 10.0 / np.finfo(np.float64).max


def to_bf16(x: np.ndarray) -> np.ndarray:
    """FP64 -> FP32 (round to nearest) -> BF16 (round to nearest even), as
    float32 values."""
    u = np.ascontiguousarray(x, dtype=np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def lowp_product(L: np.ndarray, U: np.ndarray, mode: int) -> np.ndarray:
    """L U with the operands rounded as the trailing update rounds them and
    FP32 sums (float32 matmul), returned in FP64."""
    if mode == SOLVE_FP32_IR:
        P = L.astype(np.float32) @ U.astype(np.float32)
    elif mode == SOLVE_BF16_IR:
        P = to_bf16(L) @ to_bf16(U)
    elif mode == SOLVE_BF16X3_IR:
        Lh, Uh = to_bf16(L), to_bf16(U)
        Ll = to_bf16(L - Lh.astype(np.float64))
        Ul = to_bf16(U - Uh.astype(np.float64))
        P = np.concatenate([Lh, Lh, Ll], axis=1) @ np.concatenate([Uh, Ul, Uh], axis=0)
    else:
        raise ValueError(f"mode {mode}")
    return P.astype(np.float64)


def lu_factor(A: np.ndarray, mode: int = SOLVE_FP64, nb: int = NB_LOWP):
    """(LU, perm) with P A = L U, P = identity[perm]; None when singular."""
    LU = np.array(A, dtype=np.float64, copy=True)
    n = LU.shape[0]
    perm = np.arange(n)
    for J in range(0, n, nb):
        jb = min(nb, n - J)
        for j in range(J, J + jb):
            col = np.abs(LU[j:, j])
            p = j + int(np.argmax(col))  # the first maximum
            if not (col[p - j] > TINY):
                return None
            if p != j:
                LU[[j, p], :] = LU[[p, j], :]
                perm[[j, p]] = perm[[p, j]]
            inv = 1.0 / LU[j, j]
            LU[j + 1:, j] *= inv
            LU[j + 1:, j + 1:J + jb] -= np.outer(LU[j + 1:, j], LU[j, j + 1:J + jb])
        if J + jb < n:
            LU[J:J + jb, J + jb:] = solve_triangular(LU[J:J + jb, J:J + jb], LU[J:J + jb, J + jb:], lower=True,
                                                     unit_diagonal=True)
            L21, U12 = LU[J + jb:, J:J + jb], LU[J:J + jb, J + jb:]
            if mode == SOLVE_FP64:
                LU[J + jb:, J + jb:] -= L21 @ U12
            else:
                LU[J + jb:, J + jb:] -= lowp_product(L21, U12, mode)
    return LU, perm


def lu_solve(LU: np.ndarray, perm: np.ndarray, B: np.ndarray) -> np.ndarray:
    """(LU)^-1 P B."""
    Y = solve_triangular(LU, B[perm], lower=True, unit_diagonal=True)
    return solve_triangular(LU, Y, lower=False)


def backward_error(A: np.ndarray, X: np.ndarray, B: np.ndarray, R: np.ndarray) -> float:
    """max_j ||R_j||inf / (||A||inf ||X_j||inf + ||B_j||inf) (a column with a
    zero denominator contributes ||R_j||inf)."""
    an = np.max(np.sum(np.abs(A), axis=1)) if A.size else 0.0
    e = 0.0
    for c in range(B.shape[1]):
        den = an * np.max(np.abs(X[:, c])) + np.max(np.abs(B[:, c]))
        rv = np.max(np.abs(R[:, c]))
        v = rv / den if den > 0 else rv
        if not (v <= e):
            e = v
    return e


def gmres_correction(A, LU, perm, R, m: int = GMRES_DIM, tol: float = GMRES_TOL, probe: int = GMRES_PROBE,
                     _plant: str | None = None):
    """(D, steps): D[:, c] approximately solves (LU)^-1 P A d = (LU)^-1 P R[:, c]
    by GMRES with one Krylov space per column."""
    n, nrhs = R.shape

    def prec(X):
        return X if _plant == "no_preconditioner" else lu_solve(LU, perm, X)

    V = np.zeros((m + 1, n, nrhs))
    V0 = prec(R)
    beta = np.sqrt(np.sum(V0 * V0, axis=0))
    ok = (beta > 0) & np.isfinite(beta)
    V[0] = np.where(ok, V0 / np.where(ok, beta, 1.0), 0.0)
    Rm = np.zeros((nrhs, m, m))
    cs = np.zeros((nrhs, m))
    sn = np.zeros((nrhs, m))
    g = np.zeros((nrhs, m + 1))
    g[:, 0] = beta
    kc = np.zeros(nrhs, dtype=int)
    kc[~np.isfinite(beta)] = -1
    active = ok.copy()
    steps = 0
    worst = 1.0
    k = 0
    while k < m and active.any():
        if k >= probe and worst > 0.1:
            break
        W = prec(A @ V[k])
        h = np.einsum("inc,nc->ic", V[:k + 1], W)
        W = W - np.einsum("inc,ic->nc", V[:k + 1], h)
        h2 = np.einsum("inc,nc->ic", V[:k + 1], W)
        W = W - np.einsum("inc,ic->nc", V[:k + 1], h2)
        h = h + h2
        ss = np.sum(W * W, axis=0)
        good = (ss > 0) & (ss < np.inf)
        V[k + 1] = np.where(good, W / np.sqrt(np.where(good, ss, 1.0)), 0.0)
        steps = k + 1
        worst = 0.0
        for c in np.flatnonzero(active):
            hc = np.concatenate([h[:, c], [np.sqrt(ss[c])]])
            for j in range(k):
                a, b = hc[j], hc[j + 1]
                hc[j] = cs[c, j] * a + sn[c, j] * b
                hc[j + 1] = -sn[c, j] * a + cs[c, j] * b
            a, b = hc[k], hc[k + 1]
            rr = np.hypot(a, b)
            cc, s = (a / rr, b / rr) if rr > 0 else (1.0, 0.0)
            if _plant == "givens_sign":
                s = -s
            cs[c, k], sn[c, k] = cc, s
            hc[k] = rr
            Rm[c, :k + 1, k] = hc[:k + 1]
            g[c, k + 1] = -s * g[c, k]
            g[c, k] = cc * g[c, k]
            kc[c] = k + 1
            res = abs(g[c, k + 1])
            worst = max(worst, res / beta[c])
            if not (res > tol * beta[c]) or not (hc[k + 1] >= 0.0) or not (rr > 0.0):
                active[c] = False
        k += 1
    Y = np.zeros((m + 1, nrhs))
    for c in range(nrhs):
        if kc[c] < 0:
            Y[0, c] = np.nan
            continue
        for i in range(kc[c] - 1, -1, -1):
            t = g[c, i] - np.dot(Rm[c, i, i + 1:kc[c]], Y[i + 1:kc[c], c])
            d = Rm[c, i, i]
            Y[i, c] = t / d if d != 0 else 0.0
    nv = k if k > 0 else 1
    D = np.einsum("inc,ic->nc", V[:nv], Y[:nv])
    return D, steps


def _factor_solve_refine(A, B, mode, tol, max_iter):
    st = {"iterations": 0, "gmres": 0, "krylov_steps": 0, "backward_error": float("nan")}
    f = lu_factor(A, mode)
    if f is None:
        return SINGULAR, None, st
    LU, perm = f
    X = lu_solve(LU, perm, B)
    prev = np.inf
    gmres = False
    it = 0
    while True:
        R = B - A @ X
        berr = backward_error(A, X, B, R)
        st["iterations"] = it
        st["backward_error"] = berr
        if berr <= tol:
            return OK, X, st
        if it >= max_iter or np.isnan(berr):
            return NOT_CONVERGED, X, st
        if not (berr < PLAIN_RATE * prev):
            if gmres or mode == SOLVE_FP64:
                if gmres or not (berr < 0.5 * prev):
                    return NOT_CONVERGED, X, st
            else:
                gmres = True
                st["gmres"] = 1
        prev = berr
        if gmres:
            D, k = gmres_correction(A, LU, perm, R)
            st["krylov_steps"] += k
        else:
            D = lu_solve(LU, perm, R)
        X = X + D
        it += 1


def solve_ir(A: np.ndarray, B: np.ndarray, mode: int, tol: float = 1e-14, max_iter: int = 50):
    """(rc, X, stats) of the mixed-precision solve; stats as mrm_solve_stats
    (iterations, gmres, krylov_steps, backward_error, fell_back and the
    low-precision attempt's lowp_iterations / lowp_krylov_steps /
    lowp_backward_error when it fell back)."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    rc, X, st = _factor_solve_refine(A, B, mode, tol, max_iter)
    st["fell_back"] = 0
    if rc in (NOT_CONVERGED, SINGULAR) and mode != SOLVE_FP64:
        first = st
        rc, X, st = _factor_solve_refine(A, B, SOLVE_FP64, tol, max_iter)
        st["fell_back"] = 1
        st["lowp_iterations"] = first["iterations"]
        st["lowp_krylov_steps"] = first["krylov_steps"]
        st["lowp_backward_error"] = first["backward_error"]
    return rc, (np.ascontiguousarray(X) if X is not None else None), st
