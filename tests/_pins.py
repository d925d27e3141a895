"""Known-answer constructions shared by the oracle and GPU tests.

wilkinson(n): the matrix for which partial pivoting has the largest growth
(1 on the diagonal, -1 below it, 1 in the last column; the pivot rule keeps
every diagonal pivot, and U's last column becomes 2^i), so an unrefined FP64
solve loses about (n - 53 + log2 n) bits while the matrix itself is well
conditioned: one step of FP64 refinement restores full accuracy (Skeel).
exact_solve(A, b): Gaussian elimination in rational arithmetic, the exact
solution of the FP64 system as given.
"""
from fractions import Fraction

import numpy as np


def wilkinson(n: int) -> np.ndarray:
    A = np.eye(n)
    for i in range(n):
        A[i, :i] = -1.0
    A[:, n - 1] = 1.0
    return A


def exact_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    n = len(b)
    M = [[Fraction(float(A[i, j])) for j in range(n)] + [Fraction(float(b[i]))] for i in range(n)]
    for k in range(n):
        p = max(range(k, n), key=lambda r: abs(M[r][k]))
        M[k], M[p] = M[p], M[k]
        for r in range(k + 1, n):
            f = M[r][k] / M[k][k]
            if f:
                for c in range(k, n + 1):
                    M[r][c] -= f * M[k][c]
    x = [Fraction(0)] * n
    for i in range(n - 1, -1, -1):
        s = M[i][n] - sum(M[i][j] * x[j] for j in range(i + 1, n))
        x[i] = s / M[i][i]
    return np.array([float(v) for v in x])


def forward_error(X: np.ndarray, Xe: np.ndarray) -> float:
    """max_j ||x_j - xe_j||_inf / ||xe_j||_inf over the columns."""
    return max(float(np.abs(X[:, j] - Xe[:, j]).max() / np.abs(Xe[:, j]).max()) for j in range(X.shape[1]))
