"""Host-side mirror of mrMath's covariance caller, bound to the GPU.

Covariance (src/Statistics.pas:599-639) and TCorrelation.Covariance
(src/Corr.pas:222-242): rows are observations, columns variables; the result
is width x width (or the 1 x 1 variance of a single row or column, as the
reference computes it).  Every call goes through libmrmath_gpu.so.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .stubswitch import _mat


def Covariance(dest: np.ndarray, A: np.ndarray, Unbiased: bool = True) -> np.ndarray:
    """Writes the covariance matrix of A into dest (src/Statistics.pas:600-601)."""
    pd, dw, dh, dlw = _mat(dest, "dest")
    pa, w, h, lwa = _mat(A, "A")
    want = (1, 1) if (w == 1 or h == 1) else (w, w)
    if (dh, dw) != want:
        raise ValueError(f"Dimension error: dest is {dh} x {dw}, the covariance is {want[0]} x {want[1]}")
    _lib.check(_lib.load().mrm_covariance(pd, dlw, pa, lwa, w, h, 1 if Unbiased else 0))
    return dest


class TCorrelation:
    """The object form (src/Corr.pas:222-242)."""

    @staticmethod
    def Covariance(A: np.ndarray, Unbiased: bool = True) -> np.ndarray:
        h, w = A.shape
        n = 1 if (w == 1 or h == 1) else w
        return Covariance(np.zeros((n, n)), A, Unbiased)
