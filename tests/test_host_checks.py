"""Argument checks of the host-side mirror that run before any library call
(CPU only: every case raises before the GPU would be touched)."""
import numpy as np
import pytest

from mrmath_amd import _lib
from mrmath_amd import stubswitch as ss


def _ops(h1, w1, h2, w2):
    return np.ones((h1, w1)), np.ones((h2, w2))


@pytest.mark.parametrize("fn,shapes,dest", [
    # NN: dest h1 x w2
    (ss.MatrixMult, (3, 4, 4, 5), (2, 5)),
    (ss.MatrixMultEx, (3, 4, 4, 5), (3, 4)),
    # T1 (mt1^T mt2): dest w1 x w2 -- (h1, w2) is the easy mistake
    (ss.MatrixMultT1, (6, 3, 6, 5), (6, 5)),
    (ss.MatrixMultT1Ex, (6, 3, 6, 5), (2, 5)),
    # T2 (mt1 mt2^T): dest h1 x h2
    (ss.MatrixMultT2, (3, 4, 7, 4), (3, 6)),
    (ss.MatrixMultT2Ex, (3, 4, 7, 4), (2, 7)),
])
def test_dest_without_the_product_shape_raises(fn, shapes, dest):
    A, B = _ops(*shapes)
    D = np.zeros(dest)
    with pytest.raises(ValueError, match="Dimension error"):
        fn(D, A, B)


@pytest.mark.parametrize("trans,shapes,dest", [
    (_lib.NN, (3, 4, 4, 5), (2, 5)),
    (_lib.TN, (6, 3, 6, 5), (6, 5)),
    (_lib.NT, (3, 4, 7, 4), (3, 7 - 1)),
])
def test_ozaki_dest_shape_checked(trans, shapes, dest):
    A, B = _ops(*shapes)
    with pytest.raises(ValueError, match="Dimension error"):
        ss.MatrixMultOzaki(np.zeros(dest), A, B, 6, trans)


def test_ozaki_bad_trans_raises():
    A, B = _ops(3, 4, 4, 5)
    with pytest.raises(ValueError):
        ss.MatrixMultOzaki(np.zeros((3, 5)), A, B, 6, 7)


def test_product_shapes_follow_the_reference_rules():
    assert ss._product_shape(_lib.NN, 4, 3, 5, 4) == (3, 5)
    assert ss._product_shape(_lib.TN, 3, 6, 5, 6) == (3, 5)
    assert ss._product_shape(_lib.NT, 4, 3, 4, 7) == (3, 7)
