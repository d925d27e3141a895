"""The covariance caller (Covariance, src/Statistics.pas:599-639) on the GPU
against the oracle's restatement, through the C ABI.

Bar: ||C - Cref||_F / ||Ac||_F^2 <= 4*height*2^-53 plus the rounding of the
means (the GPU sums each column in blocks of 512 rows, the reference
sequentially), checked as the GEMM bar with Ac's norm; the degenerate
single-row / single-column case matches the oracle to a few ulps.
"""
import numpy as np
import pytest

import oracle as orc
from mrmath_amd import _lib
from mrmath_amd import statistics as st

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _bar(X):
    Xc = X - X.mean(axis=0)
    m = X.shape[0]
    return 4 * m * U * np.linalg.norm(Xc) ** 2 / (m - 1)


@pytest.mark.parametrize("h,w", [(2, 2), (50, 7), (1000, 130), (5000, 300), (20000, 96)])
@pytest.mark.parametrize("unbiased", [True, False])
def test_covariance_vs_oracle(h, w, unbiased):
    X = orc.pm1_matrix(w, h, 5000 + h + w) + 3.0  # an offset the centring must remove
    C = np.zeros((w, w))
    st.Covariance(C, X, unbiased)
    ref = orc.covariance(X, unbiased)
    assert np.linalg.norm(C - ref) <= _bar(X)
    assert np.array_equal(C, C.T)


def test_covariance_strided_input_and_object_form():
    raw = orc.pm1_matrix(40, 300, 77)
    X = raw[:, 3:35]  # a padded view: LineWidth 320 bytes
    C = st.TCorrelation.Covariance(X)
    ref = orc.covariance(np.ascontiguousarray(X))
    assert np.linalg.norm(C - ref) <= _bar(np.ascontiguousarray(X))


@pytest.mark.parametrize("shape", [(1, 33), (33, 1), (1, 1)])
def test_single_row_or_column_is_its_variance(shape):
    x = orc.pm1_matrix(shape[1], shape[0], 91)
    C = np.zeros((1, 1))
    st.Covariance(C, x)
    assert C[0, 0] == pytest.approx(orc.covariance(x)[0, 0], rel=1e-15, abs=0)


def test_device_entry_pca_shape():
    """X 200000 x 512 (the PCA covariance shape) on the device entry: symmetric,
    the centring removes an offset of 100, sampled rows against the oracle's
    formula on the host (numpy of the centred data in FP64)."""
    import torch
    lib = _lib.load()
    h, w = 200000, 512
    g = torch.Generator(device="cuda").manual_seed(2001)
    X = torch.rand(h, w, dtype=torch.float64, device="cuda", generator=g) * 2 - 1 + 100.0
    C = torch.empty(w, w, dtype=torch.float64, device="cuda")
    wb = lib.mrm_covariance_workspace_bytes(w, h)
    assert wb >= h * w * 8
    W = torch.empty(wb, dtype=torch.uint8, device="cuda")
    s_ = torch.cuda.current_stream().cuda_stream
    assert lib.mrm_dev_covariance(X.data_ptr(), w * 8, w, h, C.data_ptr(), w * 8, 1, W.data_ptr(), wb, s_) == 0
    torch.cuda.synchronize()
    Cn = C.cpu().numpy()
    assert np.array_equal(Cn, Cn.T)
    Xn = X.cpu().numpy()
    Xc = Xn - Xn.mean(axis=0)
    cols = [0, 1, 300, 511]
    ref = (Xc[:, cols].T @ Xc) / (h - 1)
    err = np.linalg.norm(Cn[cols] - ref) / (np.linalg.norm(Xc[:, cols]) * np.linalg.norm(Xc) / (h - 1))
    assert err <= 4 * h * U, err
    # diagonal: the variance of U[-1,1) is 1/3
    assert np.allclose(np.diag(Cn), 1 / 3, atol=0.01)


def test_device_entry_errors():
    lib = _lib.load()
    assert lib.mrm_covariance_workspace_bytes(1, 10) == 0
    assert lib.mrm_dev_covariance(None, 8, 1, 10, None, 8, 1, None, 0, None) == _lib.MRM_BAD_DIMS
    X = np.zeros((10, 3))
    with pytest.raises(ValueError):
        st.Covariance(np.zeros((2, 2)), X)
