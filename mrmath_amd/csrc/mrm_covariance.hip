// mrm_covariance.hip -- the covariance caller of the multiply on gfx950.
//
// Replaces Covariance (reference src/Statistics.pas:599-639; the object form
// TCorrelation.Covariance, src/Corr.pas:222-242), which allocates a centred
// copy and a transposed copy of A, multiplies them through MatrixMult and
// scales the product:
//   aMean = column means of A;  ac = A - aMean;  dest = ac^T ac * (1/m)
// with m = height - 1 (unbiased) or height.  Here the centred copy is the only
// copy: the product reads it as the transposed operand of a TN GEMM, whose
// operands are one matrix, so only the tiles on and above the diagonal run and
// the long K (the observations) is split into chunks (mrm_gemm.hip).
//
// HBM traffic per call for X (h x w): the mean pass reads X once, the centring
// pass reads X and writes ac once, the product reads ac once (per K chunk its
// panels are L2-resident), the scale pass touches w^2 doubles.

#include <cstdint>

#include "mrm_common.h"
#include "mrm_host.h"

namespace {

constexpr int COV_THREADS = 256;
constexpr int COV_ROWS = 512;   // rows per partial column sum

int64_t row_blocks(int64_t height) { return (height + COV_ROWS - 1) / COV_ROWS; }

// part[b][c] = sum of X[r][c] over the rows r of row block b: wave q of the
// workgroup sums rows q, q+4, q+8, ... of the block for 64 consecutive columns
// (one per lane: each row segment is one coalesced 512-byte read), then the
// four wave sums are added in wave order.  Fixed order, so deterministic.
__global__ void __launch_bounds__(COV_THREADS) colsum_partial_kernel(const double* __restrict__ X, int64_t ldx,
                                                                    int64_t width, int64_t height,
                                                                    double* __restrict__ part) {
    __shared__ double ws[4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t c = int64_t(blockIdx.x) * 64 + lane;
    const int64_t b = blockIdx.y;
    const int64_t r0 = b * COV_ROWS;
    const int64_t r1 = r0 + COV_ROWS < height ? r0 + COV_ROWS : height;
    double s = 0.0;
    if (c < width) {
        const double* p = X + c;
        int64_t r = r0 + q;
        // four independent loads in flight per step, summed in row order
        for (; r + 12 < r1; r += 16) {
            const double a0 = p[r * ldx], a1 = p[(r + 4) * ldx], a2 = p[(r + 8) * ldx], a3 = p[(r + 12) * ldx];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; r < r1; r += 4) s += p[r * ldx];
    }
    ws[q][lane] = s;
    __syncthreads();
    if (q == 0 && c < width) part[b * width + c] = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
}

// mean[c] = (sum over the partials of column c) / height (GenericMtxMean
// divides the column sum by the height, src/SimpleMatrixOperations.pas:432-445);
// wave q sums partials q, q+4, ..., the four sums are added in wave order.
__global__ void __launch_bounds__(COV_THREADS) colmean_kernel(const double* __restrict__ part, int64_t width,
                                                             int64_t blocks, int64_t height,
                                                             double* __restrict__ mean) {
    __shared__ double ws[4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t c = int64_t(blockIdx.x) * 64 + lane;
    double s = 0.0;
    if (c < width)
        for (int64_t b = q; b < blocks; b += 4) s += part[b * width + c];
    ws[q][lane] = s;
    __syncthreads();
    if (q == 0 && c < width) mean[c] = (((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane]) / double(height);
}

// ac = X - mean (MatrixSubVec row-wise, src/Statistics.pas:625); ac is dense
// (stride = width).  Two doubles per thread.
__global__ void __launch_bounds__(COV_THREADS) centre_kernel(const double* __restrict__ X, int64_t ldx,
                                                            int64_t width, int64_t height,
                                                            const double* __restrict__ mean,
                                                            double* __restrict__ ac) {
    const int64_t half = (width + 1) / 2;  // column pairs per row
    const int64_t idx = int64_t(blockIdx.x) * COV_THREADS + threadIdx.x;
    if (idx >= half * height) return;
    const int64_t r = idx / half, c = 2 * (idx - r * half);
    const double* x = X + r * ldx + c;
    double* d = ac + r * width + c;
    d[0] = x[0] - mean[c];
    if (c + 1 < width) d[1] = x[1] - mean[c + 1];
}

// C = C * scale + 0 (MatrixScaleAndAdd(dest, ..., 0, 1/m), src/Statistics.pas:634;
// GenericMtxScaleAndAdd, src/SimpleMatrixOperations.pas:1220-1232)
__global__ void __launch_bounds__(COV_THREADS) scale_kernel(double* __restrict__ C, int64_t ldc, int64_t n,
                                                           double scale) {
    const int64_t idx = int64_t(blockIdx.x) * COV_THREADS + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t r = idx / n, c = idx - r * n;
    C[r * ldc + c] = C[r * ldc + c] * scale + 0.0;
}

// Workspace layout: [ac: h*w][partials: blocks*w][mean: w][split products].
struct CovLayout {
    size_t ac, part, mean, gemm, total;
};

CovLayout cov_layout(int64_t width, int64_t height) {
    CovLayout L;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    L.ac = 0;
    L.part = up(size_t(width) * size_t(height) * 8);
    L.mean = L.part + up(size_t(row_blocks(height)) * size_t(width) * 8);
    L.gemm = L.mean + up(size_t(width) * 8);
    L.total = L.gemm + up(mrm_dgemm_work_bytes(MRM_TN, width, width, height, true));
    return L;
}

unsigned grid(int64_t n) { return unsigned((n + COV_THREADS - 1) / COV_THREADS); }

// X (height x width, element stride ldx) -> C (width x width, element stride ldc)
int covariance_device(const double* X, int64_t ldx, int64_t width, int64_t height, double* C, int64_t ldc,
                      int unbiased, char* work, const CovLayout& L, hipStream_t s) {
    double* ac = reinterpret_cast<double*>(work + L.ac);
    double* part = reinterpret_cast<double*>(work + L.part);
    double* mean = reinterpret_cast<double*>(work + L.mean);
    const int64_t blocks = row_blocks(height);
    const unsigned col_groups = unsigned((width + 63) / 64);
    hipLaunchKernelGGL(colsum_partial_kernel, dim3(col_groups, unsigned(blocks)), dim3(COV_THREADS), 0, s, X, ldx,
                       width, height, part);
    hipLaunchKernelGGL(colmean_kernel, dim3(col_groups), dim3(COV_THREADS), 0, s, part, width, blocks, height, mean);
    hipLaunchKernelGGL(centre_kernel, dim3(grid((width + 1) / 2 * height)), dim3(COV_THREADS), 0, s, X, ldx, width,
                       height, mean, ac);
    MRM_HIP_TRY(hipGetLastError());
    // ac^T ac: both operands are ac, so the symmetric path computes the upper tiles
    int rc = mrm_launch_dgemm_ws(MRM_TN, width, width, height, ac, width * 8, ac, width * 8, C, ldc * 8, MRM_DO_NONE,
                                 work + L.gemm, L.total - L.gemm, s);
    if (rc) return rc;
    const int64_t m = unbiased ? height - 1 : height;
    hipLaunchKernelGGL(scale_kernel, dim3(grid(width * width)), dim3(COV_THREADS), 0, s, C, ldc, width,
                       1.0 / double(m));
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

// MatrixVar of a single row or column (GenericMtxVar,
// src/SimpleMatrixOperations.pas:450-520): mean by a sequential sum, then the
// sequential sum of squared deviations, divided by max(1, n - 1) or n.
double vector_variance(const double* A, int64_t LineWidthA, int64_t width, int64_t height, bool unbiased) {
    const int64_t n = width * height;
    auto at = [&](int64_t i) {
        return height == 1 ? A[i] : *reinterpret_cast<const double*>(reinterpret_cast<const char*>(A) + i * LineWidthA);
    };
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum = sum + at(i);
    const double mean = sum / double(n);
    double v = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double d = at(i) - mean;
        v = v + d * d;
    }
    return unbiased ? v / double(n - 1 > 1 ? n - 1 : 1) : v / double(n);
}

}  // namespace

extern "C" {

int64_t mrm_covariance_workspace_bytes(int64_t width, int64_t height) {
    if (width < 2 || height < 2) return 0;
    return int64_t(cov_layout(width, height).total);
}

int32_t mrm_dev_covariance(const double* X, int64_t ldx, int64_t width, int64_t height, double* C, int64_t ldc,
                           int32_t unbiased, void* work, int64_t work_bytes, void* stream) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if (width < 2 || height < 2 || ldx < width * 8 || ldc < width * 8 || ((ldx | ldc) & 7)) {
        mrm_set_error(MRM_BAD_DIMS, "covariance: needs width, height >= 2 and line widths >= width*8");
        return MRM_BAD_DIMS;
    }
    if (!X || !C || !work) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    const CovLayout L = cov_layout(width, height);
    if (work_bytes < int64_t(L.total)) {
        mrm_set_error(MRM_BAD_ARG, "covariance: workspace of %lld bytes, %zu needed", (long long)work_bytes, L.total);
        return MRM_BAD_ARG;
    }
    return covariance_device(X, ldx / 8, width, height, C, ldc / 8, unbiased, static_cast<char*>(work), L,
                             mrm_resolve_stream(stream));
}

// Covariance (src/Statistics.pas:599-639)
int32_t mrm_covariance(double* dest, int64_t destLineWidth, const double* A, int64_t LineWidthA, int64_t width,
                       int64_t height, int32_t unbiased) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if (width <= 0 || height <= 0 || LineWidthA < width * 8 || (LineWidthA & 7) || (destLineWidth & 7)) {
        mrm_set_error(MRM_BAD_DIMS, "covariance: bad dimensions or line widths");
        return MRM_BAD_DIMS;
    }
    if (!dest || !A) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    if (height == 1 || width == 1) {
        // the reference's degenerate case: the vector's variance (MatrixVar)
        *dest = vector_variance(A, LineWidthA, width, height, unbiased != 0);
        return MRM_OK;
    }
    if (destLineWidth < width * 8) {
        mrm_set_error(MRM_BAD_DIMS, "Error destlinewidth to short");
        return MRM_BAD_DIMS;
    }
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    const CovLayout L = cov_layout(width, height);
    void *dX, *dC, *dW;
    if ((rc = mrm_staging(0, size_t(width * height) * 8, &dX))) return rc;
    if ((rc = mrm_staging(2, size_t(width * width) * 8, &dC))) return rc;
    if ((rc = mrm_staging(3, L.total, &dW))) return rc;
    if ((rc = mrm_copy_h2d(dX, size_t(width) * 8, A, size_t(LineWidthA),
                           size_t(width) * 8, size_t(height), s))) return rc;
    rc = covariance_device(static_cast<const double*>(dX), width, width, height, static_cast<double*>(dC), width,
                           unbiased, static_cast<char*>(dW), L, s);
    if (rc) return rc;
    if ((rc = mrm_copy_d2h(dest, size_t(destLineWidth), dC, size_t(width) * 8,
                           size_t(width) * 8, size_t(width), s))) return rc;
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

}  // extern "C"
