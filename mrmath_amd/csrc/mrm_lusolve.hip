// mrm_lusolve.hip -- X = (LU)^-1 P B for many right-hand sides.
//
// Replaces the substitutions of MatrixLUBackSubst (reference
// src/LinAlgLU.pas:315-385), which the reference runs column by column
// (MatrixLinEQSolve calls it once per right-hand side, :481-592).  Up to 128
// unknowns the substitution is the reference's, one thread per right-hand
// side.  Beyond, all right-hand sides go together and the triangular solves
// become GEMMs (the block-inverse scheme of GPU TRSMs): the
// diagonal blocks (128 x 128) of L and U are inverted once per call, all
// blocks in parallel, and then, block row by block row,
//   lower: Y_b = Linv_b X_b;   X_{b+1..} -= L_{b+1..,b} Y_b
//   upper: X_b = Uinv_b Y_b;   Y_{..b-1} -= U_{..b-1,b} X_b
// with every product on the FP64 MFMA GEMM.  The inverses and Y live in
// stream-ordered allocations of the call.

#include <cstdint>

#include "mrm_common.h"
#include "mrm_lu.h"

namespace {

constexpr int TR = 128;  // diagonal block size

__global__ void __launch_bounds__(256) gather_kernel(const double* __restrict__ B, int64_t ldb,
                                                     double* __restrict__ Y, int64_t ldy, const int* __restrict__ perm,
                                                     int n, int ncols) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(n) * ncols) return;
    const int i = int(idx / ncols), c = int(idx % ncols);
    Y[int64_t(i) * ldy + c] = B[int64_t(perm[i]) * ldb + c];
}

// Inverse of diagonal block b (rows/columns [b*TR, b*TR + rb)) of the unit
// lower (UPPER = false) or upper factor, written as a TR x TR row-major block
// of Dinv.  Thread t computes column t by substitution on e_t; the column
// lives in LDS and is private to its thread, the factor entries are the same
// for all threads at each step.
template <bool UPPER>
__global__ void __launch_bounds__(TR) tri_inv_kernel(const double* __restrict__ LU, int64_t lda, int n,
                                                     double* __restrict__ Dinv) {
    __shared__ double X[TR][TR + 1];
    const int b = blockIdx.x, t = threadIdx.x;
    const int r0 = b * TR;
    const int rb = n - r0 < TR ? n - r0 : TR;
    const double* D = LU + int64_t(r0) * lda + r0;
    if (!UPPER) {
        for (int i = 0; i < rb; ++i) {
            double acc = i == t ? 1.0 : 0.0;
            const double* di = D + int64_t(i) * lda;
            for (int k = 0; k < i; ++k) acc -= di[k] * X[k][t];
            X[i][t] = acc;
        }
    } else {
        for (int i = rb - 1; i >= 0; --i) {
            double acc = i == t ? 1.0 : 0.0;
            const double* di = D + int64_t(i) * lda;
            for (int k = i + 1; k < rb; ++k) acc -= di[k] * X[k][t];
            X[i][t] = acc / di[i];
        }
    }
    double* out = Dinv + int64_t(b) * TR * TR;
    for (int i = 0; i < rb; ++i) out[int64_t(i) * TR + t] = t < rb ? X[i][t] : 0.0;
}

// n <= TR: plain substitution, one thread per right-hand-side column, in the
// reference's order (forward sums over j ascending, backward sums over j
// ascending then division by the diagonal, src/LinAlgLU.pas:339-383); X holds
// P B on entry.  The factor entries are the same for all threads at each step.
__global__ void __launch_bounds__(64) small_subst_kernel(const double* __restrict__ LU, int64_t lda, int n,
                                                         double* __restrict__ X, int64_t ldx, int nrhs) {
    __shared__ double Y[TR][65];
    const int t = threadIdx.x;
    const int c = blockIdx.x * 64 + t;
    const bool live = c < nrhs;
    for (int i = 0; i < n; ++i) {
        double acc = live ? X[int64_t(i) * ldx + c] : 0.0;
        const double* li = LU + int64_t(i) * lda;
        for (int k = 0; k < i; ++k) acc -= li[k] * Y[k][t];
        Y[i][t] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        double acc = Y[i][t];
        const double* ui = LU + int64_t(i) * lda;
        for (int k = i + 1; k < n; ++k) acc -= ui[k] * Y[k][t];
        Y[i][t] = acc / ui[i];
    }
    if (live)
        for (int i = 0; i < n; ++i) X[int64_t(i) * ldx + c] = Y[i][t];
}

}  // namespace

namespace mrm {

template <typename T>
int lu_solve(const T* LU, int64_t lda, int n, const int* perm, const T* B, int64_t ldb, T* X, int64_t ldx, int nrhs,
             hipStream_t s) {
    static_assert(sizeof(T) == sizeof(double), "the block inverses are FP64");
    if (n == 0 || nrhs == 0) return MRM_OK;
    if (n <= TR) {
        hipLaunchKernelGGL(gather_kernel, dim3(unsigned((int64_t(n) * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb,
                           X, ldx, perm, n, nrhs);
        hipLaunchKernelGGL(small_subst_kernel, dim3(unsigned((nrhs + 63) / 64)), dim3(64), 0, s, LU, lda, n, X, ldx,
                           nrhs);
        MRM_HIP_TRY(hipGetLastError());
        return MRM_OK;
    }
    const int nb = (n + TR - 1) / TR;
    const size_t inv_elems = size_t(nb) * TR * TR;
    const int64_t ldy = nrhs;
    void* mem = nullptr;
    MRM_HIP_TRY(hipMallocAsync(&mem, (2 * inv_elems + size_t(n) * size_t(nrhs)) * sizeof(double), s));
    double* Linv = static_cast<double*>(mem);
    double* Uinv = Linv + inv_elems;
    double* Y = Uinv + inv_elems;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned((int64_t(n) * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb, X,
                       ldx, perm, n, nrhs);
    hipLaunchKernelGGL(tri_inv_kernel<false>, dim3(nb), dim3(TR), 0, s, LU, lda, n, Linv);
    hipLaunchKernelGGL(tri_inv_kernel<true>, dim3(nb), dim3(TR), 0, s, LU, lda, n, Uinv);
    int rc = MRM_OK;
    for (int b = 0; b < nb && !rc; ++b) {
        const int r0 = b * TR, rb = n - r0 < TR ? n - r0 : TR;
        rc = gemm_t<double>(MRM_NN, rb, nrhs, rb, Linv + size_t(b) * TR * TR, TR, X + int64_t(r0) * ldx, ldx,
                            Y + int64_t(r0) * ldy, ldy, MRM_DO_NONE, s);
        if (!rc && r0 + rb < n)
            rc = gemm_t<double>(MRM_NN, n - r0 - rb, nrhs, rb, LU + int64_t(r0 + rb) * lda + r0, lda,
                                Y + int64_t(r0) * ldy, ldy, X + int64_t(r0 + rb) * ldx, ldx, MRM_DO_SUB, s);
    }
    for (int b = nb - 1; b >= 0 && !rc; --b) {
        const int r0 = b * TR, rb = n - r0 < TR ? n - r0 : TR;
        rc = gemm_t<double>(MRM_NN, rb, nrhs, rb, Uinv + size_t(b) * TR * TR, TR, Y + int64_t(r0) * ldy, ldy,
                            X + int64_t(r0) * ldx, ldx, MRM_DO_NONE, s);
        if (!rc && r0 > 0)
            rc = gemm_t<double>(MRM_NN, r0, nrhs, rb, LU + r0, lda, X + int64_t(r0) * ldx, ldx, Y, ldy, MRM_DO_SUB,
                                s);
    }
    // the allocation is released in stream order after the last use
    const hipError_t e = hipFreeAsync(mem, s);
    if (rc) return rc;
    MRM_HIP_TRY(e);
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

template int lu_solve<double>(const double*, int64_t, int, const int*, const double*, int64_t, double*, int64_t, int,
                              hipStream_t);

}  // namespace mrm
