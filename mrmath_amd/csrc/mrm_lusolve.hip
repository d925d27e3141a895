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
constexpr int SOLVE_SPLITS = 4;      // K chunks of the block products for few right-hand sides
constexpr int SOLVE_WIDE_RHS = 256;  // from this many right-hand sides the products fill the GPU unsplit

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
// lives in LDS and is private to its thread.  Each factor row is staged in LDS
// first (one broadcast value per step for all threads), and the sums run in
// four independent chains.
template <bool UPPER>
__global__ void __launch_bounds__(TR) tri_inv_kernel(const double* __restrict__ LU, int64_t lda, int n,
                                                     double* __restrict__ Dinv) {
    __shared__ double X[TR][TR + 1];
    __shared__ double drow[TR];
    const int b = blockIdx.x, t = threadIdx.x;
    const int r0 = b * TR;
    const int rb = n - r0 < TR ? n - r0 : TR;
    const double* D = LU + int64_t(r0) * lda + r0;
    for (int s = 0; s < rb; ++s) {
        const int i = UPPER ? rb - 1 - s : s;
        __syncthreads();
        if (t < rb) drow[t] = D[int64_t(i) * lda + t];
        __syncthreads();
        const int k0 = UPPER ? i + 1 : 0, k1 = UPPER ? rb : i;
        double a0 = i == t ? 1.0 : 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = k0;
        for (; k + 3 < k1; k += 4) {
            a0 -= drow[k] * X[k][t];
            a1 -= drow[k + 1] * X[k + 1][t];
            a2 -= drow[k + 2] * X[k + 2][t];
            a3 -= drow[k + 3] * X[k + 3][t];
        }
        for (; k < k1; ++k) a0 -= drow[k] * X[k][t];
        const double acc = (a0 + a1) + (a2 + a3);
        X[i][t] = UPPER ? acc / drow[i] : acc;
    }
    __syncthreads();
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

size_t lu_solve_inverse_elems(int n) {
    if (n <= TR) return 0;
    return 2 * size_t((n + TR - 1) / TR) * TR * TR;
}

int lu_solve_inverses(const double* LU, int64_t lda, int n, double* dinv, hipStream_t s) {
    if (n <= TR) return MRM_OK;
    const int nb = (n + TR - 1) / TR;
    hipLaunchKernelGGL(tri_inv_kernel<false>, dim3(nb), dim3(TR), 0, s, LU, lda, n, dinv);
    hipLaunchKernelGGL(tri_inv_kernel<true>, dim3(nb), dim3(TR), 0, s, LU, lda, n, dinv + size_t(nb) * TR * TR);
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

template <typename T>
int lu_solve(const T* LU, int64_t lda, int n, const int* perm, const T* B, int64_t ldb, T* X, int64_t ldx, int nrhs,
             hipStream_t s, const double* dinv) {
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
    const size_t inv_elems = lu_solve_inverse_elems(n);
    const int64_t ldy = nrhs;
    void* mem = nullptr;
    // Y, the partial products of the split products, and the inverses unless given
    const size_t rhs = size_t(n) * size_t(nrhs);
    const int splits = nrhs < SOLVE_WIDE_RHS ? SOLVE_SPLITS : 1;
    const size_t part_elems = splits > 1 ? size_t(splits) * rhs : 0;
    MRM_HIP_TRY(hipMallocAsync(&mem, (rhs + part_elems + (dinv ? 0 : inv_elems)) * sizeof(double), s));
    double* Y = static_cast<double*>(mem);
    double* part = Y + rhs;
    if (!dinv) {
        double* own = part + part_elems;
        int rc = lu_solve_inverses(LU, lda, n, own, s);
        if (rc) {
            (void)hipFreeAsync(mem, s);
            return rc;
        }
        dinv = own;
    }
    const double* Linv = dinv;
    const double* Uinv = dinv + size_t(nb) * TR * TR;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned((int64_t(n) * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb, X,
                       ldx, perm, n, nrhs);
    int rc = MRM_OK;
    // K of every product is one diagonal block (<= 128): with few right-hand
    // sides, cut it into SOLVE_SPLITS chunks so that the skinny products
    // occupy that many times the workgroups
    auto mm = [&](int64_t M, int64_t K, const double* Ap, int64_t lda_, const double* Bp, int64_t ldb_, double* Cp,
                  int64_t ldc_, int op) {
        return mrm_launch_dgemm_splitk(MRM_NN, M, nrhs, K, Ap, lda_ * 8, Bp, ldb_ * 8, Cp, ldc_ * 8, op,
                                       splits, splits > 1 ? part : nullptr, s);
    };
    for (int b = 0; b < nb && !rc; ++b) {
        const int r0 = b * TR, rb = n - r0 < TR ? n - r0 : TR;
        rc = mm(rb, rb, Linv + size_t(b) * TR * TR, TR, X + int64_t(r0) * ldx, ldx, Y + int64_t(r0) * ldy, ldy,
                MRM_DO_NONE);
        if (!rc && r0 + rb < n)
            rc = mm(n - r0 - rb, rb, LU + int64_t(r0 + rb) * lda + r0, lda, Y + int64_t(r0) * ldy, ldy,
                    X + int64_t(r0 + rb) * ldx, ldx, MRM_DO_SUB);
    }
    for (int b = nb - 1; b >= 0 && !rc; --b) {
        const int r0 = b * TR, rb = n - r0 < TR ? n - r0 : TR;
        rc = mm(rb, rb, Uinv + size_t(b) * TR * TR, TR, Y + int64_t(r0) * ldy, ldy, X + int64_t(r0) * ldx, ldx,
                MRM_DO_NONE);
        if (!rc && r0 > 0) rc = mm(r0, rb, LU + r0, lda, X + int64_t(r0) * ldx, ldx, Y, ldy, MRM_DO_SUB);
    }
    // the allocation is released in stream order after the last use
    const hipError_t e = hipFreeAsync(mem, s);
    if (rc) return rc;
    MRM_HIP_TRY(e);
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

template int lu_solve<double>(const double*, int64_t, int, const int*, const double*, int64_t, double*, int64_t, int,
                              hipStream_t, const double*);

}  // namespace mrm
