// mrm_lusolve.hip -- X = (LU)^-1 P B for many right-hand sides.
//
// Replaces the substitutions of MatrixLUBackSubst (reference
// src/LinAlgLU.pas:315-385), which the reference runs column by column
// (MatrixLinEQSolve calls it once per right-hand side, :481-592).  Up to 128
// unknowns the substitution is the reference's, one thread per right-hand
// side.  Beyond, all right-hand sides go together and the triangular solves
// become GEMMs (the block-inverse scheme of GPU TRSMs): the diagonal blocks
// (TB = 512) of L and U are inverted once per factorisation, and then, block
// row by block row,
//   lower: Y_b = Linv_b X_b;   X_{b+1..} -= L_{b+1..,b} Y_b
//   upper: X_b = Uinv_b Y_b;   Y_{..b-1} -= U_{..b-1,b} X_b
// with every product on the FP64 MFMA GEMM.  The sweeps are a chain of n/TB
// dependent steps each, so TB sets the latency of a solve: with 128-row
// blocks 64 right-hand sides at n = 16384 took 10.8 ms in 510 launches.
// The 512-row inverses are assembled from 128-row ones (one workgroup per
// block, substitution) by block elimination on the GEMM, every 512-row block
// at once as a strided batch:
//   lower: X_jj = D_j, X_ij = -D_i sum_{k=j..i-1} L_ik X_kj (i > j)
//   upper: X_jj = D_j, X_ij = -D_i sum_{k=i+1..j} U_ik X_kj (i < j)
// for the 128 x 128 sub-blocks, D_i the inverse of the diagonal sub-block.

#include <cstdint>

#include "mrm_common.h"
#include "mrm_lu.h"

namespace {

constexpr int TR = 128;              // sub-block of the inverses (one workgroup's substitution)
constexpr int TB = 512;              // diagonal block of the sweeps
constexpr int SUB = TB / TR;         // sub-blocks per diagonal block
constexpr int SOLVE_MAX_SPLITS = 16; // K chunks of a sweep product at most (skinny products)

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

namespace {

int blocks_of(int n, int b) { return (n + b - 1) / b; }

// The layout of the inverse workspace: the TB x TB inverses of L's and U's
// diagonal blocks (row-major, ld TB), then the TR x TR ones they are built
// from, then one TR x TR scratch block per TB block.
struct InvLayout {
    size_t lb, ub, l128, u128, scratch, total;
    explicit InvLayout(int n) {
        const size_t nbb = size_t(blocks_of(n, TB)), nb1 = size_t(blocks_of(n, TR));
        lb = 0;
        ub = nbb * TB * TB;
        l128 = 2 * ub;
        u128 = l128 + nb1 * TR * TR;
        scratch = u128 + nb1 * TR * TR;
        total = scratch + nbb * TR * TR;
    }
};

// Dst block d (TB x TB, ld TB) = 0 except its diagonal TR-blocks, which take
// the TR-block inverses (nb1 of them; the TB blocks past the matrix's end keep
// zeros): element (r, c) of every TB block in one pass.
__global__ void __launch_bounds__(256) place_diag_kernel(const double* __restrict__ D, double* __restrict__ Dst,
                                                         int nbb, int nb1) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(nbb) * TB * TB) return;
    const int64_t blk = idx / (TB * TB);
    const int r = int(idx % (TB * TB)) / TB, c = int(idx % TB);
    const int64_t d = blk * SUB + r / TR;
    Dst[idx] = (c / TR == r / TR && d < nb1) ? D[d * TR * TR + int64_t(r % TR) * TR + c % TR] : 0.0;
}

// Builds the TB-block inverses of `count` diagonal blocks of size bs (all the
// same; bs <= TB) starting at block b0, by block elimination on the sub-blocks
// (sizes TR, the last one bs - TR * (p - 1)).  Batched over the blocks.
int build_inverses(const double* LU, int64_t lda, int b0, int count, int bs, double* base, const InvLayout& L,
                   hipStream_t s) {
    if (count == 0) return MRM_OK;
    const int p = blocks_of(bs, TR);
    auto sz = [&](int i) { return i < p - 1 ? TR : bs - TR * (p - 1); };
    const int64_t slu = int64_t(TB) * (lda + 1), sinv = int64_t(TB) * TB, s128 = int64_t(SUB) * TR * TR,
                  ssc = int64_t(TR) * TR;
    const double* LUb = LU + int64_t(b0) * slu;
    double* S = base + L.scratch + size_t(b0) * TR * TR;
    int rc = MRM_OK;
    for (int upper = 0; upper < 2 && !rc; ++upper) {
        double* X = base + (upper ? L.ub : L.lb) + size_t(b0) * TB * TB;
        const double* D = base + (upper ? L.u128 : L.l128) + size_t(b0) * SUB * TR * TR;
        for (int jj = 0; jj < p && !rc; ++jj) {
            const int j = upper ? p - 1 - jj : jj;
            for (int ii = 0; ii < (upper ? j : p - 1 - j) && !rc; ++ii) {
                const int i = upper ? j - 1 - ii : j + 1 + ii;
                // S = (L_{i, j..i-1} or U_{i, i+1..j}) * X_{.., j}
                const int k0 = upper ? i + 1 : j, k1 = upper ? j + 1 : i;  // sub-block range of the sum
                int64_t K = 0;
                for (int k = k0; k < k1; ++k) K += sz(k);
                rc = mrm_launch_dgemm_batched(MRM_NN, sz(i), sz(j), K, LUb + int64_t(TR * i) * lda + TR * k0, lda * 8,
                                              slu, X + int64_t(TR * k0) * TB + TR * j, int64_t(TB) * 8, sinv, S,
                                              int64_t(TR) * 8, ssc, count, MRM_DO_NONE, s);
                // X_ij = 0 - D_i S
                if (!rc)
                    rc = mrm_launch_dgemm_batched(MRM_NN, sz(i), sz(j), sz(i), D + int64_t(i) * TR * TR,
                                                  int64_t(TR) * 8, s128, S, int64_t(TR) * 8, ssc,
                                                  X + int64_t(TR * i) * TB + TR * j, int64_t(TB) * 8, sinv, count,
                                                  MRM_DO_SUB, s);
            }
        }
    }
    return rc;
}

}  // namespace

size_t lu_solve_inverse_elems(int n) {
    if (n <= TR) return 0;
    return InvLayout(n).total;
}

int lu_solve_inverses(const double* LU, int64_t lda, int n, double* dinv, hipStream_t s) {
    if (n <= TR) return MRM_OK;
    const InvLayout L(n);
    const int nb1 = blocks_of(n, TR), nbb = blocks_of(n, TB);
    hipLaunchKernelGGL(tri_inv_kernel<false>, dim3(nb1), dim3(TR), 0, s, LU, lda, n, dinv + L.l128);
    hipLaunchKernelGGL(tri_inv_kernel<true>, dim3(nb1), dim3(TR), 0, s, LU, lda, n, dinv + L.u128);
    const unsigned g = unsigned((int64_t(nbb) * TB * TB + 255) / 256);
    hipLaunchKernelGGL(place_diag_kernel, dim3(g), dim3(256), 0, s, dinv + L.l128, dinv + L.lb, nbb, nb1);
    hipLaunchKernelGGL(place_diag_kernel, dim3(g), dim3(256), 0, s, dinv + L.u128, dinv + L.ub, nbb, nb1);
    MRM_HIP_TRY(hipGetLastError());
    const int full = n / TB, rest = n - full * TB;
    int rc = build_inverses(LU, lda, 0, full, TB, dinv, L, s);
    if (!rc && rest > 0) rc = build_inverses(LU, lda, full, 1, rest, dinv, L, s);
    return rc;
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
    const int nbb = blocks_of(n, TB);
    const InvLayout IL(n);
    const int64_t ldy = nrhs;
    const int64_t tn = (nrhs + 127) / 128;
    void* mem = nullptr;
    // Y, the partial products of the split products, and the inverses unless given
    const size_t rhs = size_t(n) * size_t(nrhs);
    // from 256 right-hand sides the products fill the GPU unsplit
    const bool skinny = nrhs < 256;
    const size_t part_elems = skinny ? size_t(SOLVE_MAX_SPLITS) * rhs : 0;
    MRM_HIP_TRY(hipMallocAsync(&mem, (rhs + part_elems + (dinv ? 0 : IL.total)) * sizeof(double), s));
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
    const double* Linv = dinv + IL.lb;
    const double* Uinv = dinv + IL.ub;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned((int64_t(n) * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb, X,
                       ldx, perm, n, nrhs);
    int rc = MRM_OK;
    // The products are skinny (nrhs columns): split K so that about one wave
    // of workgroups (two per CU) runs, chunks of at least 32
    auto mm = [&](int64_t M, int64_t K, const double* Ap, int64_t lda_, const double* Bp, int64_t ldb_, double* Cp,
                  int64_t ldc_, int op) {
        const int64_t tiles = ((M + 127) / 128) * tn;
        int64_t sp = 512 / tiles;
        sp = sp < K / 32 ? sp : K / 32;
        sp = sp < SOLVE_MAX_SPLITS ? sp : SOLVE_MAX_SPLITS;
        sp = (skinny && sp > 1) ? sp : 1;
        return mrm_launch_dgemm_splitk(MRM_NN, M, nrhs, K, Ap, lda_ * 8, Bp, ldb_ * 8, Cp, ldc_ * 8, op, int(sp),
                                       sp > 1 ? part : nullptr, s);
    };
    for (int b = 0; b < nbb && !rc; ++b) {
        const int r0 = b * TB, rb = n - r0 < TB ? n - r0 : TB;
        rc = mm(rb, rb, Linv + size_t(b) * TB * TB, TB, X + int64_t(r0) * ldx, ldx, Y + int64_t(r0) * ldy, ldy,
                MRM_DO_NONE);
        if (!rc && r0 + rb < n)
            rc = mm(n - r0 - rb, rb, LU + int64_t(r0 + rb) * lda + r0, lda, Y + int64_t(r0) * ldy, ldy,
                    X + int64_t(r0 + rb) * ldx, ldx, MRM_DO_SUB);
    }
    for (int b = nbb - 1; b >= 0 && !rc; --b) {
        const int r0 = b * TB, rb = n - r0 < TB ? n - r0 : TB;
        rc = mm(rb, rb, Uinv + size_t(b) * TB * TB, TB, Y + int64_t(r0) * ldy, ldy, X + int64_t(r0) * ldx, ldx,
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
