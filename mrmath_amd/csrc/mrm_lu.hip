// mrm_lu.hip -- LU factorisation with partial pivoting on gfx950: the driver,
// row interchanges and the unit-lower triangular blocks.
//
// Replaces the reference's recursive LU (InternalRecursiveMatrixLUDecompInPlace,
// src/LinAlgLU.pas:181-281), its row interchanges (LUSwap, :120-136), the
// unit-lower triangular solve LUBacksup (:138-169) and the Schur update through
// the blocked multiply (:211-227).
//
// Factorisation: right-looking, blocked twice.  Outer blocks of LU_NB = 512
// columns (in every precision) feed the trailing-update GEMM with K = 512
// (n = 16384: 172 ms against 175 ms with 256, whose 16-step K loops left the
// tile prologue and the read-modify-write of C a larger share); each outer
// block is factored in panels of
// LU_W = 32 columns by the panel kernel (mrm_lu_panel.hip: its workgroups agree
// on every pivot through flagged exchange words), with the reference's pivot
// rule.  The pivot sequence is recorded both as indx
// (row i was exchanged with row indx[i] at step i, the reference's format)
// and as a permutation vector perm (row i of P*A is row perm[i] of A), which
// turns the solve's row interchanges (mrm_lusolve.hip) into one parallel
// gather.

#include <cfloat>
#include <climits>
#include <cstdint>

#include "mrm_common.h"
#include "mrm_lowp.h"
#include "mrm_lu.h"

namespace {

__global__ void zero_u32_kernel(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// perm from the reference's sequential pivot list: row i of P*A is row perm[i]
// of A.  One thread replays the n exchanges on a copy held in LDS (dynamic
// shared memory of n ints) or, for larger n, in global memory.
__global__ void perm_from_indx_kernel(const int* __restrict__ indx, int* __restrict__ perm, int n, int in_lds) {
    extern __shared__ int sp[];
    int* p = in_lds ? sp : perm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) {
            const int r = indx[i];
            if (r != i) {
                const int a = p[i];
                p[i] = p[r];
                p[r] = a;
            }
        }
    }
    __syncthreads();
    if (in_lds)
        for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = sp[i];
}

__global__ void iota_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Row interchanges of rows k1..k2-1 (sequentially, as LUSwap does) over the
// column ranges [c0, c1) and [c2, c3), applied as one gather.  The rows that
// can change are D = [k1, k2) and every indx[k] outside it; the row that ends
// at position r is found by running the interchanges backwards from r.  Each
// workgroup takes SW_COLS columns: it builds that (destination, source) list
// in LDS, loads every source segment into registers, and after a barrier
// stores them, so the data movement is a handful of independent wide accesses
// instead of a chain of dependent swaps.  Same result as the sequential
// interchanges for any indx; at most SW_MAX_ROWS interchanges per launch.
constexpr int SW_COLS = 16;
constexpr int SW_THREADS = 256;
constexpr int SW_MAX_ROWS = mrm::LU_NB_LOWP;
constexpr int SW_PER = 2 * SW_MAX_ROWS / (SW_THREADS / SW_COLS);  // entries per thread

template <typename T>
__global__ void __launch_bounds__(SW_THREADS)
laswp_kernel(T* __restrict__ A, int64_t lda, int k1, int k2, const int* __restrict__ indx, int c0, int c1, int c2,
             int c3) {
    __shared__ int piv[SW_MAX_ROWS], src[2 * SW_MAX_ROWS], dst[2 * SW_MAX_ROWS];
    const int kb = k2 - k1;
    const int ne = 2 * kb;
    const int t = threadIdx.x;
    for (int k = t; k < kb; k += SW_THREADS) piv[k] = indx[k1 + k];
    __syncthreads();
    for (int e = t; e < ne; e += SW_THREADS) {
        int r = -1;
        if (e < kb) r = k1 + e;
        else {
            const int p = piv[e - kb];
            if (p < k1 || p >= k2) r = p;
        }
        int pos = r;
        if (r >= 0)
            for (int k = kb - 1; k >= 0; --k) {
                const int p = piv[k], q = k1 + k;
                pos = pos == q ? p : (pos == p ? q : pos);
            }
        const bool moves = r >= 0 && pos != r;
        dst[e] = moves ? r : -1;
        src[e] = pos;
    }
    __syncthreads();
    const int n1 = c1 - c0, ncols = n1 + (c3 - c2);
    const int ci = blockIdx.x * SW_COLS + t % SW_COLS;
    const bool live = ci < ncols;
    const int col = ci < n1 ? c0 + ci : c2 + (ci - n1);
    constexpr int RL = SW_THREADS / SW_COLS;
    T v[SW_PER];
#pragma unroll
    for (int i = 0; i < SW_PER; ++i) {
        const int e = t / SW_COLS + i * RL;
        if (e < ne && live && dst[e] >= 0) v[i] = A[int64_t(src[e]) * lda + col];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SW_PER; ++i) {
        const int e = t / SW_COLS + i * RL;
        if (e < ne && live && dst[e] >= 0) A[int64_t(dst[e]) * lda + col] = v[i];
    }
}

// X = L^-1 X for the R x ncols block X (row stride ldx), L unit lower R x R
// (row stride ldl); one thread per column, the column in registers.
template <typename T, int R>
__global__ void __launch_bounds__(128)
trsm_lunit_block_kernel(const T* __restrict__ L, int64_t ldl, T* __restrict__ X, int64_t ldx, int rows, int ncols) {
    __shared__ T Ls[R][R + 1];
    for (int e = threadIdx.x; e < rows * rows; e += blockDim.x) Ls[e / rows][e % rows] = L[int64_t(e / rows) * ldl + e % rows];
    __syncthreads();
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols) return;
    T x[R];
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = i < rows ? X[int64_t(i) * ldx + col] : T(0);
#pragma unroll
    for (int i = 1; i < R; ++i) {
        if (i < rows) {
            T s = x[i];
#pragma unroll
            for (int k = 0; k < i; ++k) s = fma(-Ls[i][k], x[k], s);
            x[i] = s;
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (i < rows) X[int64_t(i) * ldx + col] = x[i];
}

// X += D (rows x cols, element strides)
template <typename T>
__global__ void add_inplace_kernel(T* __restrict__ X, int64_t ldx, const T* __restrict__ D, int64_t ldd, int rows,
                                   int cols) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= int64_t(rows) * cols) return;
    const int i = int(idx / cols), c = int(idx % cols);
    X[int64_t(i) * ldx + c] += D[int64_t(i) * ldd + c];
}

__global__ void zero_i32_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int grid1(int64_t n, int b) { return int((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// Host-side drivers (device pointers, element strides).
// ---------------------------------------------------------------------------
namespace mrm {

template <typename T>
void laswp(T* A, int64_t lda, int k1, int k2, const int* indx, int c0, int c1, int c2, int c3, hipStream_t s) {
    const int ncols = (c1 - c0) + (c3 - c2);
    if (ncols <= 0 || k2 <= k1) return;
    // callers pass at most one outer block of interchanges
    for (int k = k1; k < k2; k += SW_MAX_ROWS) {
        const int ke = k2 - k < SW_MAX_ROWS ? k2 : k + SW_MAX_ROWS;
        hipLaunchKernelGGL(laswp_kernel<T>, dim3(grid1(ncols, SW_COLS)), dim3(SW_THREADS), 0, s, A, lda, k, ke, indx,
                           c0, c1, c2, c3);
    }
}

// X (rows x ncols) = L^-1 X, L unit lower rows x rows; blocks of LU_TRSM_R rows,
// the off-diagonal part through the GEMM.
template <typename T>
int trsm_lower_unit(const T* L, int64_t ldl, T* X, int64_t ldx, int rows, int ncols, hipStream_t s) {
    constexpr int R = LU_TRSM_R;
    for (int b = 0; b < rows; b += R) {
        const int rb = rows - b < R ? rows - b : R;
        hipLaunchKernelGGL((trsm_lunit_block_kernel<T, R>), dim3(grid1(ncols, 128)), dim3(128), 0, s,
                           L + int64_t(b) * ldl + b, ldl, X + int64_t(b) * ldx, ldx, rb, ncols);
        if (b + rb < rows) {
            int rc = gemm_t<T>(MRM_NN, rows - b - rb, ncols, rb, L + int64_t(b + rb) * ldl + b, ldl,
                               X + int64_t(b) * ldx, ldx, X + int64_t(b + rb) * ldx, ldx, MRM_DO_SUB, s);
            if (rc) return rc;
        }
    }
    return MRM_OK;
}

template <typename T>
int lu_factor(T* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s, int mode) {
    // indx starts at 0 as in the reference (src/LinAlgLU.pas:289); entries a
    // singular stop leaves unset stay in range for the row interchanges
    zero_i32(indx, n, s);
    // no flag of an earlier factorisation may survive in the exchange area
    const int xw = int(lu_xch_bytes() / 4);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid1(xw, 256)), dim3(256), 0, s, static_cast<unsigned*>(ws.xch), xw);
    const int nb = lu_outer_block(mode);
    for (int J = 0; J < n; J += nb) {
        const int jb = n - J < nb ? n - J : nb;
        for (int j = J; j < J + jb; j += LU_W) {
            const int wb = J + jb - j < LU_W ? J + jb - j : LU_W;
            int rc = lu_panel<T>(A + int64_t(j) * lda + j, lda, n - j, wb, j, indx, perm, ws, s);
            if (rc) return rc;
            laswp<T>(A, lda, j, j + wb, indx, J, j, j + wb, J + jb, s);
            if (j + wb < J + jb) {
                rc = trsm_lower_unit<T>(A + int64_t(j) * lda + j, lda, A + int64_t(j) * lda + j + wb, lda, wb,
                                        J + jb - j - wb, s);
                if (rc) return rc;
                if (j + wb < n) {
                    rc = gemm_t<T>(MRM_NN, n - j - wb, J + jb - j - wb, wb, A + int64_t(j + wb) * lda + j, lda,
                                   A + int64_t(j) * lda + j + wb, lda, A + int64_t(j + wb) * lda + j + wb, lda,
                                   MRM_DO_SUB, s);
                    if (rc) return rc;
                }
            }
        }
        laswp<T>(A, lda, J, J + jb, indx, 0, J, J + jb, n, s);
        if (J + jb < n) {
            int rc = trsm_lower_unit<T>(A + int64_t(J) * lda + J, lda, A + int64_t(J) * lda + J + jb, lda, jb,
                                        n - J - jb, s);
            if (rc) return rc;
            if (mode == MRM_SOLVE_FP64)
                rc = gemm_t<T>(MRM_NN, n - J - jb, n - J - jb, jb, A + int64_t(J + jb) * lda + J, lda,
                               A + int64_t(J) * lda + J + jb, lda, A + int64_t(J + jb) * lda + J + jb, lda,
                               MRM_DO_SUB, s);
            else
                rc = mrm_launch_lowp_update(mode, n - J - jb, n - J - jb, jb, A + int64_t(J + jb) * lda + J, lda,
                                            A + int64_t(J) * lda + J + jb, lda, A + int64_t(J + jb) * lda + J + jb,
                                            lda, ws.lowp, ws.lowp_bytes, s);
            if (rc) return rc;
        }
    }
    if (perm) {
        // the replay's copy in LDS up to 64 KiB (16384 rows), in global memory beyond
        const int in_lds = n <= 16384;
        hipLaunchKernelGGL(perm_from_indx_kernel, dim3(1), dim3(256), in_lds ? size_t(n) * sizeof(int) : 0, s, indx,
                           perm, n, in_lds);
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

template int lu_factor<double>(double*, int64_t, int, int*, int*, LuWork&, hipStream_t, int);

size_t lu_lowp_bytes(int n, int mode) {
    const int nb = lu_outer_block(mode);
    if (mode == MRM_SOLVE_FP64 || n <= nb) return 0;
    return mrm_lowp_workspace(n - nb, n - nb, nb, mode);
}

int lu_lowp_reserve(LuWork& ws, int n, int mode) {
    const size_t need = lu_lowp_bytes(n, mode);
    if (need <= ws.lowp_bytes) return MRM_OK;
    if (ws.lowp) MRM_HIP_TRY(hipFree(ws.lowp));
    ws.lowp = nullptr;
    ws.lowp_bytes = 0;
    if (hipMalloc(&ws.lowp, need) != hipSuccess) {
        mrm_set_error(MRM_NO_MEMORY, "lu: cannot allocate %zu bytes of packed-operand workspace", need);
        return MRM_NO_MEMORY;
    }
    ws.lowp_bytes = need;
    return MRM_OK;
}

template <typename T>
void add_inplace(T* X, int64_t ldx, const T* D, int64_t ldd, int rows, int cols, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(add_inplace_kernel<T>, dim3(grid1(int64_t(rows) * cols, 256)), dim3(256), 0, s, X, ldx, D, ldd,
                       rows, cols);
}
template void add_inplace<double>(double*, int64_t, const double*, int64_t, int, int, hipStream_t);

void zero_i32(int* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(zero_i32_kernel, dim3(grid1(n, 256)), dim3(256), 0, s, p, n);
}

}  // namespace mrm
