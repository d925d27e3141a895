// mrm_lu.hip -- LU factorisation with partial pivoting and triangular solves on gfx950.
//
// Replaces the reference's recursive LU (InternalRecursiveMatrixLUDecompInPlace,
// src/LinAlgLU.pas:181-281), its row interchanges (LUSwap, :120-136), the
// unit-lower triangular solve LUBacksup (:138-169), the Schur update through the
// blocked multiply (:211-227) and the substitutions of MatrixLUBackSubst
// (:315-385).
//
// Factorisation: right-looking, blocked twice.  Outer blocks of LU_NB = 256
// columns feed the FP64 MFMA GEMM with K = 256 for the trailing update; each
// outer block is factored in panels of LU_W = 32 columns by one cooperative
// kernel that keeps the panel's rows in registers, spread over as many
// workgroups as the panel height needs, and takes one grid-wide barrier per
// column to agree on the pivot.  The pivot rule is the reference's: the first
// row (in the current row order) of maximal |a| in the column, singular when
// |pivot| <= 10/MaxDouble; the column is scaled by the reciprocal of the pivot.
// The pivot sequence is recorded both as indx (row i was exchanged with row
// indx[i] at step i, the reference's format) and as a permutation vector
// perm (row i of P*A is row perm[i] of A), which turns the solve's row
// interchanges into one parallel gather.
//
// The arithmetic type is a template parameter: double for the FP64 path, float
// for the low-precision factorisation of the mixed-precision solver.

#include <cfloat>
#include <climits>
#include <cstdint>

#include "mrm_common.h"
#include "mrm_lu.h"

namespace {

constexpr int PANEL_THREADS = 256;

template <typename T>
struct Lim;
template <>
struct Lim<double> {
    // src/LinAlgLU.pas:262: abs(maxVal) > 10/MaxDouble
    static __device__ __forceinline__ double tiny() { return 10.0 / DBL_MAX; }
};
template <>
struct Lim<float> {
    static __device__ __forceinline__ float tiny() { return 10.0f / FLT_MAX; }
};

// Larger |value| wins; on equal values the smaller row index (earlier row) wins.
template <typename T>
__device__ __forceinline__ bool better(T va, int ra, T vb, int rb) {
    return va > vb || (va == vb && ra < rb);
}

template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(v, off, 64);
        const int orow = __shfl_xor(r, off, 64);
        if (better(ov, orow, v, r)) {
            v = ov;
            r = orow;
        }
    }
}

// Grid-wide barrier for a cooperative launch: monotonic arrival counter.
// Every workgroup leaves after `deadline` (wall clock, 100 MHz) at the latest
// and then reports MRM_LU_TIMEOUT, so no wave can spin for ever.
__device__ __forceinline__ bool grid_barrier(unsigned* counter, unsigned target, uint64_t deadline,
                                             int* status) {
    __syncthreads();
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        atomicAdd(counter, 1u);
        while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (wall_clock64() > deadline) {
                atomicExch(status, MRM_LU_TIMEOUT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <typename T, int W>
struct PanelShared {
    T prow[W];                  // pivot row of the current column
    T drow[W];                  // row j before the exchange
    T wv[PANEL_THREADS / 64];   // per-wave reduction scratch
    int wr[PANEL_THREADS / 64];
    int ww[PANEL_THREADS / 64];
    int piv, pwg, sing;
};

// Exchange area of one column parity: [G values][G rows][G*W row data][W diag row].
template <typename T, int W>
struct Exchange {
    T* val;
    int* row;
    T* data;
    T* diag;
    __device__ Exchange(T* xch, int G, int par) {
        const size_t per = size_t(G) * (W + 2) + W;
        T* base = xch + par * per;
        val = base;
        row = reinterpret_cast<int*>(base + G);
        data = base + 2 * G;
        diag = base + 2 * G + size_t(G) * W;
    }
};

// One column j of the panel (j is a compile-time constant after unrolling, so
// v stays in registers).  Returns false when the grid stops (singular column or
// barrier timeout).
template <typename T, int W, int RPT>
__device__ __forceinline__ bool panel_column(const int j, T (&v)[RPT][W], const int (&lr)[RPT], int m, int g, int G,
                                             int t, int lane, int wave, int piv_base, int* __restrict__ indx,
                                             int* __restrict__ perm, T* __restrict__ xch,
                                             unsigned* __restrict__ counter, int* __restrict__ status,
                                             uint64_t deadline, PanelShared<T, W>& sh) {
    Exchange<T, W> x(xch, G, j & 1);
    // (a) best candidate of this workgroup among rows j..m-1
    T best = T(-1);
    int brow = INT_MAX;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] >= j && lr[r] < m) {
            const T a = fabs(v[r][j]);
            if (better(a, lr[r], best, brow)) {
                best = a;
                brow = lr[r];
            }
        }
    }
    wave_argmax(best, brow);
    if (lane == 0) {
        sh.wv[wave] = best;
        sh.wr[wave] = brow;
    }
    __syncthreads();
    if (t == 0) {
        for (int q = 1; q < PANEL_THREADS / 64; ++q)
            if (better(sh.wv[q], sh.wr[q], best, brow)) {
                best = sh.wv[q];
                brow = sh.wr[q];
            }
        sh.piv = brow;
        __hip_atomic_store(&x.val[g], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&x.row[g], brow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int wgrow = sh.piv;
    // (b) publish the candidate row and, from its owner, the current row j
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] == wgrow) {
#pragma unroll
            for (int c = 0; c < W; ++c)
                __hip_atomic_store(&x.data[size_t(g) * W + c], v[r][c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lr[r] == j) {
#pragma unroll
            for (int c = 0; c < W; ++c)
                __hip_atomic_store(&x.diag[c], v[r][c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // (c) all workgroups agree
    if (!grid_barrier(counter, unsigned(G) * unsigned(j + 1), deadline, status)) return false;
    // (d) global pivot: max |value|, earliest row on ties; carry the owning workgroup
    T gb = T(-1);
    int gr = INT_MAX, gw = 0;
    for (int q = t; q < G; q += PANEL_THREADS) {
        const T cv = __hip_atomic_load(&x.val[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int cr = __hip_atomic_load(&x.row[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (better(cv, cr, gb, gr)) {
            gb = cv;
            gr = cr;
            gw = q;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(gb, off, 64);
        const int orow = __shfl_xor(gr, off, 64);
        const int ow = __shfl_xor(gw, off, 64);
        if (better(ov, orow, gb, gr)) {
            gb = ov;
            gr = orow;
            gw = ow;
        }
    }
    if (lane == 0) {
        sh.wv[wave] = gb;
        sh.wr[wave] = gr;
        sh.ww[wave] = gw;
    }
    __syncthreads();
    if (t == 0) {
        T b0 = sh.wv[0];
        int r0 = sh.wr[0], w0 = sh.ww[0];
        for (int q = 1; q < PANEL_THREADS / 64; ++q)
            if (better(sh.wv[q], sh.wr[q], b0, r0)) {
                b0 = sh.wv[q];
                r0 = sh.wr[q];
                w0 = sh.ww[q];
            }
        sh.piv = r0;
        sh.pwg = w0;
        sh.sing = !(b0 > Lim<T>::tiny());
    }
    __syncthreads();
    if (sh.sing) {
        if (g == 0 && t == 0) atomicExch(status, MRM_SINGULAR);
        return false;
    }
    const int prow = sh.piv;
    if (t < W) {
        sh.prow[t] = __hip_atomic_load(&x.data[size_t(sh.pwg) * W + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.drow[t] = __hip_atomic_load(&x.diag[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (g == 0 && t == 0) {
        indx[piv_base + j] = piv_base + prow;
        if (perm && prow != j) {
            const int a = perm[piv_base + j];
            perm[piv_base + j] = perm[piv_base + prow];
            perm[piv_base + prow] = a;
        }
    }
    // (e) exchange rows j and prow
    if (prow != j) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            if (lr[r] == prow) {
#pragma unroll
                for (int c = 0; c < W; ++c) v[r][c] = sh.drow[c];
            } else if (lr[r] == j) {
#pragma unroll
                for (int c = 0; c < W; ++c) v[r][c] = sh.prow[c];
            }
        }
    }
    // (f) scale by the reciprocal pivot (src/LinAlgLU.pas:264) and update the panel row
    const T inv = T(1) / sh.prow[j];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] > j && lr[r] < m) {
            const T l = v[r][j] * inv;
            v[r][j] = l;
#pragma unroll
            for (int c = j + 1; c < W; ++c) v[r][c] = fma(-l, sh.prow[c], v[r][c]);
        }
    }
    return true;
}

// Columns J..W-1 by compile-time recursion, so every index into v is a constant.
template <typename T, int W, int RPT, int J>
struct PanelColumns {
    static __device__ __forceinline__ void run(int w, bool& alive, T (&v)[RPT][W], const int (&lr)[RPT], int m, int g,
                                               int G, int t, int lane, int wave, int piv_base, int* indx, int* perm,
                                               T* xch, unsigned* counter, int* status, uint64_t deadline,
                                               PanelShared<T, W>& sh) {
        if (J < w && alive)
            alive = panel_column<T, W, RPT>(J, v, lr, m, g, G, t, lane, wave, piv_base, indx, perm, xch, counter,
                                            status, deadline, sh);
        PanelColumns<T, W, RPT, J + 1>::run(w, alive, v, lr, m, g, G, t, lane, wave, piv_base, indx, perm, xch,
                                            counter, status, deadline, sh);
    }
};
template <typename T, int W, int RPT>
struct PanelColumns<T, W, RPT, W> {
    static __device__ __forceinline__ void run(int, bool&, T (&)[RPT][W], const int (&)[RPT], int, int, int, int, int,
                                               int, int, int*, int*, T*, unsigned*, int*, uint64_t,
                                               PanelShared<T, W>&) {}
};

// Factor the m x w panel at P (row stride lda elements) in place, w <= W.
// Workgroup g owns panel rows [g*RPT*256, (g+1)*RPT*256); thread t owns rows
// g*RPT*256 + r*256 + t.  Pivot rows are absolute: piv_base + local row.
template <typename T, int W, int RPT>
__global__ void __launch_bounds__(PANEL_THREADS)
lu_panel_kernel(T* __restrict__ P, int64_t lda, int m, int w, int piv_base, int* __restrict__ indx,
                int* __restrict__ perm, T* __restrict__ xch, unsigned* __restrict__ counter,
                int* __restrict__ status, uint64_t timeout_ticks) {
    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    __shared__ PanelShared<T, W> sh;
    // an earlier panel found the matrix singular: nothing left to do (uniform)
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    T v[RPT][W];
    int lr[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        lr[r] = g * RPT * PANEL_THREADS + r * PANEL_THREADS + t;
        const bool ok = lr[r] < m;
        const T* src = P + int64_t(ok ? lr[r] : 0) * lda;
#pragma unroll
        for (int c = 0; c < W; ++c) v[r][c] = (ok && c < w) ? src[c] : T(0);
    }
    const uint64_t deadline = wall_clock64() + timeout_ticks;
    bool alive = true;

    PanelColumns<T, W, RPT, 0>::run(w, alive, v, lr, m, g, G, t, lane, wave, piv_base, indx, perm, xch, counter,
                                    status, deadline, sh);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] < m) {
            T* dst = P + int64_t(lr[r]) * lda;
#pragma unroll
            for (int c = 0; c < W; ++c)
                if (c < w) dst[c] = v[r][c];
        }
    }
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void iota_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Row interchanges of rows k1..k2-1 (sequentially, as LUSwap does) over the
// column ranges [c0, c1) and [c2, c3); one thread per column.
template <typename T>
__global__ void laswp_kernel(T* __restrict__ A, int64_t lda, int k1, int k2, const int* __restrict__ indx, int c0,
                             int c1, int c2, int c3) {
    const int n1 = c1 - c0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int col;
    if (i < n1) col = c0 + i;
    else if (i - n1 < c3 - c2) col = c2 + (i - n1);
    else return;
    for (int k = k1; k < k2; ++k) {
        const int p = indx[k];
        if (p != k) {
            T* a = A + int64_t(k) * lda + col;
            T* b = A + int64_t(p) * lda + col;
            const T x = *a;
            *a = *b;
            *b = x;
        }
    }
}

// X = L^-1 X for the R x ncols block X (row stride ldx), L unit lower R x R
// (row stride ldl); one thread per column, the column in registers.
template <typename T, int R>
__global__ void trsm_lunit_block_kernel(const T* __restrict__ L, int64_t ldl, T* __restrict__ X, int64_t ldx,
                                        int rows, int ncols) {
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
            for (int k = 0; k < i; ++k) s = fma(-L[int64_t(i) * ldl + k], x[k], s);
            x[i] = s;
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (i < rows) X[int64_t(i) * ldx + col] = x[i];
}

// X = U^-1 X for the R x ncols block, U upper with its diagonal (divides, as
// MatrixLUBackSubst does, src/LinAlgLU.pas:362-384).
template <typename T, int R>
__global__ void trsm_upper_block_kernel(const T* __restrict__ U, int64_t ldu, T* __restrict__ X, int64_t ldx,
                                        int rows, int ncols) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols) return;
    T x[R];
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = i < rows ? X[int64_t(i) * ldx + col] : T(0);
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {
        if (i < rows) {
            T s = x[i];
#pragma unroll
            for (int k = i + 1; k < R; ++k)
                if (k < rows) s = fma(-U[int64_t(i) * ldu + k], x[k], s);
            x[i] = s / U[int64_t(i) * ldu + i];
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (i < rows) X[int64_t(i) * ldx + col] = x[i];
}

// Y[i][c] = B[perm[i]][c]
template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ B, int64_t ldb, T* __restrict__ Y, int64_t ldy,
                                   const int* __restrict__ perm, int n, int ncols) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= int64_t(n) * ncols) return;
    const int i = int(idx / ncols), c = int(idx % ncols);
    Y[int64_t(i) * ldy + c] = B[int64_t(perm[i]) * ldb + c];
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
int lu_panel(T* P, int64_t lda, int m, int w, int piv_base, int* indx, int* perm, LuWork& ws, hipStream_t s) {
    constexpr int W = LU_W;
    int rpt = 1;
    int G = grid1(m, PANEL_THREADS);
    if (G > LU_MAX_PANEL_WG) {
        rpt = 2;
        G = grid1(m, 2 * PANEL_THREADS);
    }
    if (G > LU_MAX_PANEL_WG) {
        mrm_set_error(MRM_BAD_DIMS, "lu: matrix too tall for the panel kernel (%d rows)", m);
        return MRM_BAD_DIMS;
    }
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, ws.counter, 1);
    T* xch = reinterpret_cast<T*>(ws.xch);
    uint64_t timeout = 2ull * 100000000ull;  // 2 s of wall clock (100 MHz)
    void* args[] = {&P, &lda, &m, &w, &piv_base, &indx, &perm, &xch, &ws.counter, &ws.status, &timeout};
    hipError_t e;
    if (rpt == 1)
        e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&lu_panel_kernel<T, W, 1>), dim3(G),
                                       dim3(PANEL_THREADS), args, 0, s);
    else
        e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&lu_panel_kernel<T, W, 2>), dim3(G),
                                       dim3(PANEL_THREADS), args, 0, s);
    if (e != hipSuccess) {
        mrm_set_error(MRM_HIP_ERROR, "lu panel launch (G=%d): %s", G, hipGetErrorString(e));
        return MRM_HIP_ERROR;
    }
    return MRM_OK;
}

template <typename T>
void laswp(T* A, int64_t lda, int k1, int k2, const int* indx, int c0, int c1, int c2, int c3, hipStream_t s) {
    const int ncols = (c1 - c0) + (c3 - c2);
    if (ncols <= 0 || k2 <= k1) return;
    hipLaunchKernelGGL(laswp_kernel<T>, dim3(grid1(ncols, 256)), dim3(256), 0, s, A, lda, k1, k2, indx, c0, c1,
                       c2, c3);
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
int trsm_upper(const T* U, int64_t ldu, T* X, int64_t ldx, int rows, int ncols, hipStream_t s) {
    constexpr int R = LU_TRSM_R;
    // bottom block first
    int last = ((rows - 1) / R) * R;
    for (int b = last; b >= 0; b -= R) {
        const int rb = rows - b < R ? rows - b : R;
        hipLaunchKernelGGL((trsm_upper_block_kernel<T, R>), dim3(grid1(ncols, 128)), dim3(128), 0, s,
                           U + int64_t(b) * ldu + b, ldu, X + int64_t(b) * ldx, ldx, rb, ncols);
        if (b > 0) {
            // rows above: X[0:b] -= U[0:b, b:b+rb] * X[b:b+rb]
            int rc = gemm_t<T>(MRM_NN, b, ncols, rb, U + b, ldu, X + int64_t(b) * ldx, ldx, X, ldx, MRM_DO_SUB, s);
            if (rc) return rc;
        }
    }
    return MRM_OK;
}

template <typename T>
int lu_factor(T* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s) {
    // indx starts at 0 as in the reference (src/LinAlgLU.pas:289); entries a
    // singular stop leaves unset stay in range for the row interchanges
    zero_i32(indx, n, s);
    if (perm) hipLaunchKernelGGL(iota_kernel, dim3(grid1(n, 256)), dim3(256), 0, s, perm, n);
    for (int J = 0; J < n; J += LU_NB) {
        const int jb = n - J < LU_NB ? n - J : LU_NB;
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
            rc = gemm_t<T>(MRM_NN, n - J - jb, n - J - jb, jb, A + int64_t(J + jb) * lda + J, lda,
                           A + int64_t(J) * lda + J + jb, lda, A + int64_t(J + jb) * lda + J + jb, lda, MRM_DO_SUB,
                           s);
            if (rc) return rc;
        }
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

// X (n x nrhs) = (LU)^-1 P B, with P given by perm.
template <typename T>
int lu_solve(const T* LU, int64_t lda, int n, const int* perm, const T* B, int64_t ldb, T* X, int64_t ldx, int nrhs,
             hipStream_t s) {
    if (n == 0 || nrhs == 0) return MRM_OK;
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(grid1(int64_t(n) * nrhs, 256)), dim3(256), 0, s, B, ldb, X, ldx,
                       perm, n, nrhs);
    int rc = trsm_lower_unit<T>(LU, lda, X, ldx, n, nrhs, s);
    if (rc) return rc;
    rc = trsm_upper<T>(LU, lda, X, ldx, n, nrhs, s);
    if (rc) return rc;
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

template int lu_factor<double>(double*, int64_t, int, int*, int*, LuWork&, hipStream_t);
template int lu_solve<double>(const double*, int64_t, int, const int*, const double*, int64_t, double*, int64_t, int,
                              hipStream_t);

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

size_t lu_xch_bytes() {
    // two parities of [G values][G rows][G*W data][W diag], doubles, G <= LU_MAX_PANEL_WG
    return 2 * (size_t(LU_MAX_PANEL_WG) * (LU_W + 2) + LU_W) * sizeof(double);
}

}  // namespace mrm
