// mrm_lu.hip -- LU factorisation with partial pivoting on gfx950: the driver,
// row interchanges and the unit-lower triangular blocks.
//
// Replaces the reference's recursive LU (InternalRecursiveMatrixLUDecompInPlace,
// src/LinAlgLU.pas:181-281), its row interchanges (LUSwap, :120-136), the
// unit-lower triangular solve LUBacksup (:138-169) and the Schur update through
// the blocked multiply (:211-227).
//
// Factorisation: right-looking, blocked twice.  Outer blocks of LU_NB = 256
// columns (in every precision) feed the trailing-update GEMM with K = 256;
// each outer block is factored in panels of
// LU_W = 32 columns by the panel kernel (mrm_lu_panel.hip: its workgroups agree
// on every pivot through flagged exchange words), with the reference's pivot
// rule.  From n = LU_LA_MIN_N on, a lookahead overlaps the factorisation of
// block k + 1 (on 32 CUs of their own) with the update of the columns beyond
// it by block k (on the other 224): the panel chain is the critical path, and
// the trailing updates run beside it instead of between its steps.  The pivot sequence is recorded both as indx
// (row i was exchanged with row indx[i] at step i, the reference's format)
// and as a permutation vector perm (row i of P*A is row perm[i] of A), which
// turns the solve's row interchanges (mrm_lusolve.hip) into one parallel
// gather.

#include <cfloat>
#include <climits>
#include <cstdint>
#include <mutex>

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
// perm[r] = the row that ends at position r when the interchanges
// indx[0 .. n-1] are applied in order (as LUSwap does): run them backwards
// from r.  One thread per position; indx passes through LDS in chunks.
constexpr int PERM_CHUNK = 4096;
__global__ void __launch_bounds__(256) perm_from_indx_kernel(const int* __restrict__ indx, int* __restrict__ perm,
                                                             int n) {
    __shared__ int c[PERM_CHUNK];
    const int r = blockIdx.x * 256 + threadIdx.x;
    int pos = r;
    for (int hi = n; hi > 0; hi -= PERM_CHUNK) {
        const int lo = hi > PERM_CHUNK ? hi - PERM_CHUNK : 0;
        __syncthreads();
        for (int i = lo + int(threadIdx.x); i < hi; i += 256) c[i - lo] = indx[i];
        __syncthreads();
        for (int i = hi - 1; i >= lo; --i) {
            const int t = c[i - lo];
            pos = pos == i ? t : pos == t ? i : pos;
        }
    }
    if (r < n) perm[r] = pos;
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

namespace {

// Factors the columns J .. J + jb - 1 (rows J .. n - 1) in panels of LU_W:
// each panel, then its interchanges on the block's other columns, the block
// row of U right of it and the GEMM update of the block's remaining columns.
template <typename T>
int factor_block(T* A, int64_t lda, int n, int J, int jb, int* indx, int* perm, LuWork& ws, hipStream_t s) {
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
    return MRM_OK;
}

// U12 = L11^-1 A12 for the columns c0 .. c0 + cols - 1 right of the factored
// block J .. J + jb - 1, then the Schur update of rows J + jb .. n - 1 of
// those columns: A22 -= L21 U12 (src/LinAlgLU.pas:211-227), in FP64 or, for
// the refined modes, in the low precision.
template <typename T>
int block_row_update(T* A, int64_t lda, int n, int J, int jb, int c0, int cols, int mode, void* work,
                     size_t work_bytes, hipStream_t s) {
    if (cols <= 0) return MRM_OK;
    int rc = trsm_lower_unit<T>(A + int64_t(J) * lda + J, lda, A + int64_t(J) * lda + c0, lda, jb, cols, s);
    if (rc || J + jb >= n) return rc;
    const int rows = n - J - jb;
    if (mode == MRM_SOLVE_FP64)
        return gemm_t<T>(MRM_NN, rows, cols, jb, A + int64_t(J + jb) * lda + J, lda, A + int64_t(J) * lda + c0, lda,
                         A + int64_t(J + jb) * lda + c0, lda, MRM_DO_SUB, s);
    return mrm_launch_lowp_update(mode, rows, cols, jb, A + int64_t(J + jb) * lda + J, lda, A + int64_t(J) * lda + c0,
                                  lda, A + int64_t(J + jb) * lda + c0, lda, work, work_bytes, s);
}

// Bit i of a stream's CU mask selects, in the runtime's order, either CU
// slot i / 8 of XCD i % 8 or CU slot i % 32 of XCD i / 32: the panel mask
// takes bits x, x + 8, ... of the 32-bit word of XCD x, which are
// LU_LA_CUS_PER_XCD CUs of XCD x under either reading, and the update stream
// the complement, so that each stream has CUs in every XCD (workgroups are
// dealt to the XCDs in turn; one dealt to an XCD where its stream has no CU
// would never start).
struct Lookahead {
    int state = 0;   // 0 not tried, 1 ready, -1 not available
    int device = -1; // the device the streams belong to
    hipStream_t panel = nullptr, update = nullptr;
    hipEvent_t ev[7] = {};  // fork, block factored x2, block ready x2, join x2
};
Lookahead g_la;
std::mutex g_la_mu;

void destroy_lookahead() {
    for (hipEvent_t& e : g_la.ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (g_la.panel) (void)hipStreamDestroy(g_la.panel);
    if (g_la.update) (void)hipStreamDestroy(g_la.update);
    g_la.panel = g_la.update = nullptr;
}

}  // namespace

int lu_lookahead_init() {
    static_assert(LU_LA_CUS_PER_XCD >= 1 && LU_LA_CUS_PER_XCD <= 4, "bits x + 8 t of XCD x's word, t < 4");
    std::lock_guard<std::mutex> lk(g_la_mu);
    if (g_la.state) return g_la.state;
    g_la.state = -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus != 256) {
        (void)hipGetLastError();
        return g_la.state;
    }
    uint32_t pm[8] = {}, um[8], got[8];
    for (int x = 0; x < 8; ++x)
        for (int t = 0; t < LU_LA_CUS_PER_XCD; ++t) pm[x] |= 1u << (x + 8 * t);
    for (int w = 0; w < 8; ++w) um[w] = ~pm[w];
    bool ok = hipExtStreamCreateWithCUMask(&g_la.panel, 8, pm) == hipSuccess &&
              hipExtStreamCreateWithCUMask(&g_la.update, 8, um) == hipSuccess;
    ok = ok && hipExtStreamGetCUMask(g_la.panel, 8, got) == hipSuccess;
    for (int w = 0; ok && w < 8; ++w) ok = got[w] == pm[w];
    ok = ok && hipExtStreamGetCUMask(g_la.update, 8, got) == hipSuccess;
    for (int w = 0; ok && w < 8; ++w) ok = got[w] == um[w];
    for (hipEvent_t& e : g_la.ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        destroy_lookahead();
        (void)hipGetLastError();
        return g_la.state;
    }
    g_la.device = dev;
    g_la.state = 1;
    return g_la.state;
}

bool lu_lookahead_usable() {
    if (lu_lookahead_init() != 1) return false;
    int dev = -1;
    return hipGetDevice(&dev) == hipSuccess && dev == g_la.device;
}

void lu_lookahead_release() {
    std::lock_guard<std::mutex> lk(g_la_mu);
    destroy_lookahead();
    g_la.state = 0;
    g_la.device = -1;
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
    int rc = MRM_OK;
    if (!ws.coop && n >= LU_LA_MIN_N && n <= LU_LA_MAX_N && lu_lookahead_usable()) {
        // Lookahead.  Iteration k, block k = columns J .. J1 - 1, block k + 1 =
        // J1 .. J2 - 1:
        //   update stream (224 CUs): wait "block k factored"; block k's
        //       interchanges, U row block and Schur update on block k + 1's
        //       columns; signal "block k + 1 ready"; then the same on every
        //       column left of block k and from J2 on
        //   panel stream (32 CUs): wait "block k + 1 ready"; factor block
        //       k + 1; signal "block k + 1 factored"
        // The streams never write the same columns between two signals, and
        // every column receives the sequential schedule's operations in its
        // order (only the column split of the updates differs).
        hipStream_t P = g_la.panel, U = g_la.update;
        hipEvent_t* ev = g_la.ev;  // 0 fork, 1-2 block factored, 3-4 block ready, 5-6 join
        ws.panel_cus = 8 * LU_LA_CUS_PER_XCD;
        // signal: record e on `from`, make `to` wait for it
        auto sig = [&rc](hipEvent_t e, hipStream_t from, hipStream_t to) {
            hipError_t r = hipEventRecord(e, from);
            if (r == hipSuccess) r = hipStreamWaitEvent(to, e, 0);
            if (r != hipSuccess && !rc) {
                mrm_set_error(MRM_HIP_ERROR, "lu lookahead: stream signal: %s", hipGetErrorString(r));
                rc = MRM_HIP_ERROR;
            }
            return rc == MRM_OK;
        };
        if (sig(ev[0], s, P) && sig(ev[0], s, U))
            rc = factor_block<T>(A, lda, n, 0, n < nb ? n : nb, indx, perm, ws, P);
        for (int J = 0, k = 0; !rc && J < n; J += nb, ++k) {
            const int jb = n - J < nb ? n - J : nb;
            const int J1 = J + jb, jb1 = n - J1 < nb ? n - J1 : nb, J2 = J1 + jb1;
            if (!sig(ev[1 + (k & 1)], P, U)) break;  // block k factored
            if (J1 >= n) {
                laswp<T>(A, lda, J, J1, indx, 0, J, n, n, U);
                break;
            }
            laswp<T>(A, lda, J, J1, indx, J1, J2, n, n, U);
            if ((rc = block_row_update<T>(A, lda, n, J, jb, J1, jb1, mode, ws.lowp, ws.lowp_bytes, U))) break;
            if (!sig(ev[3 + (k & 1)], U, P)) break;  // block k + 1 ready
            if ((rc = factor_block<T>(A, lda, n, J1, jb1, indx, perm, ws, P))) break;
            laswp<T>(A, lda, J, J1, indx, 0, J, J2, n, U);
            if ((rc = block_row_update<T>(A, lda, n, J, jb, J2, n - J2, mode, ws.lowp, ws.lowp_bytes, U))) break;
        }
        ws.panel_cus = 0;
        // join, also after an error: nothing of this factorisation may run
        // after the caller's stream has moved on
        sig(ev[5], P, s);
        sig(ev[6], U, s);
        if (rc) return rc;
    } else {
        for (int J = 0; J < n; J += nb) {
            const int jb = n - J < nb ? n - J : nb;
            if ((rc = factor_block<T>(A, lda, n, J, jb, indx, perm, ws, s))) return rc;
            laswp<T>(A, lda, J, J + jb, indx, 0, J, J + jb, n, s);
            if ((rc = block_row_update<T>(A, lda, n, J, jb, J + jb, n - J - jb, mode, ws.lowp, ws.lowp_bytes, s)))
                return rc;
        }
    }
    if (perm && n > 0)
        hipLaunchKernelGGL(perm_from_indx_kernel, dim3((n + 255) / 256), dim3(256), 0, s, indx, perm, n);
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
