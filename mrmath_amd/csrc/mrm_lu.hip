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
// workgroups as the panel height needs; per column every workgroup publishes
// its candidate in flagged words and polls everybody else's (one memory round
// trip, no grid barrier) to agree on the pivot.  The pivot rule is the reference's: the first
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
#include "mrm_lowp.h"
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

// Cross-workgroup exchange of the panel: every 8-byte word written carries 32
// bits of data and a 32-bit flag naming the column it belongs to, so a reader
// polls the words it needs until their flags match and needs no barrier, no
// counter and no cache maintenance: one memory round trip per column.  Aligned
// 8-byte stores and loads are single-copy atomic; the words are relaxed
// agent-scope atomics (coherent across the XCDs).  Two parities alternate by
// column: a workgroup can only publish column j+2 after every workgroup has
// published j+1, which each did after reading all of column j.
template <typename T>
struct Unit {
    static constexpr int per = int(sizeof(T) / 4);  // 32-bit units per value
    static __device__ __forceinline__ uint32_t part(T v, int h) {
        if constexpr (sizeof(T) == 8) {
            const uint64_t b = uint64_t(__double_as_longlong(v));
            return uint32_t(h ? (b >> 32) : b);
        } else {
            return __float_as_uint(v);
        }
    }
    static __device__ __forceinline__ T value(const uint32_t* u) {
        if constexpr (sizeof(T) == 8) return __longlong_as_double((long long)((uint64_t(u[1]) << 32) | u[0]));
        else return __uint_as_float(u[0]);
    }
};

// Record of one workgroup for one column: [|candidate| (per units)][row]
// [candidate row (W values)]; after the G records, row j (W values) from the
// workgroup that owns it.
template <typename T, int W>
struct Rec {
    static constexpr int UPV = Unit<T>::per;
    static constexpr int RS = UPV + 1 + W * UPV;   // units per record
    static constexpr int DU = W * UPV;             // units of row j
    static __host__ __device__ constexpr size_t parity_units(int G) { return size_t(G) * RS + DU; }
};

template <typename T, int W>
struct PanelShared {
    T mine[W];                  // this workgroup's candidate row
    T drow[W];                  // row j before the exchange (its owner's copy)
    T pv[W];                    // the chosen pivot row, decoded
    T rj[W];                    // row j, decoded
    T wv[PANEL_THREADS / 64];   // per-wave candidates
    int wr[PANEL_THREADS / 64];
    int bad;                    // a poll ran past the deadline
    uint32_t raw[mrm::LU_MAX_PANEL_WG * Rec<T, W>::RS + Rec<T, W>::DU];  // every record of the column
};

__device__ __forceinline__ void st_unit(uint64_t* p, uint32_t flag, uint32_t v) {
    __hip_atomic_store(p, (uint64_t(flag) << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_unit(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One column j of the panel (j is a compile-time constant after unrolling, so
// v stays in registers).  Critical path: local argmax, publication of the
// workgroup's record (and of row j by its owner), one polled read of all
// records into LDS, redundant selection by every thread.  Returns false when
// the grid stops (singular column or a poll past the deadline).
template <typename T, int W, int RPT>
__device__ __forceinline__ bool panel_column(const int j, T (&v)[RPT][W], const int (&lr)[RPT], int m, int g, int G,
                                             int t, int lane, int wave, int piv_base, int* __restrict__ indx,
                                             uint64_t* __restrict__ xch, uint32_t flag_base,
                                             int* __restrict__ status, uint64_t deadline, PanelShared<T, W>& sh) {
    constexpr int NW = PANEL_THREADS / 64;
    using R = Rec<T, W>;
    constexpr int UPV = R::UPV;
    static_assert(R::RS <= 128 && R::DU <= PANEL_THREADS - 128, "record does not fit the publishing threads");
    const uint32_t flag = flag_base + uint32_t(j);
    uint64_t* recs = xch + size_t(j & 1) * R::parity_units(G);
    uint64_t* diag = recs + size_t(G) * R::RS;
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
    best = sh.wv[0];
    brow = sh.wr[0];
#pragma unroll
    for (int q = 1; q < NW; ++q)
        if (better(sh.wv[q], sh.wr[q], best, brow)) {
            best = sh.wv[q];
            brow = sh.wr[q];
        }
    // (b) the owners put the candidate row and row j into LDS ...
    const bool own_j = (j >= g * RPT * PANEL_THREADS) && (j < (g + 1) * RPT * PANEL_THREADS);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] == brow) {
#pragma unroll
            for (int c = 0; c < W; ++c) sh.mine[c] = v[r][c];
        }
        if (lr[r] == j) {
#pragma unroll
            for (int c = 0; c < W; ++c) sh.drow[c] = v[r][c];
        }
    }
    __syncthreads();
    // ... and publish them, one flagged word per thread
    if (t < R::RS) {
        uint32_t u;
        if (t < UPV) u = Unit<T>::part(best, t);
        else if (t == UPV) u = uint32_t(brow);
        else u = Unit<T>::part(sh.mine[(t - UPV - 1) / UPV], (t - UPV - 1) % UPV);
        st_unit(recs + size_t(g) * R::RS + t, flag, u);
    } else if (own_j && t >= 128 && t < 128 + R::DU) {
        const int e = t - 128;
        st_unit(diag + e, flag, Unit<T>::part(sh.drow[e / UPV], e % UPV));
    }
    // (c) poll every record (and row j) into LDS
    const int nrec = G * R::RS;
    const int total = nrec + (own_j ? 0 : R::DU);
    bool late = false;
    for (int e = t; e < total; e += PANEL_THREADS) {
        const uint64_t* p = e < nrec ? recs + e : diag + (e - nrec);
        uint64_t u = ld_unit(p);
        unsigned polls = 0;
        while (uint32_t(u >> 32) != flag) {
            if ((++polls & 63u) == 0 && wall_clock64() > deadline) {
                late = true;
                break;
            }
            u = ld_unit(p);
        }
        sh.raw[e] = uint32_t(u);
    }
    if (late) {
        sh.bad = 1;
        atomicExch(status, MRM_LU_TIMEOUT);
    }
    __syncthreads();
    if (sh.bad) return false;
    // every wave selects the pivot itself (lane q looks at record q): max
    // |value|, earliest row on ties
    T gb = T(-1);
    int gr = INT_MAX;
    for (int q = lane; q < G; q += 64) {
        const T cv = Unit<T>::value(&sh.raw[q * R::RS]);
        const int cr = int(sh.raw[q * R::RS + UPV]);
        if (better(cv, cr, gb, gr)) {
            gb = cv;
            gr = cr;
        }
    }
    wave_argmax(gb, gr);
    if (!(gb > Lim<T>::tiny())) {
        if (g == 0 && t == 0) atomicExch(status, MRM_SINGULAR);
        return false;
    }
    const int prow = gr;
    // the record holding the pivot: the workgroup that owns row prow
    const int gw = prow / (RPT * PANEL_THREADS);
    if (g == 0 && t == 0) indx[piv_base + j] = piv_base + prow;
    // decode the pivot row and row j once into LDS
    if (t < W) {
        sh.pv[t] = Unit<T>::value(&sh.raw[gw * R::RS + UPV + 1 + t * UPV]);
    } else if (t < 2 * W) {
        const int c = t - W;
        sh.rj[c] = own_j ? sh.drow[c] : Unit<T>::value(&sh.raw[nrec + c * UPV]);
    }
    __syncthreads();
    // (d) exchange rows j and prow
    if (prow != j) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            if (lr[r] == prow) {
#pragma unroll
                for (int c = 0; c < W; ++c) v[r][c] = sh.rj[c];
            } else if (lr[r] == j) {
#pragma unroll
                for (int c = 0; c < W; ++c) v[r][c] = sh.pv[c];
            }
        }
    }
    // (e) scale by the reciprocal pivot (src/LinAlgLU.pas:264) and update the panel row
    const T inv = T(1) / sh.pv[j];
    T pvc[W];
#pragma unroll
    for (int c = j + 1; c < W; ++c) pvc[c] = sh.pv[c];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lr[r] > j && lr[r] < m) {
            const T l = v[r][j] * inv;
            v[r][j] = l;
#pragma unroll
            for (int c = j + 1; c < W; ++c) v[r][c] = fma(-l, pvc[c], v[r][c]);
        }
    }
    // the next column writes sh.mine / sh.drow only after its first
    // __syncthreads and sh.pv / sh.rj after its second, when every thread is
    // done with this column's values
    return true;
}

// Columns J..W-1 by compile-time recursion, so every index into v is a constant.
template <typename T, int W, int RPT, int J>
struct PanelColumns {
    static __device__ __forceinline__ void run(int w, bool& alive, T (&v)[RPT][W], const int (&lr)[RPT], int m, int g,
                                               int G, int t, int lane, int wave, int piv_base, int* indx,
                                               uint64_t* xch, uint32_t flag_base, int* status, uint64_t deadline,
                                               PanelShared<T, W>& sh) {
        if (J < w && alive)
            alive = panel_column<T, W, RPT>(J, v, lr, m, g, G, t, lane, wave, piv_base, indx, xch, flag_base, status,
                                            deadline, sh);
        PanelColumns<T, W, RPT, J + 1>::run(w, alive, v, lr, m, g, G, t, lane, wave, piv_base, indx, xch, flag_base,
                                            status, deadline, sh);
    }
};
template <typename T, int W, int RPT>
struct PanelColumns<T, W, RPT, W> {
    static __device__ __forceinline__ void run(int, bool&, T (&)[RPT][W], const int (&)[RPT], int, int, int, int, int,
                                               int, int, int*, uint64_t*, uint32_t, int*, uint64_t,
                                               PanelShared<T, W>&) {}
};

// Factor the m x w panel at P (row stride lda elements) in place, w <= W.
// Workgroup g owns panel rows [g*RPT*256, (g+1)*RPT*256); thread t owns rows
// g*RPT*256 + r*256 + t.  Pivot rows are absolute: piv_base + local row.
template <typename T, int W, int RPT>
__global__ void __launch_bounds__(PANEL_THREADS)
lu_panel_kernel(T* __restrict__ P, int64_t lda, int m, int w, int piv_base, int* __restrict__ indx,
                uint64_t* __restrict__ xch, uint32_t flag_base, int* __restrict__ status, uint64_t timeout_ticks) {
    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    __shared__ PanelShared<T, W> sh;
    // an earlier panel found the matrix singular: nothing left to do (uniform)
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    if (t == 0) sh.bad = 0;

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

    PanelColumns<T, W, RPT, 0>::run(w, alive, v, lr, m, g, G, t, lane, wave, piv_base, indx, xch, flag_base, status,
                                    deadline, sh);
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

// X = U^-1 X for the R x ncols block, U upper with its diagonal (divides, as
// MatrixLUBackSubst does, src/LinAlgLU.pas:362-384).
template <typename T, int R>
__global__ void __launch_bounds__(128)
trsm_upper_block_kernel(const T* __restrict__ U, int64_t ldu, T* __restrict__ X, int64_t ldx, int rows, int ncols) {
    __shared__ T Us[R][R + 1];
    for (int e = threadIdx.x; e < rows * rows; e += blockDim.x) Us[e / rows][e % rows] = U[int64_t(e / rows) * ldu + e % rows];
    __syncthreads();
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
                if (k < rows) s = fma(-Us[i][k], x[k], s);
            x[i] = s / Us[i][i];
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
int lu_panel(T* P, int64_t lda, int m, int w, int piv_base, int* indx, int* perm, LuWork& ws, hipStream_t s) {
    constexpr int W = LU_W;
    // Every column reads every workgroup's record, so rows are packed up to 4
    // per thread before more workgroups are added.
    int rpt = 1;
    int G = grid1(m, PANEL_THREADS);
    while (G > LU_PANEL_WG_TARGET && rpt < 4) {
        rpt *= 2;
        G = grid1(m, rpt * PANEL_THREADS);
    }
    if (G > LU_MAX_PANEL_WG) {
        mrm_set_error(MRM_BAD_DIMS, "lu: matrix too tall for the panel kernel (%d rows)", m);
        return MRM_BAD_DIMS;
    }
    uint64_t* xch = static_cast<uint64_t*>(ws.xch);
    // flags are unique within a factorisation (the exchange area is zeroed at
    // its start): column j of this panel is flagged piv_base + j + 1
    uint32_t flag_base = uint32_t(piv_base) + 1u;
    uint64_t timeout = 2ull * 100000000ull;  // 2 s of wall clock (100 MHz)
    void* args[] = {&P, &lda, &m, &w, &piv_base, &indx, &xch, &flag_base, &ws.status, &timeout};
    hipError_t e;
    const void* fn = rpt == 1   ? reinterpret_cast<const void*>(&lu_panel_kernel<T, W, 1>)
                     : rpt == 2 ? reinterpret_cast<const void*>(&lu_panel_kernel<T, W, 2>)
                                : reinterpret_cast<const void*>(&lu_panel_kernel<T, W, 4>);
    e = hipLaunchCooperativeKernel(fn, dim3(G), dim3(PANEL_THREADS), args, 0, s);
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
int lu_factor(T* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s, int mode) {
    // indx starts at 0 as in the reference (src/LinAlgLU.pas:289); entries a
    // singular stop leaves unset stay in range for the row interchanges
    zero_i32(indx, n, s);
    // no flag of an earlier factorisation may survive in the exchange area
    const int xw = int(lu_xch_bytes() / 4);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid1(xw, 256)), dim3(256), 0, s, static_cast<unsigned*>(ws.xch), xw);
    // the low-precision path takes twice the outer block: half as many
    // read-modify-write passes of the FP64 trailing matrix
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
        const int in_lds = n <= 32768;
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

size_t lu_xch_bytes() {
    // two parities of G flagged records plus row j, 8-byte words, G <= LU_MAX_PANEL_WG
    return 2 * Rec<double, LU_W>::parity_units(LU_MAX_PANEL_WG) * sizeof(uint64_t);
}

}  // namespace mrm
