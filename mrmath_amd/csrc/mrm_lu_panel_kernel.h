// mrm_lu_panel_kernel.h -- the panel kernel of the LU factorisation, included
// by one translation unit per rows-per-thread variant (so the variants
// compile in parallel) and by the launcher (for the record layout).
//
// Factors an m x w (w <= LU_W) column panel in place with partial pivoting:
// the left-looking column loop of the reference's recursive LU at its base
// case (src/LinAlgLU.pas:236-271: pivot search, LUSwap, scaling by the
// reciprocal pivot, rank-1 update).  The panel's rows stay in registers,
// spread over as many workgroups as its height needs; per column every
// workgroup publishes its candidate in flagged words and polls everybody
// else's (one memory round trip, no grid barrier) to agree on the pivot.  The
// pivot rule is the reference's: the first row (in the current row order) of
// maximal |a| in the column, singular when |pivot| <= 10/MaxDouble.

#pragma once

#include <cfloat>
#include <climits>
#include <cstdint>

#include "mrm_common.h"
#include "mrm_lu.h"
#include "mrm_wave.h"

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

using mrm::wave::better;

template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& r) {
    mrm::wave::argmax(v, r);
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
    // (b) the owners put the candidate row and row j into LDS (only the waves
    // holding one of the two rows take the branch) ...
    const bool own_j = (j >= g * RPT * PANEL_THREADS) && (j < (g + 1) * RPT * PANEL_THREADS);
    bool holds = false;
#pragma unroll
    for (int r = 0; r < RPT; ++r) holds |= (lr[r] == brow) || (lr[r] == j);
    if (__any(holds)) {
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
    // (c) poll every record (and row j) into LDS: a thread's words are read
    // together, and each round re-reads together the words whose flag is not
    // there yet, so a poll costs one memory round trip however many words a
    // thread watches
    const int nrec = G * R::RS;
    const int total = nrec + (own_j ? 0 : R::DU);
    // deadline 0: the host asked every wait to give up at once
    bool late = deadline == 0;
    constexpr int PB = 8;  // words in flight per thread
    for (int e0 = t; e0 < total && !late; e0 += PB * PANEL_THREADS) {
        uint64_t u[PB];
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int e = e0 + i * PANEL_THREADS;
            u[i] = e < total ? ld_unit(e < nrec ? recs + e : diag + (e - nrec)) : (uint64_t(flag) << 32);
        }
        unsigned polls = 0;
        for (;;) {
            bool pending = false;
#pragma unroll
            for (int i = 0; i < PB; ++i) pending |= uint32_t(u[i] >> 32) != flag;
            if (!pending) break;
            if ((++polls & 63u) == 0 && wall_clock64() > deadline) {
                late = true;
                break;
            }
#pragma unroll
            for (int i = 0; i < PB; ++i) {
                const int e = e0 + i * PANEL_THREADS;
                if (uint32_t(u[i] >> 32) != flag) u[i] = ld_unit(e < nrec ? recs + e : diag + (e - nrec));
            }
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            const int e = e0 + i * PANEL_THREADS;
            if (e < total) sh.raw[e] = uint32_t(u[i]);
        }
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
    // (d) exchange rows j and prow (only the waves holding one of them)
    bool swaps = false;
#pragma unroll
    for (int r = 0; r < RPT; ++r) swaps |= (lr[r] == prow) || (lr[r] == j);
    if (prow != j && __any(swaps)) {
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
    const uint64_t deadline = timeout_ticks ? wall_clock64() + timeout_ticks : 0;
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

}  // namespace

// The kernel of one rows-per-thread variant, as a launchable address.
namespace mrm {
template <int RPT>
const void* lu_panel_fn();
}  // namespace mrm
