// mrm_batched_lu_wave.hip -- small linear solves of up to 64 unknowns, one
// system per wave, the matrix in registers.
//
// The per-system solve of the reference's MatrixLinEQSolve (LU with partial
// pivoting, src/LinAlgLU.pas:181-281, then MatrixLUBackSubst, :315-385) for
// the batched plans.  Lane i holds row i of A (NMAX values, the column index
// a compile-time register index after unrolling); rows are never moved: each
// lane keeps the position its row has in the pivoted order, so an
// interchange is two position updates.  Per column j: the pivot is the wave
// argmax of |a[j]| over the rows not yet used (ties to the earlier position,
// the reference's "first maximum in the current row order"); the pivot row
// is broadcast through a per-wave LDS row; every remaining row scales its
// entry by the reciprocal pivot (src/LinAlgLU.pas:264) and subtracts the
// multiple of the pivot row.  No barrier: each wave works alone.  The
// substitutions run in position order with the values of one row broadcast
// by readlane per step.  Mode 1 factors in FP32 and refines in FP64 against
// A read again from memory: x += (LU)^-1 P (b - A x).

#include <cfloat>
#include <type_traits>

#include "mrm_batched_lu.h"
#include "mrm_wave.h"

namespace {

using mrm::BatchLuP;
using mrm::BL_THREADS;

template <typename T>
__device__ __forceinline__ T tiny_pivot();
template <>
__device__ __forceinline__ double tiny_pivot<double>() { return 10.0 / DBL_MAX; }  // src/LinAlgLU.pas:262
template <>
__device__ __forceinline__ float tiny_pivot<float>() { return 10.0f / FLT_MAX; }

// The wave's LDS writes are complete before any lane reads them.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// The lane whose row sits at position j (exactly one live lane).
__device__ __forceinline__ int lane_at(int pos, int j) {
    return int(__builtin_ctzll(__ballot(pos == j)));
}

// rv (this lane's right-hand-side entry, in its row) := the solution entry of
// the position the row ends at: forward substitution with the unit lower
// factor (multipliers a[j] of the rows eliminated at step j), then backward
// with the upper factor (a[k], k >= pos).  x_j is written to xs[j] (LDS).
template <typename T, int NMAX, typename R>
__device__ __forceinline__ void wave_solve(const T (&a)[NMAX], int pos, int n, R rv, double* xs) {
    mrm::static_for<0, NMAX>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j >= n) return;
        const R yj = mrm::wave::readlane_v(rv, lane_at(pos, j));
        if (pos > j) rv -= R(a[j]) * yj;
    });
    mrm::static_for<0, NMAX>([&](auto ic) {
        constexpr int j = NMAX - 1 - decltype(ic)::value;
        if (j >= n) return;
        const int q = lane_at(pos, j);
        if (pos == j) rv = rv / R(a[j]);
        const R xj = mrm::wave::readlane_v(rv, q);
        if (pos < j) rv -= R(a[j]) * xj;
        if (int(threadIdx.x % 64) == q) xs[j] = double(xj);
    });
    wave_lds_sync();
}

template <int MODE, int NMAX>
__global__ void __launch_bounds__(BL_THREADS) bl_wave_kernel(const BatchLuP* __restrict__ sys, int count,
                                                             unsigned* __restrict__ counter, int* __restrict__ status,
                                                             int refine) {
    using T = typename std::conditional<MODE == 0, double, float>::type;
    constexpr int NW = BL_THREADS / 64;
    __shared__ __attribute__((aligned(16))) T urow_all[NW][NMAX];
    __shared__ double xs_all[NW][NMAX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T* urow = urow_all[wave];
    double* xs = xs_all[wave];
    for (;;) {
        int item = 0;
        if (lane == 0) item = int(atomicAdd(counter, 1u));
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= count) return;
        const BatchLuP s = sys[item];
        const int n = int(s.n);
        const bool live = lane < n;
        T a[NMAX];
        {
            const double* row = s.A + int64_t(live ? lane : 0) * s.lda;
            mrm::static_for<0, NMAX>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                a[k] = (live && k < n) ? T(row[k]) : T(0);
            });
        }
        const double bl = live ? s.b[lane] : 0.0;
        int pos = lane;        // position of this lane's row in the pivoted order
        bool used = !live;     // the row has been a pivot row (or does not exist)
        bool ok = true;
        mrm::static_for<0, NMAX>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j >= n || !ok) return;
            T v = used ? T(-1) : fabs(a[j]);
            int key = pos * 64 + lane;   // ties go to the earlier position
            mrm::wave::argmax(v, key);
            if (!(v > tiny_pivot<T>())) {
                ok = false;
                return;
            }
            const int p = key & 63, pp = key >> 6;
            // interchange positions j and pp (rows stay in their lanes)
            if (!used && pos == j) pos = pp;
            if (lane == p) {
                pos = j;
                used = true;
#pragma unroll
                for (int k = j; k < NMAX; ++k) urow[k] = a[k];
            }
            wave_lds_sync();
            if (!used) {
                const T l = a[j] * (T(1) / urow[j]);
                a[j] = l;
#pragma unroll
                for (int k = j + 1; k < NMAX; ++k) a[k] = fma(-l, urow[k], a[k]);
            }
            wave_lds_sync();
        });
        if (ok) {
            if (MODE == 0) {
                wave_solve<T, NMAX, double>(a, pos, n, bl, xs);
            } else {
                wave_solve<T, NMAX, float>(a, pos, n, float(bl), xs);
                for (int it = 0; it < refine; ++it) {
                    // r = b - A x in FP64 (this lane's row of A from memory)
                    double r = bl;
                    if (live) {
                        const double* row = s.A + int64_t(lane) * s.lda;
                        for (int k = 0; k < n; ++k) r -= row[k] * xs[k];
                    }
                    const double x_old = live ? xs[lane] : 0.0;
                    wave_lds_sync();
                    wave_solve<T, NMAX, float>(a, pos, n, float(r), xs);  // xs := the correction
                    if (live) xs[lane] = x_old + xs[lane];
                    wave_lds_sync();
                }
            }
            if (live) s.x[lane] = xs[lane];
        }
        if (lane == 0 && status) status[s.idx] = ok ? MRM_OK : MRM_SINGULAR;
        wave_lds_sync();
    }
}

}  // namespace

namespace mrm {

const void* batched_lu_wave_fn(int mode, int nmax) {
    if (mode == 0)
        return nmax <= 32 ? reinterpret_cast<const void*>(&bl_wave_kernel<0, 32>)
                          : reinterpret_cast<const void*>(&bl_wave_kernel<0, 64>);
    return nmax <= 32 ? reinterpret_cast<const void*>(&bl_wave_kernel<1, 32>)
                      : reinterpret_cast<const void*>(&bl_wave_kernel<1, 64>);
}

}  // namespace mrm
