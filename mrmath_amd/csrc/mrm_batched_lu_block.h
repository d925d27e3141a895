// mrm_batched_lu_block.h -- small linear solves of 65..128 unknowns, one
// system per workgroup of eight waves, the matrix in registers.
//
// The same solve as mrm_batched_lu_wave.hip (the per-system MatrixLinEQSolve
// of the reference, src/LinAlgLU.pas:181-385) for matrices too large for one
// wave's registers.  Wave w holds rows 64 * (w % 2) .. + 63 (one row per
// lane) and the columns c = q + 4 k (k = 0 .. NMAX/4 - 1) of them, q = w / 2:
// every thread keeps a quarter row, the columns interleaved so that every
// quarter has the same compile-time loop bounds at every step and the work
// stays balanced between the quarters.  Rows stay in their threads and carry
// their position in the pivoted order.  Per column j, two barriers:
//   1. the two waves holding column j put it into LDS with their wave argmax
//      of |a| over the unused rows (ties to the earlier position) and the
//      winner's reciprocal (every lane divides its own entry while the argmax
//      runs, so the division is off the chain);
//   2. after the first barrier every thread knows the pivot, its row's
//      multiplier (column value times the reciprocal pivot, src/LinAlgLU.pas:264)
//      and its new position; the four threads holding the pivot row put it
//      into LDS; after the second barrier every unused row subtracts the
//      multiple of the pivot row from its quarter.
// The right-hand side rides along as one more column (forward substitution
// for free).  The substitutions take one barrier per unknown: the four threads
// of the row at position j post their quarters' partial sums, after the
// barrier every thread forms the unknown and the owner quarter of column j
// adds its term to its own partial sums.  The backward pass multiplies by the
// pivot reciprocal kept from the factorisation.  Mode 1 factors in FP32 and
// refines in FP64 against A read again from memory: x += (LU)^-1 P (b - A x).

// One instantiation per source file (mrm_batched_lu_block_{d,s}{96,128}.hip)
// so that the four compile in parallel.
#pragma once

#include <cfloat>
#include <type_traits>

#include "mrm_batched_lu.h"
#include "mrm_wave.h"

namespace mrm {
namespace blk {

using mrm::BatchLuP;
using mrm::BL_BLOCK_THREADS;

constexpr int NQ = 4;  // column quarters

template <typename T>
__device__ __forceinline__ T tiny_pivot();
template <>
__device__ __forceinline__ double tiny_pivot<double>() { return 10.0 / DBL_MAX; }  // src/LinAlgLU.pas:262
template <>
__device__ __forceinline__ float tiny_pivot<float>() { return 10.0f / FLT_MAX; }

template <typename T, int NMAX>
struct BlockShared {
    T col[2][128];           // column j, all rows (by parity of j)
    T urow[2][NQ][NMAX / NQ];// the pivot row by quarter (by parity)
    T ub[2];                 // ... and its right-hand side
    T cv[2][2];              // per row wave: its candidate |a| ...
    int ck[2][2];            // ... key (position * 128 + row) ...
    T ci[2][2];              // ... and the candidate's reciprocal
    double fp[2][NQ];        // substitutions: partial sums of the row at position j
    double fb[2];            // ... its right-hand side
    double fd[2];            // ... its pivot reciprocal
    double xs[128];          // solution / correction by unknown
    double part[NQ][128];    // residual partial sums by quarter
    int item;
};

// Solves L U x = (the right-hand side c of this thread's row, in its
// position's slot) with the factors held by the workgroup: forward with the
// unit lower factor first when `forward` (c is then the right-hand side of the
// row; otherwise c has already been eliminated), then backward.  Only the
// quarter-0 thread of a row needs c.  Result: sh.xs[j] = x_j, R arithmetic.
template <typename T, int NMAX, typename R>
__device__ __forceinline__ void block_solve(const T (&a)[NMAX / NQ], int pos, T dinv, int n, int q, int tid, R c,
                                            bool forward, BlockShared<T, NMAX>& sh) {
    if (forward) {
        R acc = R(0);
        mrm::static_for<0, NMAX>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            constexpr int par = j & 1, qj = j % NQ, jj = j / NQ;
            if (j >= n) return;
            if (pos == j) {
                sh.fp[par][q] = double(acc);
                if (q == 0) sh.fb[par] = double(c);
            }
            __syncthreads();
            const R yj = R(sh.fb[par]) - (R(sh.fp[par][0]) + R(sh.fp[par][1]) + R(sh.fp[par][2]) + R(sh.fp[par][3]));
            if (pos == j) c = yj;
            if (q == qj && pos > j) acc += R(a[jj]) * yj;
        });
    }
    R acc = R(0);
    mrm::static_for<0, NMAX>([&](auto ic) {
        constexpr int j = NMAX - 1 - decltype(ic)::value;
        constexpr int par = j & 1, qj = j % NQ, jj = j / NQ;
        if (j >= n) return;
        if (pos == j) {
            sh.fp[par][q] = double(acc);
            if (q == 0) {
                sh.fb[par] = double(c);
                sh.fd[par] = double(dinv);
            }
        }
        __syncthreads();
        const R xj = (R(sh.fb[par]) - (R(sh.fp[par][0]) + R(sh.fp[par][1]) + R(sh.fp[par][2]) + R(sh.fp[par][3]))) *
                     R(sh.fd[par]);
        if (tid == 0) sh.xs[j] = double(xj);
        if (q == qj && pos < j) acc += R(a[jj]) * xj;
    });
    __syncthreads();
}

template <int MODE, int NMAX>
__global__ void __launch_bounds__(BL_BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
bl_block_kernel(const BatchLuP* __restrict__ sys, int count, unsigned* __restrict__ counter, int* __restrict__ status,
                int refine) {
    using T = typename std::conditional<MODE == 0, double, float>::type;
    constexpr int QW = NMAX / NQ;
    static_assert(NMAX <= 128 && NMAX % NQ == 0, "two row waves of 64, four column quarters");
    static_assert(BL_BLOCK_THREADS == 2 * 64 * NQ, "one wave per (row half, column quarter)");
    __shared__ __attribute__((aligned(16))) BlockShared<T, NMAX> sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // wave-uniform values go through readfirstlane so that branches on them
    // stay scalar
    const int rh = __builtin_amdgcn_readfirstlane(wave & 1);
    const int r = lane + 64 * rh;                              // this thread's row
    const int q = __builtin_amdgcn_readfirstlane(wave >> 1);  // its columns: q + NQ * k
    for (;;) {
        if (tid == 0) sh.item = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(sh.item);
        if (item >= count) return;
        const BatchLuP s = sys[item];
        const int n = int(s.n);
        const bool live = r < n;
        T a[QW];
        {
            const double* row = s.A + int64_t(live ? r : 0) * s.lda + q;
            mrm::static_for<0, QW>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                a[k] = (live && q + NQ * k < n) ? T(row[NQ * k]) : T(0);
            });
        }
        const double br = live ? s.b[r] : 0.0;
        T bb = T(br);   // the right-hand side, eliminated along in every quarter
        T dinv = T(0);  // the reciprocal pivot of this row's step
        int pos = r;
        bool used = !live;
        bool ok = true;
        mrm::static_for<0, NMAX>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j >= n || !ok) return;
            constexpr int par = j & 1, qj = j % NQ, jj = j / NQ;
            if (q == qj) {
                const T cj = a[jj];
                sh.col[par][r] = cj;
                const T inv = T(1) / cj;
                T v = used ? T(-1) : fabs(cj);
                int key = pos * 128 + r;
                mrm::wave::argmax(v, key);
                const T winv = mrm::wave::readlane_v(inv, key & 63);
                if (lane == 0) {
                    sh.cv[par][rh] = v;
                    sh.ck[par][rh] = key;
                    sh.ci[par][rh] = winv;
                }
            }
            __syncthreads();
            T bv = sh.cv[par][0], bi = sh.ci[par][0];
            int bk = sh.ck[par][0];
            {
                const T v1 = sh.cv[par][1];
                const int k1 = sh.ck[par][1];
                if (v1 > bv || (v1 == bv && k1 < bk)) {
                    bv = v1;
                    bk = k1;
                    bi = sh.ci[par][1];
                }
            }
            if (!__builtin_amdgcn_readfirstlane(int(bv > tiny_pivot<T>()))) {  // the same in every thread
                ok = false;
                return;
            }
            const int p = bk & 127, pp = bk >> 7;
            const T l = sh.col[par][r] * bi;
            if (!used && pos == j) pos = pp;
            if (r == p) {
                pos = j;
                used = true;
                dinv = bi;
                // the pivot row's columns from j on in this quarter (and its right-hand side)
#pragma unroll
                for (int k = jj; k < QW; ++k) sh.urow[par][q][k] = a[k];
                if (q == 0) sh.ub[par] = bb;
            }
            __syncthreads();
            if (!used) {
#pragma unroll
                for (int k = jj + 1; k < QW; ++k) a[k] = fma(-l, sh.urow[par][q][k], a[k]);
                // column q + NQ jj: after column j for quarters above qj, column j itself at qj
                const T upd = fma(-l, sh.urow[par][q][jj], a[jj]);
                a[jj] = q > qj ? upd : q == qj ? l : a[jj];
                bb = fma(-l, sh.ub[par], bb);
            }
        });
        if (ok) {
            block_solve<T, NMAX, T>(a, pos, dinv, n, q, tid, bb, false, sh);
            if (MODE == 1) {
                for (int it = 0; it < refine; ++it) {
                    // r = b - A x in FP64: each quarter sums a contiguous quarter of row r
                    double acc = 0.0;
                    if (live) {
                        const int k0 = q * QW, k1 = n < k0 + QW ? n : k0 + QW;
                        const double* row = s.A + int64_t(r) * s.lda;
                        for (int k = k0; k < k1; ++k) acc += row[k] * sh.xs[k];
                    }
                    sh.part[q][r] = acc;
                    const double x_old = tid < n ? sh.xs[tid] : 0.0;
                    __syncthreads();
                    double rr = 0.0;
                    if (q == 0 && live) {
                        rr = br;
#pragma unroll
                        for (int qq = 0; qq < NQ; ++qq) rr -= sh.part[qq][r];
                    }
                    block_solve<T, NMAX, float>(a, pos, dinv, n, q, tid, float(rr), true, sh);
                    if (tid < n) sh.xs[tid] = x_old + sh.xs[tid];
                    __syncthreads();
                }
            }
            if (tid < n) s.x[tid] = sh.xs[tid];
        }
        if (tid == 0 && status) status[s.idx] = ok ? MRM_OK : MRM_SINGULAR;
        __syncthreads();
    }
}

}  // namespace blk
}  // namespace mrm
