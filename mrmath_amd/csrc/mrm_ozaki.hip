// mrm_ozaki.hip -- FP64-accurate GEMM by split-precision slicing on the INT8
// matrix cores (Ozaki scheme with integer slices; Ozaki, Ogita, Oishi, Rump 2012,
// and the integer-unit variant of Ootomo, Ozaki and Yokota 2024).
//
// C (op)= A * B for FP64 A (M x K), B (K x N), any of NN / TN / NT:
//  1. every row i of A gets a power-of-two scale 2^ea_i with max_k |A_ik| < 2^ea_i,
//     every column j of B a scale 2^eb_j; the scaled values lie in (-1, 1);
//  2. each scaled value y is cut into s signed 7-bit digits
//     d_t = trunc(128 y_t), y_{t+1} = 128 y_t - d_t (all steps exact in FP64),
//     so y = sum_t d_t 2^(-7(t+1)) + r with |r| < 2^(-7s);
//  3. A*B = 2^(ea_i + eb_j) sum_{t,u} 2^(-7(t+u+2)) (D^A_t D^B_u)_ij; the pairs
//     with t + u = d are summed exactly in one int32 accumulator over a K chunk
//     of at most MRM_OZAKI_KCHUNK (|sum| <= (d+1) KCHUNK 127^2 < 2^31), pairs
//     with t + u >= s are dropped (below the truncation error), and the
//     diagonals are accumulated in FP64 in the epilogue, largest first.
// The INT8 products run on v_mfma_i32_32x32x32_i8 (lane l supplies row (A) /
// column (B) l%32 and one 16-byte k chunk chosen by l/32; the result register
// v of lane l is row 8(v/4) + 4(l/32) + v%4, column l%32).
//
// Slices are stored in 64-byte k blocks, [slice][k/64][row][k%64], K padded to
// a multiple of 64 and rows to a multiple of the 256-row tile with zeros, for
// A by rows and for B by columns: one K step of a 256-row panel is 16 KiB of
// contiguous memory, and the GEMM kernel has no edge cases except in its
// epilogue.

#include <cmath>
#include <cstdint>

#include "mrm_common.h"
#include "mrm_host.h"
#include "mrm_ozaki.h"

namespace {

typedef int mrm_i4 __attribute__((ext_vector_type(4)));
typedef int mrm_i16 __attribute__((ext_vector_type(16)));
typedef const __attribute__((address_space(1))) void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

constexpr int OZ_BK = 64;                 // bytes (= int8 k-values) per K step
constexpr int OZ_THREADS = 256;

// ---------------------------------------------------------------------------
// Scaling exponents, stored biased: ev[r] = e + OZ_EBIAS with
// max_k |X(r,k)| < 2^e, or 0 for an all-zero row (decoded e = 0).  The biased
// form makes "no nonzero seen yet" the smallest value, so partial maxima over
// K pieces combine with atomicMax.  X(r,k) = X[r*rs + k*ks].
// ---------------------------------------------------------------------------
constexpr int OZ_EBIAS = 4096;

__device__ __forceinline__ int oz_biased_exponent(double amax) {
    if (!(amax > 0.0)) return 0;
    int e;
    (void)frexp(amax, &e);   // amax = f * 2^e, f in [0.5, 1)
    return e + OZ_EBIAS;
}
__device__ __forceinline__ int oz_exponent(int ev) { return ev ? ev - OZ_EBIAS : 0; }

__global__ void __launch_bounds__(256) oz_zero_kernel(int* __restrict__ e, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) e[i] = 0;
}

// rows contiguous in k (ks == 1): one workgroup per row
__global__ void __launch_bounds__(256) oz_exp_rows_kernel(const double* __restrict__ X, int64_t rs, int64_t R, int64_t K,
                                                          int* __restrict__ ev) {
    const int64_t r = blockIdx.x;
    const double* row = X + r * rs;
    double m = 0.0;
    for (int64_t k = threadIdx.x; k < K; k += blockDim.x) m = fmax(m, fabs(row[k]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) ev[r] = oz_biased_exponent(fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3])));
}

// rows strided (rs == 1, k at stride ks): a 64-row x 256-k piece per
// workgroup (64 rows x 4 k-lanes), combined over pieces with atomicMax on the
// zeroed biased exponents
__global__ void __launch_bounds__(256) oz_exp_cols_kernel(const double* __restrict__ X, int64_t ks, int64_t R, int64_t K,
                                                          int* __restrict__ ev) {
    const int64_t r = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
    const int kl = threadIdx.x >> 6;
    const int64_t kb = int64_t(blockIdx.y) * 256, ke = kb + 256 < K ? kb + 256 : K;
    double m = 0.0;
    if (r < R)
        for (int64_t k = kb + kl; k < ke; k += 4) m = fmax(m, fabs(X[k * ks + r]));
    __shared__ double part[4][64];
    part[kl][threadIdx.x & 63] = m;
    __syncthreads();
    if (kl == 0 && r < R) {
        const int c = threadIdx.x;
        const int v = oz_biased_exponent(fmax(fmax(part[0][c], part[1][c]), fmax(part[2][c], part[3][c])));
        if (v) atomicMax(ev + r, v);
    }
}

// ---------------------------------------------------------------------------
// Slicing: a 64 (r) x 64 (k) tile per workgroup; S[t][k/64][r][k%64] (k-block
// pitch Rp*64 bytes, slice pitch spitch bytes).  Padding rows / k beyond R / K
// are written as 0.  Each thread cuts 4 consecutive k of one row and stores
// their 4 digits per slice as one 32-bit word.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) oz_slice_kernel(const double* __restrict__ X, int64_t rs, int64_t ks, int64_t R,
                                                       int64_t K, const int* __restrict__ ev, int8_t* __restrict__ S,
                                                       int64_t Rp, int64_t spitch, int s) {
    __shared__ __attribute__((aligned(16))) int8_t tile[MRM_OZAKI_MAX_SLICES][64][80];
    const int64_t r0 = int64_t(blockIdx.y) * 64, k0 = int64_t(blockIdx.x) * 64;
    const bool kfast = ks == 1;
    for (int q = threadIdx.x; q < 64 * 16; q += 256) {
        // coalesce the FP64 reads along whichever index is contiguous
        const int rr = kfast ? q / 16 : q % 64;
        const int kq = kfast ? q % 16 : q / 64;
        const int64_t r = r0 + rr;
        double y[4];
        const int e = r < R ? oz_exponent(ev[r]) : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t k = k0 + 4 * kq + i;
            y[i] = (r < R && k < K) ? ldexp(X[r * rs + k * ks], -e) : 0.0;
        }
        for (int t = 0; t < s; ++t) {
            uint32_t w = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                y[i] *= 128.0;
                const double d = trunc(y[i]);
                y[i] -= d;
                w |= uint32_t(uint8_t(int8_t(d))) << (8 * i);
            }
            *reinterpret_cast<uint32_t*>(&tile[t][rr][4 * kq]) = w;
        }
    }
    __syncthreads();
    // 64 rows x 64 bytes per slice: one contiguous 4 KiB block
    for (int q = threadIdx.x; q < s * 64 * 4; q += 256) {
        const int t = q / 256, rr = (q / 4) % 64, c = q % 4;
        const int64_t r = r0 + rr;
        const int4 v = *reinterpret_cast<const int4*>(&tile[t][rr][16 * c]);
        *reinterpret_cast<int4*>(S + t * spitch + (k0 / 64) * Rp * 64 + r * 64 + 16 * c) = v;
    }
}

// ---------------------------------------------------------------------------
// INT8 GEMM over one diagonal d: acc = sum_{t=0..d} A_t * B_(d-t)^T (exact
// int32), then C (op)= acc * 2^(ea_i + eb_j - 7(d+2)).
// Workgroup tile BM x BN = 64 PM x 64 PN, K step 64 bytes; 4 waves in 2 x 2,
// each PM x PN blocks of 32 x 32 (v_mfma_i32_32x32x32_i8, which measures
// ~4.5 POPS against ~3.5 for the 16x16x64 form).  S-stage LDS ring
// filled by direct global->LDS loads, one barrier per K step.
// LDS panels [row][64 bytes], 16-byte chunk c of row r at slot c ^ ((r/4) % 4).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int oz_slot(int r, int c) { return c ^ ((r >> 2) & 3); }

// XCD-aware tile order: consecutive workgroups go to different XCDs, so XCD x
// gets a contiguous run of logical tiles, walked in groups of 4 tile rows.
__device__ __forceinline__ void oz_tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int NXCD = 8, GROUP_M = 4;
    const int total = tiles_m * tiles_n;
    const int xcd = bid % NXCD;
    const int q = total / NXCD, rem = total % NXCD;
    const int logical = xcd * q + (xcd < rem ? xcd : rem) + bid / NXCD;
    const int per_group = GROUP_M * tiles_n;
    const int group = logical / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = (tiles_m - first_m) < GROUP_M ? (tiles_m - first_m) : GROUP_M;
    const int in_group = logical - group * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

template <int PM, int PN, int S>
__global__ void __launch_bounds__(OZ_THREADS, (PM * PN > 8 ? 1 : 2))
oz_gemm_kernel(const int8_t* __restrict__ SA, const int8_t* __restrict__ SB, int64_t kbpa, int64_t kbpb, int64_t spa, int64_t spb,
               int64_t kstep0, int64_t nk, int d, const int* __restrict__ ea, const int* __restrict__ eb,
               double* __restrict__ C, int64_t ldc, int64_t M, int64_t N, int op, int tiles_m, int tiles_n) {
    constexpr int BM = 64 * PM, BN = 64 * PN;
    constexpr int LA = BM / 64, LB = BN / 64;  // 16-row load instructions per wave
    constexpr int STAGE = (BM + BN) * OZ_BK;
    __shared__ __attribute__((aligned(16))) int8_t smem[S * STAGE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    oz_tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;

    // direct loads: wave w writes panel rows [w*BM/4, (w+1)*BM/4) as 16-row
    // chunks; lane L covers row L/4 of a chunk, LDS slot L%4 <- HBM chunk
    // (L%4) ^ ((L/16)%4)
    const int lrow = lane >> 2;
    const int gchunk = (lane & 3) ^ ((lane >> 4) & 3);
    const int64_t steps = int64_t(d + 1) * nk;
    const int8_t* pa = SA + kstep0 * kbpa + (m0 + wave * (BM / 4) + lrow) * OZ_BK + 16 * gchunk;
    const int8_t* pb = SB + int64_t(d) * spb + kstep0 * kbpb + (n0 + wave * (BN / 4) + lrow) * OZ_BK + 16 * gchunk;
    int64_t kleft = nk;

    auto issue = [&](int stage) {
        int8_t* la = smem + stage * STAGE + wave * (BM / 4) * OZ_BK;
        int8_t* lb = smem + stage * STAGE + BM * OZ_BK + wave * (BN / 4) * OZ_BK;
#pragma unroll
        for (int i = 0; i < LA; ++i)
            __builtin_amdgcn_global_load_lds((gptr_t)(pa + 16 * i * OZ_BK), (lptr_t)(la + 16 * i * OZ_BK), 16, 0, 0);
#pragma unroll
        for (int i = 0; i < LB; ++i)
            __builtin_amdgcn_global_load_lds((gptr_t)(pb + 16 * i * OZ_BK), (lptr_t)(lb + 16 * i * OZ_BK), 16, 0, 0);
        pa += kbpa;
        pb += kbpb;
        if (--kleft == 0) {  // next slice pair (t+1, d-t-1)
            kleft = nk;
            pa += spa - nk * kbpa;
            pb -= spb + nk * kbpb;
        }
    };

    // v_mfma_i32_32x32x32_i8: lane l supplies row (A) / column (B) l%32 and the
    // 16 k-values of chunk l/32 of the 32-wide k half; result register v of
    // lane l is row 8(v/4) + 4(l/32) + v%4, column l%32
    mrm_i16 acc[PM][PN];
#pragma unroll
    for (int p = 0; p < PM; ++p)
#pragma unroll
        for (int c = 0; c < PN; ++c) acc[p][c] = mrm_i16{};

    // Fragments of one 32-wide k half: lane l reads row frow of each block,
    // 16-byte chunk 2h + l/32.
    const int frow = lane & 31, fchunk = lane >> 5;
    mrm_i4 fa[2][PM], fb[2][PN];
    auto load_frags = [&](int stage, int h, mrm_i4* a, mrm_i4* b) {
        const int8_t* As = smem + stage * STAGE;
        const int8_t* Bs = As + BM * OZ_BK;
#pragma unroll
        for (int c = 0; c < PN; ++c) {
            const int r = wn * (BN / 2) + 32 * c + frow;
            b[c] = *reinterpret_cast<const mrm_i4*>(Bs + r * OZ_BK + 16 * oz_slot(r, 2 * h + fchunk));
        }
#pragma unroll
        for (int p = 0; p < PM; ++p) {
            const int r = wm * (BM / 2) + 32 * p + frow;
            a[p] = *reinterpret_cast<const mrm_i4*>(As + r * OZ_BK + 16 * oz_slot(r, 2 * h + fchunk));
        }
    };
    auto mma = [&](const mrm_i4* a, const mrm_i4* b) {
#pragma unroll
        for (int p = 0; p < PM; ++p)
#pragma unroll
            for (int c = 0; c < PN; ++c) acc[p][c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[p], b[c], acc[p][c], 0, 0, 0);
    };
    // Wait until this thread's loads of step `ready` have landed, given that
    // the loads of `later` further steps may stay in flight (an immediate
    // count; fewer in flight near the end, then wait for all).
    auto wait_loads = [&](int64_t later) {
        if (later >= S - 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((S - 1) * (LA + LB)) : "memory");
        else if (later == S - 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((S - 2) * (LA + LB)) : "memory");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    };
    // A plain s_barrier (__syncthreads() would also wait for the loads still
    // in flight for later stages).
    auto barrier = [&]() {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's LDS reads are done
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };

    // Ring of S stages: step st lives in stage st % S.  Each K step computes
    // its first half while the second half's fragments load, then waits for
    // step st+1, refills the stage just consumed with step st+S and loads the
    // next step's first fragments while the second half computes.
#pragma unroll
    for (int i = 0; i < S; ++i)
        if (i < steps) issue(i);
    wait_loads(steps - 1 < S - 1 ? steps - 1 : S - 1);
    barrier();
    load_frags(0, 0, fa[0], fb[0]);
    // (sched_barrier keeps the compiler from sinking the fragment reads
    // behind the MFMAs they are meant to overlap; the lgkmcnt(0) before a
    // batch of reads costs nothing -- the previous batch had a half step of
    // MFMAs to land -- and keeps at most 8 reads in flight, which the
    // compiler's own wait counting handles without falling back to a full wait)
    for (int64_t st = 0; st + 1 < steps; ++st) {
        const int cur = int(st % S);
        __builtin_amdgcn_s_waitcnt(0xC07F);
        load_frags(cur, 1, fa[1], fb[1]);
        __builtin_amdgcn_sched_barrier(0);
        mma(fa[0], fb[0]);
        // steps issued so far: up to min(steps-1, st+S-1)
        const int64_t last = steps - 1 < st + S - 1 ? steps - 1 : st + S - 1;
        wait_loads(last - (st + 1));
        barrier();
        if (st + S < steps) issue(cur);
        load_frags(int((st + 1) % S), 0, fa[0], fb[0]);
        __builtin_amdgcn_sched_barrier(0);
        mma(fa[1], fb[1]);
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);
    load_frags(int((steps - 1) % S), 1, fa[1], fb[1]);
    mma(fa[0], fb[0]);
    mma(fa[1], fb[1]);

    const int scale = -7 * (d + 2);
#pragma unroll
    for (int p = 0; p < PM; ++p) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t row = m0 + wm * (BM / 2) + 32 * p + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (row >= M) continue;
            const int er = oz_exponent(ea[row]) + scale;
            double* crow = C + row * ldc;
#pragma unroll
            for (int c = 0; c < PN; ++c) {
                const int64_t col = n0 + wn * (BN / 2) + 32 * c + (lane & 31);
                if (col < N) {
                    const double x = ldexp(double(acc[p][c][v]), er + oz_exponent(eb[col]));
                    if (op == MRM_DO_NONE) crow[col] = x;
                    else if (op == MRM_DO_ADD) crow[col] = crow[col] + x;
                    else crow[col] = crow[col] - x;
                }
            }
        }
    }
}

// 128 x 256 workgroup tiles, three stages (72 KiB of LDS), two workgroups per
// CU: measured 10% faster than 256 x 256 with four stages and one workgroup
// per CU, the extra occupancy hiding more of the L2 latency.
constexpr int OZ_PM = 2, OZ_PN = 4, OZ_STAGES = 3;
constexpr int OZ_TILE = 256;  // slices are padded to whole tiles of either side
static_assert(64 * OZ_PM <= OZ_TILE && 64 * OZ_PN <= OZ_TILE && OZ_TILE % (64 * OZ_PM) == 0, "tile padding");

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

// C (M x N) (op)= op(A) op(B) in FP64 through `slices` int8 slices per operand.
// Device pointers, byte strides; enqueued on `stream`; `work` holds
// mrm_ozaki_workspace(M, N, K, slices) bytes.
int mrm_launch_dgemm_ozaki(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double* C, int64_t ldc, int op, int slices, void* work, size_t work_bytes,
                           hipStream_t stream) {
    if (slices < 1 || slices > MRM_OZAKI_MAX_SLICES) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: slice count %d outside 1..%d", slices, MRM_OZAKI_MAX_SLICES);
        return MRM_BAD_ARG;
    }
    if (M < 0 || N < 0 || K < 0 || op < 0 || op > 2 || trans < 0 || trans > 2 || ((lda | ldb | ldc) & 7)) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: negative dimension, bad op/trans or stride not a multiple of 8 bytes");
        return MRM_BAD_ARG;
    }
    if (M == 0 || N == 0) return MRM_OK;
    if (K == 0) return mrm_launch_dgemm(trans, M, N, 0, A, lda, B, ldb, C, ldc, op, stream);
    if (M >= (int64_t(1) << 22) || N >= (int64_t(1) << 22) || K >= (int64_t(1) << 24)) {
        mrm_set_error(MRM_BAD_DIMS, "ozaki: M or N of 2^22 or more, or K of 2^24 or more (launch grid limits)");
        return MRM_BAD_DIMS;
    }
    const int64_t Mp = round_up(M, OZ_TILE), Np = round_up(N, OZ_TILE), Kp = round_up(K, OZ_BK);
    const size_t need = mrm_ozaki_workspace(M, N, K, slices);
    if (!work || work_bytes < need) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: workspace of %zu bytes needed, %zu given", need, work_bytes);
        return MRM_BAD_ARG;
    }
    const int64_t ela = lda / 8, elb = ldb / 8, ec = ldc / 8;
    char* w = static_cast<char*>(work);
    int* ea = reinterpret_cast<int*>(w);
    int* eb = ea + Mp;
    int8_t* SA = reinterpret_cast<int8_t*>(w + round_up((Mp + Np) * 4, 256));
    int8_t* SB = SA + slices * Mp * Kp;
    const int64_t spa = Mp * Kp, spb = Np * Kp;
    // logical A(i, k) = A[i*ars + k*aks] and B^T(j, k) = B[j*brs + k*bks] (element strides)
    int64_t ars, aks, brs, bks;
    if (trans == MRM_TN) { ars = 1; aks = ela; } else { ars = ela; aks = 1; }
    if (trans == MRM_NT) { brs = elb; bks = 1; } else { brs = 1; bks = elb; }
    const unsigned kpieces = unsigned((K + 255) / 256);
    if (aks != 1 || bks != 1)
        hipLaunchKernelGGL(oz_zero_kernel, dim3(unsigned((Mp + Np + 255) / 256)), dim3(256), 0, stream, ea, Mp + Np);
    if (aks == 1) hipLaunchKernelGGL(oz_exp_rows_kernel, dim3(unsigned(M)), dim3(256), 0, stream, A, ars, M, K, ea);
    else hipLaunchKernelGGL(oz_exp_cols_kernel, dim3(unsigned((M + 63) / 64), kpieces), dim3(256), 0, stream, A, aks,
                            M, K, ea);
    if (bks == 1) hipLaunchKernelGGL(oz_exp_rows_kernel, dim3(unsigned(N)), dim3(256), 0, stream, B, brs, N, K, eb);
    else hipLaunchKernelGGL(oz_exp_cols_kernel, dim3(unsigned((N + 63) / 64), kpieces), dim3(256), 0, stream, B, bks,
                            N, K, eb);
    // slices; padding rows up to Mp / Np and k up to Kp are written as zero
    hipLaunchKernelGGL(oz_slice_kernel, dim3(unsigned(Kp / 64), unsigned(Mp / 64)), dim3(256), 0, stream, A, ars, aks,
                       M, K, ea, SA, Mp, spa, slices);
    hipLaunchKernelGGL(oz_slice_kernel, dim3(unsigned(Kp / 64), unsigned(Np / 64)), dim3(256), 0, stream, B, brs, bks,
                       N, K, eb, SB, Np, spb, slices);
    // per K chunk, diagonals largest first; the very first launch applies op,
    // every later one accumulates with the sign op asks for
    const int tiles_m = int(Mp / (64 * OZ_PM)), tiles_n = int(Np / (64 * OZ_PN));
    const int64_t ksteps = Kp / OZ_BK, chunk = MRM_OZAKI_KCHUNK / OZ_BK;
    bool first = true;
    for (int64_t k0 = 0; k0 < ksteps; k0 += chunk) {
        const int64_t nk = ksteps - k0 < chunk ? ksteps - k0 : chunk;
        for (int d = 0; d < slices; ++d) {
            const int o = first ? op : (op == MRM_DO_SUB ? MRM_DO_SUB : MRM_DO_ADD);
            first = false;
            hipLaunchKernelGGL((oz_gemm_kernel<OZ_PM, OZ_PN, OZ_STAGES>), dim3(unsigned(tiles_m) * unsigned(tiles_n)),
                               dim3(OZ_THREADS), 0, stream, SA, SB, Mp * OZ_BK, Np * OZ_BK, spa, spb, k0, nk, d, ea, eb, C, ec, M, N, o,
                               tiles_m, tiles_n);
        }
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

size_t mrm_ozaki_workspace(int64_t M, int64_t N, int64_t K, int slices) {
    const int64_t Mp = round_up(M, OZ_TILE), Np = round_up(N, OZ_TILE), Kp = round_up(K, OZ_BK);
    return size_t(round_up((Mp + Np) * 4, 256)) + size_t(slices) * size_t(Mp + Np) * size_t(Kp);
}
