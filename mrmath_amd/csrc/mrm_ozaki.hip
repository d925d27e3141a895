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
// v of lane l is row 8(v/4) + 4(l/32) + v%4, column l%32).  The same digits
// held as BF16 run on v_mfma_f32_32x32x16_bf16 with the same layouts and exact
// FP32 partial sums folded into the int32 ones (MRM_OZAKI_BF16): the same
// result bit for bit.
//
// Slices are stored in 32-byte k blocks, [slice][k block][row][32 bytes] (32
// INT8 or 16 BF16 digits a block), K padded to a multiple of 64 and rows to a
// multiple of the 128-row tile with zeros, for A by rows and for B by columns:
// one K step of a 128-row panel is 4 KiB of contiguous memory, and the GEMM
// kernel has no edge cases except in its epilogue.  One launch forms the slice
// products of up to four diagonals (oz_fused_kernel).

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "mrm_common.h"
#include "mrm_host.h"
#include "mrm_ozaki.h"

namespace {

typedef int mrm_i4 __attribute__((ext_vector_type(4)));
typedef int mrm_i16 __attribute__((ext_vector_type(16)));
typedef float mrm_f16 __attribute__((ext_vector_type(16)));
typedef __bf16 mrm_bf8 __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(1))) void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

constexpr int OZ_KB = 32;                 // bytes (= int8 k-values) per k block of the slice layout
constexpr int OZ_KPAD = 64;               // K is padded to a multiple of this

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
// Slicing: a 64 (r) x 64 (k) tile per workgroup; S[t][kb][r][32 bytes] with
// kb the 32-byte k block (32 INT8 digits, or 16 BF16 digits when BF), k-block
// pitch Rp*32 bytes, slice pitch spitch bytes.  Padding rows / k beyond R / K
// are written as 0.  Each thread cuts 4 consecutive k of one row and stores
// their 4 digits per slice as one 32-bit (INT8) or 64-bit (BF16) word; a
// digit |d| <= 127 is exact in BF16 (8 significant bits), its bits are those
// of the FP32 value's upper half.
// ---------------------------------------------------------------------------
template <bool BF>
__global__ void __launch_bounds__(256) oz_slice_kernel(const double* __restrict__ X, int64_t rs, int64_t ks, int64_t R,
                                                       int64_t K, const int* __restrict__ ev, int8_t* __restrict__ S,
                                                       int64_t Rp, int64_t spitch, int s) {
    constexpr int DB = BF ? 2 : 1;          // bytes per digit
    constexpr int ROWB = 64 * DB + 16;      // tile row: 64 digits + padding
    constexpr int KBLK = 64 * DB / OZ_KB;   // 32-byte k blocks per 64 k
    __shared__ __attribute__((aligned(16))) uint8_t tile[MRM_OZAKI_MAX_SLICES][64][ROWB];
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
            uint64_t w = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                y[i] *= 128.0;
                const double d = trunc(y[i]);
                y[i] -= d;
                if (BF) w |= uint64_t(__float_as_uint(float(d)) >> 16) << (16 * i);
                else w |= uint64_t(uint8_t(int8_t(d))) << (8 * i);
            }
            if (BF) *reinterpret_cast<uint64_t*>(&tile[t][rr][8 * kq]) = w;
            else *reinterpret_cast<uint32_t*>(&tile[t][rr][4 * kq]) = uint32_t(w);
        }
    }
    __syncthreads();
    // per slice KBLK k blocks of 64 rows x 32 bytes, each one contiguous 2 KiB run
    const int64_t kb0 = k0 * DB / OZ_KB;
    for (int q = threadIdx.x; q < s * 64 * KBLK * 2; q += 256) {
        const int t = q / (64 * KBLK * 2), ch = q & 1, rr = (q / 2) % 64, kb = (q / 128) % KBLK;
        const int64_t r = r0 + rr;
        const int4 v = *reinterpret_cast<const int4*>(&tile[t][rr][32 * kb + 16 * ch]);
        *reinterpret_cast<int4*>(S + t * spitch + (kb0 + kb) * Rp * OZ_KB + r * OZ_KB + 16 * ch) = v;
    }
}

// ---------------------------------------------------------------------------
// Several diagonals in one launch: diagonals D0 .. D0+ND-1 of one K chunk.
// Every pair (t, u) with D0 <= t + u < D0 + ND uses slices t, u < T = D0 + ND
// of either side, so one K step loads the T slice panels of A and of B once
// and forms all the pairs from LDS, with one int32 accumulator set per
// diagonal (exact: |sum| <= (d+1) KCHUNK 127^2 < 2^31).  Workgroup tile 128 x 128, eight
// waves in 4 x 2 of 32 x 64 (1 x 2 blocks of 32 x 32) or four in 2 x 2 of
// 64 x 64 (2 x 2 blocks); K step 32 bytes (one
// v_mfma_i32_32x32x32_i8 k extent), NSTAGE-stage LDS ring
// [stage][A slices, B slices][128 rows][32 B], the 16-byte chunk c of row r at
// slot c ^ ((r/8) % 2) so a fragment read of 16 consecutive rows covers all
// banks.  The epilogue adds the ND diagonals into C in order, with the same
// FP64 roundings as one launch per diagonal.
// ---------------------------------------------------------------------------
constexpr bool oz_pair(int D0, int ND, int t, int u) { return t + u >= D0 && t + u < D0 + ND; }
// largest UH with at least half of the group's pairs at u >= UH (half 1, the
// half computed after the barrier, while the global loads and the next
// step's half-0 reads issue)
constexpr int oz_split_u(int D0, int ND) {
    const int T = D0 + ND;
    int total = 0;
    for (int u = 0; u < T; ++u)
        for (int t = 0; t < T; ++t) total += oz_pair(D0, ND, t, u) ? 1 : 0;
    int acc = 0;
    for (int u = T - 1; u > 0; --u) {
        for (int t = 0; t < T; ++t) acc += oz_pair(D0, ND, t, u) ? 1 : 0;
        if (2 * acc >= total) return u;
    }
    return 1;
}
// slice t of A takes part in a pair of half h (0: u < UH, 1: u >= UH)
constexpr bool oz_a_used(int D0, int ND, int UH, int h, int t) {
    for (int u = h ? UH : 0; u < (h ? D0 + ND : UH); ++u)
        if (oz_pair(D0, ND, t, u)) return true;
    return false;
}
// fragment reads (ds_read_b128) and MFMAs of half h for PBA 32-row A blocks
constexpr int oz_half_reads(int D0, int ND, int PBA, int h) {
    const int UH = oz_split_u(D0, ND), T = D0 + ND;
    int r = 0;
    for (int t = 0; t < T; ++t) r += oz_a_used(D0, ND, UH, h, t) ? PBA : 0;
    for (int u = 0; u < T; ++u) r += ((u < UH) == (h == 0)) ? 2 : 0;
    return r;
}
constexpr int oz_half_mfmas(int D0, int ND, int PBA, int h) {
    const int UH = oz_split_u(D0, ND), T = D0 + ND;
    int m = 0;
    for (int u = 0; u < T; ++u)
        if ((u < UH) == (h == 0))
            for (int t = 0; t < T; ++t) m += oz_pair(D0, ND, t, u) ? 2 * PBA : 0;
    return m;
}
// Instruction order of one half: MFMA, then up to one global load and up to
// two fragment reads, repeated; the last MFMA takes what is left.  Issuing
// the reads and loads as one burst ahead of the MFMAs stalled the waves on a
// full LDS queue with the matrix pipe idle (8-12% slower per group at 8192^3).
template <int NV, int NR, int NM>
__device__ __forceinline__ void oz_interleave() {
    if constexpr (NM > 0) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        constexpr int V = NM == 1 ? NV : (NV > 0 ? 1 : 0);
        constexpr int R = NM == 1 ? NR : (NR < 2 ? NR : 2);
        if constexpr (V > 0) __builtin_amdgcn_sched_group_barrier(0x010, V, 0);  // VMEM
        if constexpr (R > 0) __builtin_amdgcn_sched_group_barrier(0x100, R, 0);  // DS read
        oz_interleave<NV - V, NR - R, NM - 1>();
    }
}

// BF: the slices are BF16 digits (16 per 32-byte k block) multiplied on
// v_mfma_f32_32x32x16_bf16 with the same fragment and result layouts as the
// INT8 instruction; the FP32 sums of 8 K steps (128 k) are exact
// (|sum| <= 8 * 128 * 127^2 < 2^24 for diagonals up to 7) and are added into
// the int32 accumulators after every 8 steps, so the int32 diagonal sums, and
// the result, equal those of the INT8 slices bit for bit.
template <int D0, int ND, int NSTAGE, int NWAVE, bool BF>
__global__ void __launch_bounds__(NWAVE * 64, 1)
oz_fused_kernel(const int8_t* __restrict__ SA, const int8_t* __restrict__ SB, int64_t kbpa, int64_t kbpb, int64_t spa,
                int64_t spb, int64_t kstep0, int64_t nk, const int* __restrict__ ea, const int* __restrict__ eb,
                double* __restrict__ C, int64_t ldc, int64_t M, int64_t N, int op, int tiles_m, int tiles_n) {
    static_assert(NWAVE == 4 || NWAVE == 8, "2 x 2 waves of 64 x 64, or 4 x 2 of 32 x 64");
    constexpr int T = D0 + ND;              // slices 0..T-1 of each operand take part
    constexpr int RB = OZ_KB;               // bytes of k per row and step
    constexpr int PANEL = 128 * RB;         // one slice of one operand, one step
    constexpr int STAGE = 2 * T * PANEL;
    constexpr int PBA = NWAVE == 4 ? 2 : 1; // 32-row A blocks per wave (B: two 32-column blocks)
    constexpr int LPS = NWAVE == 4 ? 2 * T : T;  // direct loads per thread and step
    __shared__ __attribute__((aligned(16))) int8_t smem[NSTAGE * STAGE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = NWAVE == 4 ? (wave & 1) : (wave & 3), wn = NWAVE == 4 ? (wave >> 1) : (wave >> 2);
    int tm, tn;
    {
        // workgroups are dealt round-robin to the 8 XCDs: XCD x gets a
        // contiguous run of logical tiles, walked in groups of 8 tile rows so
        // the workgroups resident on one XCD share slice panels in its L2
        constexpr int NXCD = 8, GROUP_M = 8;
        const int bid = blockIdx.x, total = tiles_m * tiles_n;
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
    const int64_t m0 = int64_t(tm) * 128, n0 = int64_t(tn) * 128;

    // direct loads: one instruction of a wave writes rows 32q .. 32q+31 of a
    // panel (1 KiB), q = wave % 4; lane L covers row 32q + L/2, LDS slot L%2 <-
    // HBM chunk (L%2) ^ ((L/16)%2) (= the slot rule for that row).  Four waves:
    // each loads its rows of every panel; eight: waves 0-3 the A panels, 4-7
    // the B panels.
    const int lq = wave & 3;
    const int lrow = 32 * lq + (lane >> 1);
    const int gch = (lane & 1) ^ ((lane >> 4) & 1);
    const int8_t* pa = SA + kstep0 * kbpa + (m0 + lrow) * OZ_KB + 16 * gch;
    const int8_t* pb = SB + kstep0 * kbpb + (n0 + lrow) * OZ_KB + 16 * gch;
    const int64_t steps = nk;

    auto issue = [&](int stage, int64_t st) {
        const int64_t oa = st * kbpa;
        const int64_t ob = st * kbpb;
        int8_t* l = smem + stage * STAGE + lq * (32 * RB);
        if (NWAVE == 4 || wave < 4) {
#pragma unroll
            for (int t = 0; t < T; ++t)
                __builtin_amdgcn_global_load_lds((gptr_t)(pa + t * spa + oa), (lptr_t)(l + t * PANEL), 16, 0, 0);
        }
        if (NWAVE == 4 || wave >= 4) {
#pragma unroll
            for (int u = 0; u < T; ++u)
                __builtin_amdgcn_global_load_lds((gptr_t)(pb + u * spb + ob), (lptr_t)(l + (T + u) * PANEL), 16, 0,
                                                 0);
        }
    };

    mrm_i16 acc[ND][PBA][2];
#pragma unroll
    for (int e = 0; e < ND; ++e)
#pragma unroll
        for (int p = 0; p < PBA; ++p)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[e][p][c] = mrm_i16{};
    constexpr int NDF = BF ? ND : 1;        // FP32 partial sums (BF16 slices only)
    mrm_f16 accf[NDF][PBA][2];
#pragma unroll
    for (int e = 0; e < NDF; ++e)
#pragma unroll
        for (int p = 0; p < PBA; ++p)
#pragma unroll
            for (int c = 0; c < 2; ++c) accf[e][p][c] = mrm_f16{};
    auto flush = [&]() {
#pragma unroll
        for (int e = 0; e < NDF; ++e)
#pragma unroll
            for (int p = 0; p < PBA; ++p)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
#pragma unroll
                    for (int v = 0; v < 16; ++v) acc[e][p][c][v] += int(accf[e][p][c][v]);
                    accf[e][p][c] = mrm_f16{};
                }
    };

    // fragment of block p (A) / c (B): lane l reads row l%32 of the block,
    // 16-byte chunk l/32
    const int frow = lane & 31, fch = lane >> 5;
    auto frag = [&](const int8_t* P, int r) {
        return *reinterpret_cast<const mrm_i4*>(P + r * RB + 16 * (fch ^ ((r >> 3) & 1)));
    };
    const int ra0 = wm * (32 * PBA) + frow, rb0 = wn * 64 + frow;

    // Wait until this thread's loads of the current step have landed while
    // those of `later` further steps stay in flight (LPS instructions a step).
    static_assert(NSTAGE >= 2 && NSTAGE <= 3 && 2 * LPS < 64, "vmcnt range");
    auto wait_loads = [&](int64_t later) {
        if (NSTAGE >= 3 && later >= 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * LPS) : "memory");
        else if (later >= 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(LPS) : "memory");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    };
    auto barrier = [&]() {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };

    // The pairs are split by u into two halves of about equal size (half 0:
    // u < UH).  A K step computes half 0 while half 1's fragments load, then
    // waits for the next step, refills the stage just consumed and loads the
    // next step's half-0 fragments while half 1 computes (the MFMAs keep
    // running across the barrier).
    constexpr int UH = oz_split_u(D0, ND);
    mrm_i4 fa[2][T][PBA], fb[2][T][2];
    auto load_half = [&](int stage, int h) {
        const int8_t* As = smem + stage * STAGE;
        const int8_t* Bs = As + T * PANEL;
#pragma unroll
        for (int t = 0; t < T; ++t)
            if (oz_a_used(D0, ND, UH, h, t)) {
#pragma unroll
                for (int p = 0; p < PBA; ++p) fa[h][t][p] = frag(As + t * PANEL, ra0 + 32 * p);
            }
#pragma unroll
        for (int u = 0; u < T; ++u)
            if ((u < UH) == (h == 0)) {
                fb[h][u][0] = frag(Bs + u * PANEL, rb0);
                fb[h][u][1] = frag(Bs + u * PANEL, rb0 + 32);
            }
    };
    auto mma_half = [&](int h) {
#pragma unroll
        for (int u = 0; u < T; ++u) {
            if ((u < UH) != (h == 0)) continue;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                if (!oz_pair(D0, ND, t, u)) continue;
                const int e = t + u - D0;
#pragma unroll
                for (int p = 0; p < PBA; ++p)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        if constexpr (BF)
                            accf[e][p][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                __builtin_bit_cast(mrm_bf8, fa[h][t][p]), __builtin_bit_cast(mrm_bf8, fb[h][u][c]),
                                accf[e][p][c], 0, 0, 0);
                        else
                            acc[e][p][c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[h][t][p], fb[h][u][c],
                                                                                 acc[e][p][c], 0, 0, 0);
                    }
            }
        }
    };

#pragma unroll
    for (int i = 0; i < NSTAGE; ++i)
        if (i < steps) issue(i, i);
    wait_loads(steps - 1 < NSTAGE - 1 ? steps - 1 : NSTAGE - 1);
    barrier();
    load_half(0, 0);
    constexpr int R0 = oz_half_reads(D0, ND, PBA, 0), R1 = oz_half_reads(D0, ND, PBA, 1);
    constexpr int M0 = oz_half_mfmas(D0, ND, PBA, 0), M1 = oz_half_mfmas(D0, ND, PBA, 1);
    // one K step; `refill` (compile time) reloads the stage just consumed
    // with step st + NSTAGE, so the steps that refill and the last NSTAGE - 1
    // that do not run as two loops without a branch inside
    auto step = [&](int64_t st, auto refill) {
        constexpr bool kRefill = decltype(refill)::value;
        const int cur = int(st % NSTAGE);
        __builtin_amdgcn_s_waitcnt(0xC07F);
        load_half(cur, 1);
        mma_half(0);
        oz_interleave<0, R1, M0>();
        // steps issued so far: up to min(steps-1, st+NSTAGE-1)
        const int64_t last = steps - 1 < st + NSTAGE - 1 ? steps - 1 : st + NSTAGE - 1;
        wait_loads(last - (st + 1));
        barrier();
        if constexpr (kRefill) issue(cur, st + NSTAGE);
        load_half(int((st + 1) % NSTAGE), 0);
        mma_half(1);
        oz_interleave<kRefill ? LPS : 0, R0, M1>();
        if constexpr (BF) {
            if ((st & 7) == 7) flush();
        }
    };
    int64_t st = 0;
    for (; st + NSTAGE < steps; ++st) step(st, std::true_type());
    for (; st + 1 < steps; ++st) step(st, std::false_type());
    __builtin_amdgcn_s_waitcnt(0xC07F);
    load_half(int((steps - 1) % NSTAGE), 1);
    mma_half(0);
    mma_half(1);
    if constexpr (BF) flush();

    // C (op)= diagonal D0, then (+/-)= D0+1, ...: the roundings of one launch
    // per diagonal.  Result register v of lane l is row 8(v/4) + 4(l/32) + v%4,
    // column l%32 of its block.
    const int later = op == MRM_DO_SUB ? MRM_DO_SUB : MRM_DO_ADD;
#pragma unroll
    for (int p = 0; p < PBA; ++p) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t row = m0 + wm * (32 * PBA) + 32 * p + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (row >= M) continue;
            const int er = oz_exponent(ea[row]);
            double* crow = C + row * ldc;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int64_t col = n0 + wn * 64 + 32 * c + (lane & 31);
                if (col >= N) continue;
                const int ec = er + oz_exponent(eb[col]);
                double x = op == MRM_DO_NONE ? 0.0 : crow[col];
#pragma unroll
                for (int e = 0; e < ND; ++e) {
                    const double y = ldexp(double(acc[e][p][c][v]), ec - 7 * (D0 + e + 2));
                    const int o = e == 0 ? op : later;
                    x = o == MRM_DO_NONE ? y : (o == MRM_DO_ADD ? x + y : x - y);
                }
                crow[col] = x;
            }
        }
    }
}

// Three LDS stages while they fit in the 160 KiB of a CU, else two.
constexpr int oz_fstages(int T) { return 3 * 2 * T * 128 * 32 <= 160 * 1024 ? 3 : 2; }

// Waves per workgroup: eight (32 x 64 each, two per SIMD; 1-6% faster per
// group at 8192^3) except for (3, 4), whose eight-wave form spills registers
// in its K loop (four waves of 64 x 64 time the same, 9.7 ms).
template <int D0, int ND>
constexpr int oz_waves() { return (D0 == 3 && ND == 4) ? 4 : 8; }

template <int D0, int ND, bool BF>
const void* oz_fused_fn() {
    return reinterpret_cast<const void*>(&oz_fused_kernel<D0, ND, oz_fstages(D0 + ND), oz_waves<D0, ND>(), BF>);
}

// Diagonal groups per slice count, one launch each.  A group of ND diagonals
// holds ND int32 accumulator sets of 64 registers per lane (four waves) or 32
// (eight), so ND <= 4 (five and six spill); within that, larger groups load
// fewer slice panels per pair.  Chosen from per-group times at 8192^3 (ms,
// four-wave kernels without the interleaved schedule): (0,4) 5.26 against
// (0,3)+(3,1) 6.65 and (0,2)+(2,2) 6.37; s = 5: (0,2)+(2,3) 8.38 against
// (0,3)+(3,2) 8.78; s = 6: (0,2)+(2,4) 11.11 against (0,3)+(3,3) 11.57 and
// (0,4)+(4,2) 11.94; s = 7: (0,3)+(3,4) 14.18 against (0,4)+(4,3) 14.52.
struct OzGroup {
    int d0, nd;
};
constexpr int OZ_MAX_GROUPS = 3;
constexpr OzGroup OZ_GROUPS[MRM_OZAKI_MAX_SLICES + 1][OZ_MAX_GROUPS] = {
    {},
    {{0, 1}},
    {{0, 2}},
    {{0, 3}},
    {{0, 4}},
    {{0, 2}, {2, 3}},
    {{0, 2}, {2, 4}},
    {{0, 3}, {3, 4}},
    {{0, 3}, {3, 4}, {7, 1}},
};
static_assert(MRM_OZAKI_MAX_SLICES == 8, "a group table row per slice count");

// BF16 slices: a diagonal holds an int32 and an FP32 accumulator set, so the
// groups have at most two diagonals (three spill, and so does (6,2) in its K
// loop: s = 8 ends with (6,1) and (7,1)).
constexpr int OZ_MAX_GROUPS_BF = 5;
constexpr OzGroup OZ_GROUPS_BF[MRM_OZAKI_MAX_SLICES + 1][OZ_MAX_GROUPS_BF] = {
    {},
    {{0, 1}},
    {{0, 2}},
    {{0, 2}, {2, 1}},
    {{0, 2}, {2, 2}},
    {{0, 2}, {2, 2}, {4, 1}},
    {{0, 2}, {2, 2}, {4, 2}},
    {{0, 2}, {2, 2}, {4, 2}, {6, 1}},
    {{0, 2}, {2, 2}, {4, 2}, {6, 1}, {7, 1}},
};

// The kernel of group (d0, nd) and its workgroup size.
const void* oz_fused_for(int d0, int nd, bool bf, int& threads) {
    if (bf) {
        threads = 512;
        switch (d0 * 8 + nd) {
#define OZ_CASE(a, b) \
    case a * 8 + b: return oz_fused_fn<a, b, true>();
            OZ_CASE(0, 1)
            OZ_CASE(0, 2)
            OZ_CASE(2, 1)
            OZ_CASE(2, 2)
            OZ_CASE(4, 1)
            OZ_CASE(4, 2)
            OZ_CASE(6, 1)
            OZ_CASE(7, 1)
#undef OZ_CASE
            default: return nullptr;
        }
    }
    switch (d0 * 8 + nd) {
#define OZ_CASE(a, b)                      \
    case a * 8 + b:                        \
        threads = 64 * oz_waves<a, b>();   \
        return oz_fused_fn<a, b, false>();
        OZ_CASE(0, 1)
        OZ_CASE(0, 2)
        OZ_CASE(0, 3)
        OZ_CASE(0, 4)
        OZ_CASE(2, 3)
        OZ_CASE(2, 4)
        OZ_CASE(3, 4)
        OZ_CASE(7, 1)
#undef OZ_CASE
        default: return nullptr;
    }
}

constexpr int OZ_TILE = 128;  // slices are padded to whole tiles of either side

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

// C (M x N) (op)= op(A) op(B) in FP64 through `slices` INT8 (kind
// MRM_OZAKI_INT8) or BF16 (MRM_OZAKI_BF16) slices per operand.  Device
// pointers, byte strides; enqueued on `stream`; `work` holds
// mrm_ozaki_workspace(M, N, K, slices, kind) bytes.
int mrm_launch_dgemm_ozaki(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double* C, int64_t ldc, int op, int slices, int kind, void* work,
                           size_t work_bytes, hipStream_t stream) {
    if (slices < 1 || slices > MRM_OZAKI_MAX_SLICES) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: slice count %d outside 1..%d", slices, MRM_OZAKI_MAX_SLICES);
        return MRM_BAD_ARG;
    }
    if (kind != MRM_OZAKI_INT8 && kind != MRM_OZAKI_BF16) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: slice kind %d is neither MRM_OZAKI_INT8 nor MRM_OZAKI_BF16", kind);
        return MRM_BAD_ARG;
    }
    const bool bf = kind == MRM_OZAKI_BF16;
    const int db = bf ? 2 : 1;  // bytes per digit
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
    const int64_t Mp = round_up(M, OZ_TILE), Np = round_up(N, OZ_TILE), Kp = round_up(K, OZ_KPAD);
    const size_t need = mrm_ozaki_workspace(M, N, K, slices, kind);
    if (!work || work_bytes < need) {
        mrm_set_error(MRM_BAD_ARG, "ozaki: workspace of %zu bytes needed, %zu given", need, work_bytes);
        return MRM_BAD_ARG;
    }
    const int64_t ela = lda / 8, elb = ldb / 8, ec = ldc / 8;
    char* w = static_cast<char*>(work);
    int* ea = reinterpret_cast<int*>(w);
    int* eb = ea + Mp;
    int8_t* SA = reinterpret_cast<int8_t*>(w + round_up((Mp + Np) * 4, 256));
    int8_t* SB = SA + slices * Mp * Kp * db;
    const int64_t spa = Mp * Kp * db, spb = Np * Kp * db;
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
    auto slice = bf ? oz_slice_kernel<true> : oz_slice_kernel<false>;
    hipLaunchKernelGGL(slice, dim3(unsigned(Kp / 64), unsigned(Mp / 64)), dim3(256), 0, stream, A, ars, aks, M, K, ea,
                       SA, Mp, spa, slices);
    hipLaunchKernelGGL(slice, dim3(unsigned(Kp / 64), unsigned(Np / 64)), dim3(256), 0, stream, B, brs, bks, N, K, eb,
                       SB, Np, spb, slices);
    // per K chunk, the diagonal groups of OZ_GROUPS[slices] (OZ_GROUPS_BF
    // for BF16 slices) in order; the very first diagonal applies op, every
    // later one accumulates with the sign op asks for.  A K step is one
    // 32-byte k block: 32 INT8 or 16 BF16 digits.
    const int tiles_m = int(Mp / 128), tiles_n = int(Np / 128);
    const int64_t ksteps = Kp * db / OZ_KB, chunk = MRM_OZAKI_KCHUNK * db / OZ_KB;
    const OzGroup* groups = bf ? OZ_GROUPS_BF[slices] : OZ_GROUPS[slices];
    const int max_groups = bf ? OZ_MAX_GROUPS_BF : OZ_MAX_GROUPS;
    int64_t kbpa = Mp * OZ_KB, kbpb = Np * OZ_KB;
    bool first = true;
    for (int64_t k0 = 0; k0 < ksteps; k0 += chunk) {
        int64_t nk = ksteps - k0 < chunk ? ksteps - k0 : chunk;
        int64_t kk = k0;
        for (int g = 0; g < max_groups && groups[g].nd > 0; ++g) {
            const int d0 = groups[g].d0, nd = groups[g].nd;
            int o = first ? op : (op == MRM_DO_SUB ? MRM_DO_SUB : MRM_DO_ADD);
            first = false;
            int threads = 0;
            const void* fn = oz_fused_for(d0, nd, bf, threads);
            if (!fn) {
                mrm_set_error(MRM_BAD_ARG, "ozaki: no kernel for diagonals %d..%d", d0, d0 + nd - 1);
                return MRM_BAD_ARG;
            }
            int64_t Mv = M, Nv = N, ecv = ec;
            int tmv = tiles_m, tnv = tiles_n;
            const int* eav = ea;
            const int* ebv = eb;
            const int8_t* sav = SA;
            const int8_t* sbv = SB;
            int64_t spav = spa, spbv = spb;
            double* cv = C;
            void* args[] = {&sav, &sbv, &kbpa, &kbpb, &spav, &spbv, &kk, &nk, &eav, &ebv, &cv, &ecv, &Mv, &Nv, &o, &tmv, &tnv};
            MRM_HIP_TRY(hipLaunchKernel(fn, dim3(unsigned(tiles_m) * unsigned(tiles_n)), dim3(unsigned(threads)),
                                        args, 0, stream));
        }
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

size_t mrm_ozaki_workspace(int64_t M, int64_t N, int64_t K, int slices, int kind) {
    const int64_t Mp = round_up(M, OZ_TILE), Np = round_up(N, OZ_TILE), Kp = round_up(K, OZ_KPAD);
    const size_t db = kind == MRM_OZAKI_BF16 ? 2 : 1;
    return size_t(round_up((Mp + Np) * 4, 256)) + db * size_t(slices) * size_t(Mp + Np) * size_t(Kp);
}
