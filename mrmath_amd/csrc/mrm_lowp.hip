// mrm_lowp.hip -- low-precision trailing update of the mixed-precision LU:
// C (FP64) -= A * B with A (M x K) and B (K x N) rounded to FP32 or BF16 and
// multiplied on the matrix cores with FP32 accumulation.  The split-BF16 mode
// rounds each operand to a high BF16 piece and its remainder to a low one,
// x = xh + xl + O(2^-16 |x|), and forms Ah Bh + Ah Bl + Al Bh (Al Bl, about
// 2^-18 |A||B|, is dropped) as one BF16 product over three times K: the
// pieces are packed one after the other along K, A as [Ah | Ah | Al] and B as
// [Bh ; Bl ; Bh].
//
// The matrix being factored stays in FP64 (the panel, the pivot search and the
// residuals of the refinement are FP64); only the Schur-complement product
// that carries almost all of the factorisation's flops,
// A22 -= L21 * U12 (reference src/LinAlgLU.pas:211-227, there in FP64 through
// the blocked multiply), is computed in the low precision.
//
// Operands are first packed into 64-byte k blocks, [k/64 bytes][row][64 bytes]
// (16 FP32 or 32 BF16 values of one row), A by rows and B by columns, rows
// padded to 256 rows and k to the block with zeros; one K step of a 256-row
// panel is then 16 KiB of contiguous memory, moved by direct global->LDS loads
// through a 3-stage ring.  Workgroup tile 128 x 256 (two per CU, so one
// workgroup's FP64 read-modify-write of C overlaps the other's products), 4
// waves of 64 x 128 (2 x 4 blocks of 32 x 32 on v_mfma_f32_32x32x2_f32 or
// v_mfma_f32_32x32x16_bf16).

#include <cstdint>
#include <type_traits>

#include "mrm_common.h"
#include "mrm_lowp.h"

namespace {

typedef float mrm_f16v __attribute__((ext_vector_type(16)));
typedef int mrm_i4 __attribute__((ext_vector_type(4)));
typedef __bf16 mrm_bf8 __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(1))) void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

constexpr int LP_BK = 64;       // bytes per row per K step
constexpr int LP_THREADS = 256;
constexpr int LP_TILE = 256;
constexpr int LP_STAGES = 3;
constexpr int LP_PM = 2, LP_PN = 4;  // 128 x 256 tiles, two workgroups per CU

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Packs X (logical R x K, element (r, k) at X[r*rs + k*ks]) into 64-byte k
// blocks.  64 rows x one k block per workgroup; reads coalesce along whichever
// index is contiguous.
// PIECE (BF16): 0 the value rounded, 1 its low piece x - bf16(x) rounded (the
// high piece is PIECE 0).
template <int MODE, int PIECE = 0>
__global__ void __launch_bounds__(256) lp_pack_kernel(const double* __restrict__ X, int64_t rs, int64_t ks, int64_t R,
                                                      int64_t K, uint8_t* __restrict__ P, int64_t Rp) {
    constexpr int EPB = MODE == MRM_LOWP_FP32 ? 16 : 32;  // elements per 64-byte block row
    constexpr int ESZ = 64 / EPB;
    __shared__ __attribute__((aligned(16))) uint8_t tile[64][64 + 16];
    const int64_t r0 = int64_t(blockIdx.y) * 64, k0 = int64_t(blockIdx.x) * EPB;
    const bool kfast = ks == 1;
    for (int q = threadIdx.x; q < 64 * EPB; q += 256) {
        const int rr = kfast ? q / EPB : q % 64;
        const int kk = kfast ? q % EPB : q / 64;
        const int64_t r = r0 + rr, k = k0 + kk;
        const double x = (r < R && k < K) ? X[r * rs + k * ks] : 0.0;
        if (MODE == MRM_LOWP_FP32) {
            *reinterpret_cast<float*>(&tile[rr][kk * ESZ]) = float(x);
        } else {
            const uint16_t h = to_bf16(float(x));
            *reinterpret_cast<uint16_t*>(&tile[rr][kk * ESZ]) =
                PIECE == 0 ? h : to_bf16(float(x - double(__uint_as_float(uint32_t(h) << 16))));
        }
    }
    __syncthreads();
    const int rr = threadIdx.x / 4, c = threadIdx.x % 4;
    const int4 v = *reinterpret_cast<const int4*>(&tile[rr][16 * c]);
    *reinterpret_cast<int4*>(P + int64_t(blockIdx.x) * Rp * 64 + (r0 + rr) * 64 + 16 * c) = v;
}

__device__ __forceinline__ int lp_slot(int r, int c) { return c ^ ((r >> 2) & 3); }

// Instruction order of one half step: MFMA, then up to one global load and up
// to one fragment read, repeated; the last MFMA takes what is left.
template <int NV, int NR, int NM>
__device__ __forceinline__ void lp_interleave() {
    if constexpr (NM > 0) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        constexpr int V = NM == 1 ? NV : (NV > 0 ? 1 : 0);
        constexpr int R = NM == 1 ? NR : (NR > 0 ? 1 : 0);
        if constexpr (V > 0) __builtin_amdgcn_sched_group_barrier(0x010, V, 0);  // VMEM
        if constexpr (R > 0) __builtin_amdgcn_sched_group_barrier(0x100, R, 0);  // DS read
        lp_interleave<NV - V, NR - R, NM - 1>();
    }
}

// XCD-aware tile order (consecutive workgroups go to different XCDs; each XCD
// walks a contiguous run of tiles in groups of 4 tile rows).
__device__ __forceinline__ void lp_tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
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

// C (M x N, FP64, element stride ldc) -= PA * PB^T over `steps` 64-byte k
// blocks.  Fragment of one 16-byte chunk per lane: lane l supplies row l%32 of
// a 32-row block and chunk 2h + l/32 of the block row (FP32: 4 values, one per
// v_mfma_f32_32x32x2_f32; BF16: 8 values, one v_mfma_f32_32x32x16_bf16).
// Result register v of lane l is row 8(v/4) + 4(l/32) + v%4, column l%32.
template <int MODE>
__global__ void __launch_bounds__(LP_THREADS, 2)
lp_update_kernel(const uint8_t* __restrict__ PA, const uint8_t* __restrict__ PB, int64_t kbpa, int64_t kbpb,
                 int64_t steps, double* __restrict__ C, int64_t ldc, int64_t M, int64_t N, int tiles_m, int tiles_n) {
    constexpr int PM = LP_PM, PN = LP_PN, S = LP_STAGES;
    constexpr int BM = 64 * PM, BN = 64 * PN;
    constexpr int LA = BM / 64, LB = BN / 64;
    constexpr int STAGE = (BM + BN) * LP_BK;
    __shared__ __attribute__((aligned(16))) uint8_t smem[S * STAGE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    lp_tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;

    const int lrow = lane >> 2;
    const int gchunk = (lane & 3) ^ ((lane >> 4) & 3);
    const uint8_t* pa = PA + (m0 + wave * (BM / 4) + lrow) * LP_BK + 16 * gchunk;
    const uint8_t* pb = PB + (n0 + wave * (BN / 4) + lrow) * LP_BK + 16 * gchunk;
    auto issue = [&](int stage) {
        uint8_t* la = smem + stage * STAGE + wave * (BM / 4) * LP_BK;
        uint8_t* lb = smem + stage * STAGE + BM * LP_BK + wave * (BN / 4) * LP_BK;
#pragma unroll
        for (int i = 0; i < LA; ++i)
            __builtin_amdgcn_global_load_lds((gptr_t)(pa + 16 * i * LP_BK), (lptr_t)(la + 16 * i * LP_BK), 16, 0, 0);
#pragma unroll
        for (int i = 0; i < LB; ++i)
            __builtin_amdgcn_global_load_lds((gptr_t)(pb + 16 * i * LP_BK), (lptr_t)(lb + 16 * i * LP_BK), 16, 0, 0);
        pa += kbpa;
        pb += kbpb;
    };

    mrm_f16v acc[PM][PN];
#pragma unroll
    for (int p = 0; p < PM; ++p)
#pragma unroll
        for (int c = 0; c < PN; ++c) acc[p][c] = mrm_f16v{};

    const int frow = lane & 31, fchunk = lane >> 5;
    mrm_i4 fa[2][PM], fb[2][PN];
    auto load_frags = [&](int stage, int h, mrm_i4* a, mrm_i4* b) {
        const uint8_t* As = smem + stage * STAGE;
        const uint8_t* Bs = As + BM * LP_BK;
#pragma unroll
        for (int c = 0; c < PN; ++c) {
            const int r = wn * (BN / 2) + 32 * c + frow;
            b[c] = *reinterpret_cast<const mrm_i4*>(Bs + r * LP_BK + 16 * lp_slot(r, 2 * h + fchunk));
        }
#pragma unroll
        for (int p = 0; p < PM; ++p) {
            const int r = wm * (BM / 2) + 32 * p + frow;
            a[p] = *reinterpret_cast<const mrm_i4*>(As + r * LP_BK + 16 * lp_slot(r, 2 * h + fchunk));
        }
    };
    auto mma = [&](const mrm_i4* a, const mrm_i4* b) {
#pragma unroll
        for (int p = 0; p < PM; ++p)
#pragma unroll
            for (int c = 0; c < PN; ++c) {
                if constexpr (MODE == MRM_LOWP_FP32) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        acc[p][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(__int_as_float(a[p][e]),
                                                                          __int_as_float(b[c][e]), acc[p][c], 0, 0, 0);
                } else {
                    acc[p][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mrm_bf8, a[p]),
                                                                         __builtin_bit_cast(mrm_bf8, b[c]), acc[p][c],
                                                                         0, 0, 0);
                }
            }
    };
    auto wait_loads = [&](int64_t later) {
        if (later >= S - 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((S - 1) * (LA + LB)) : "memory");
        else if (later == S - 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((S - 2) * (LA + LB)) : "memory");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    };
    auto barrier = [&]() {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's LDS reads are done
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };

    // Same ring as the INT8 slice GEMM: step st in stage st % S, each step's
    // second-half fragments load under its first-half MFMAs, the next step's
    // first half (and the refill of the stage just consumed) under its second
    // half, the loads and reads interleaved with the MFMAs.
#pragma unroll
    for (int i = 0; i < S; ++i)
        if (i < steps) issue(i);
    wait_loads(steps - 1 < S - 1 ? steps - 1 : S - 1);
    barrier();
    load_frags(0, 0, fa[0], fb[0]);
    constexpr int NM = PM * PN * (MODE == MRM_LOWP_FP32 ? 4 : 1);  // MFMAs per half step
    auto step = [&](int64_t st, auto refill) {
        constexpr bool kRefill = decltype(refill)::value;
        const int cur = int(st % S);
        __builtin_amdgcn_s_waitcnt(0xC07F);
        load_frags(cur, 1, fa[1], fb[1]);
        mma(fa[0], fb[0]);
        lp_interleave<0, PM + PN, NM>();
        const int64_t last = steps - 1 < st + S - 1 ? steps - 1 : st + S - 1;
        wait_loads(last - (st + 1));
        barrier();
        if constexpr (kRefill) issue(cur);
        load_frags(int((st + 1) % S), 0, fa[0], fb[0]);
        mma(fa[1], fb[1]);
        lp_interleave<kRefill ? LA + LB : 0, PM + PN, NM>();
    };
    int64_t st = 0;
    for (; st + S < steps; ++st) step(st, std::true_type());
    for (; st + 1 < steps; ++st) step(st, std::false_type());
    __builtin_amdgcn_s_waitcnt(0xC07F);
    load_frags(int((steps - 1) % S), 1, fa[1], fb[1]);
    mma(fa[0], fb[0]);
    mma(fa[1], fb[1]);

#pragma unroll
    for (int p = 0; p < PM; ++p) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t row = m0 + wm * (BM / 2) + 32 * p + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (row >= M) continue;
            double* crow = C + row * ldc;
#pragma unroll
            for (int c = 0; c < PN; ++c) {
                const int64_t col = n0 + wn * (BN / 2) + 32 * c + (lane & 31);
                if (col < N) crow[col] -= double(acc[p][c][v]);
            }
        }
    }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

size_t mrm_lowp_workspace(int64_t M, int64_t N, int64_t K, int mode) {
    const int64_t epb = mode == MRM_LOWP_FP32 ? 16 : 32;
    const int64_t kb = (K + epb - 1) / epb * (mode == MRM_LOWP_BF16X3 ? 3 : 1);
    return size_t(kb) * size_t(round_up(M, LP_TILE) + round_up(N, LP_TILE)) * 64;
}

// C (M x N) -= A * B with A(i, k) = A[i*lda + k], B(k, j) = B[k*ldb + j]
// (element strides), all FP64 in device memory; the product in `mode`.
int mrm_launch_lowp_update(int mode, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double* C, int64_t ldc, void* work, size_t work_bytes, hipStream_t stream) {
    if (mode != MRM_LOWP_FP32 && mode != MRM_LOWP_BF16 && mode != MRM_LOWP_BF16X3) {
        mrm_set_error(MRM_BAD_ARG, "lowp update: unknown mode %d", mode);
        return MRM_BAD_ARG;
    }
    if (M <= 0 || N <= 0 || K <= 0) return MRM_OK;
    if (M >= (int64_t(1) << 22) || N >= (int64_t(1) << 22) || K >= (int64_t(1) << 20)) {
        mrm_set_error(MRM_BAD_DIMS, "lowp update: dimensions beyond the launch grid limits");
        return MRM_BAD_DIMS;
    }
    if (!work || work_bytes < mrm_lowp_workspace(M, N, K, mode)) {
        mrm_set_error(MRM_BAD_ARG, "lowp update: workspace too small");
        return MRM_BAD_ARG;
    }
    const int64_t epb = mode == MRM_LOWP_FP32 ? 16 : 32;
    const int64_t kb1 = (K + epb - 1) / epb;  // k blocks of one piece
    const int64_t kb = kb1 * (mode == MRM_LOWP_BF16X3 ? 3 : 1);
    const int64_t Mp = round_up(M, LP_TILE), Np = round_up(N, LP_TILE);
    uint8_t* PA = static_cast<uint8_t*>(work);
    uint8_t* PB = PA + kb * Mp * 64;
    const dim3 ga(unsigned(kb1), unsigned(Mp / 64)), gb(unsigned(kb1), unsigned(Np / 64));
    if (mode == MRM_LOWP_FP32) {
        hipLaunchKernelGGL(lp_pack_kernel<MRM_LOWP_FP32>, ga, dim3(256), 0, stream, A, lda, int64_t(1), M, K, PA, Mp);
        hipLaunchKernelGGL(lp_pack_kernel<MRM_LOWP_FP32>, gb, dim3(256), 0, stream, B, int64_t(1), ldb, N, K, PB, Np);
    } else if (mode == MRM_LOWP_BF16) {
        hipLaunchKernelGGL(lp_pack_kernel<MRM_LOWP_BF16>, ga, dim3(256), 0, stream, A, lda, int64_t(1), M, K, PA, Mp);
        hipLaunchKernelGGL(lp_pack_kernel<MRM_LOWP_BF16>, gb, dim3(256), 0, stream, B, int64_t(1), ldb, N, K, PB, Np);
    } else {
        // three pieces along K, so that one BF16 product over 3K forms
        // Ah Bh + Ah Bl + Al Bh: A as [Ah | Ah | Al], B as [Bh ; Bl ; Bh]
        const int64_t sa = kb1 * Mp * 64, sb = kb1 * Np * 64;
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 0>), ga, dim3(256), 0, stream, A, lda, int64_t(1), M, K, PA,
                           Mp);
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 0>), ga, dim3(256), 0, stream, A, lda, int64_t(1), M, K,
                           PA + sa, Mp);
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 1>), ga, dim3(256), 0, stream, A, lda, int64_t(1), M, K,
                           PA + 2 * sa, Mp);
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 0>), gb, dim3(256), 0, stream, B, int64_t(1), ldb, N, K, PB,
                           Np);
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 1>), gb, dim3(256), 0, stream, B, int64_t(1), ldb, N, K,
                           PB + sb, Np);
        hipLaunchKernelGGL((lp_pack_kernel<MRM_LOWP_BF16, 0>), gb, dim3(256), 0, stream, B, int64_t(1), ldb, N, K,
                           PB + 2 * sb, Np);
    }
    const int tiles_m = int(Mp / (64 * LP_PM)), tiles_n = int(Np / (64 * LP_PN));
    const dim3 grid(unsigned(tiles_m) * unsigned(tiles_n));
    if (mode == MRM_LOWP_FP32)
        hipLaunchKernelGGL(lp_update_kernel<MRM_LOWP_FP32>, grid, dim3(LP_THREADS), 0, stream, PA, PB, Mp * 64,
                           Np * 64, kb, C, ldc, M, N, tiles_m, tiles_n);
    // BF16 and split BF16: one BF16 product over kb k blocks
    else
        hipLaunchKernelGGL(lp_update_kernel<MRM_LOWP_BF16>, grid, dim3(LP_THREADS), 0, stream, PA, PB, Mp * 64,
                           Np * 64, kb, C, ldc, M, N, tiles_m, tiles_n);
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}
