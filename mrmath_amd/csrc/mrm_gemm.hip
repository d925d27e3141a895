// mrm_gemm.hip -- FP64 GEMM on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the reference's blocked FMA product (BlockMatrixMultiplication[T1/T2],
// src/BlockedMult.pas:350-724, and its micro-kernel FMAMatrixMultTransposed,
// src/FMAMatrixOperations.pas:55-92).  One kernel family covers NN, TN (MultT1)
// and NT (MultT2): the operand layout only changes how a 128x16 panel is
// moved from HBM into LDS and addressed there, so the tile transposes the
// reference performs per K-block (src/BlockedMult.pas:435, 560-561) disappear.
// The destination operation (doNone / doAdd / doSub, src/BlockedMult.pas:459-463)
// is fused into the epilogue.
//
// Workgroup: 256 threads = 4 waves in a 2x2 arrangement; workgroup tile
// BM x BN = 128 x 128, K step BK = 16; each wave owns a 64x64 block of C as
// 4x4 blocks of 16x16.  LDS holds two stages of both panels (two workgroups
// per CU).  Interior K steps move each panel straight from HBM into LDS with
// global_load_lds_dwordx4 (no register round trip, no LDS store
// instructions); the load of step t+1 is issued before step t is consumed and
// one barrier per step publishes it.  Edge tiles and partial K steps take a
// predicated register path, after the step's compute, that writes the same
// LDS layout.
//
// LDS panel layouts (r = the panel's M or N index):
//   KC (operand K-contiguous in HBM): [r][k], 16 doubles per row, the eight
//       16-byte chunks of row r stored at slot (chunk ^ (r % 8)), so that the
//       16-byte fragment reads of eight consecutive rows hit distinct banks;
//   RC (operand r-contiguous in HBM): [k][r], row stride SR = 136 doubles
//       (k-rows 2 apart start 32 banks apart: the two lane groups of a
//       64-lane 8-byte read touch disjoint banks).
// One direct-load instruction of a wave writes 1 KiB of LDS (64 lanes x 16 B):
// eight KC rows, or one RC k-row.
//
// Matrix instruction: v_mfma_f64_16x16x4_f64 (lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; result register v of lane l is row l/16 + 4v, column
// l%16).  Each A and B value feeds 16 products per instruction, four times
// the operand reuse of the four-block 4x4x4 form, so a K step reads 16 KiB of
// fragments per wave from LDS instead of 40.  Pure-MFMA loops reach 76.6
// TFLOP/s (97.5% of peak) with it, but only when the accumulators stay in
// VGPRs: in the compiler's default AGPR form every accumulator is copied
// between the register files around each instruction (36 TFLOP/s), so this
// file is compiled with -mllvm -amdgpu-mfma-vgpr-form.
//
// K ordering inside a step: the MFMA's k index is the lane group g = l/16; a
// step is two halves m, and in half m the first MFMA of a pair takes
// k = 8m + 2g, the second k = 8m + 2g + 1, for A and B alike, so every
// product A[m,k]*B[k,n] is formed exactly once and a lane's two k of a KC row
// are one 16-byte chunk (see frag_pair).  Both halves' fragments are read at
// the start of a step; the second half's land while the first half computes.

#include <type_traits>

#include "mrm_common.h"

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int NTHREADS = 256;
constexpr int SR = BM + 8;                          // RC panels: [k][r] row stride in doubles
constexpr int KC_PANEL = BM * BK;                   // KC panels: 128 rows x 16 doubles
constexpr int RC_PANEL = BK * SR;
constexpr int PANEL = KC_PANEL > RC_PANEL ? KC_PANEL : RC_PANEL;   // doubles per panel stage

typedef const __attribute__((address_space(1))) void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

__device__ __forceinline__ int kc_swizzle(int r) { return r & 7; }

// LDS element index of (r, k) in a panel.
template <bool KC>
__device__ __forceinline__ int panel_index(int r, int k) {
    if (KC) return r * BK + 2 * ((k >> 1) ^ kc_swizzle(r)) + (k & 1);
    return k * SR + r;
}

// The fragments of one K step are read as pairs: half step m (0, 1), lane
// group g = lane/16 takes k = 8m + 2g (.x) and 8m + 2g + 1 (.y), for A and B
// alike, so every product A[r,k]*B[k,n] is formed exactly once.  In a KC panel
// the pair is one 16-byte chunk (one ds_read_b128; eight consecutive rows hit
// distinct banks), in an RC panel two k-rows.
template <bool KC>
__device__ __forceinline__ double2 frag_pair(const double* __restrict__ P, int r, int m, int g) {
    if (KC) return *reinterpret_cast<const double2*>(P + r * BK + 2 * ((4 * m + g) ^ kc_swizzle(r)));
    double2 v;
    v.x = P[(8 * m + 2 * g) * SR + r];
    v.y = P[(8 * m + 2 * g + 1) * SR + r];
    return v;
}

// Moves one 128 x 16 panel per K step.  Wave w, instruction i, lane L:
//   KC: panel row r = 32w + 8i + L/8, LDS slot L%8 <- HBM chunk (L%8) ^ swz(r);
//       swz(r + 8) = swz(r), so every instruction reads the same chunk offset
//   RC: panel k-row 4w + i, rows r = 2L, 2L+1
// base points at the operand's element (0, 0); ld in elements.
template <bool KC, bool VEC>
struct PanelMover {
    const double* p;     // this lane's source for instruction 0 at the current step
    int64_t istep;       // element offset between instructions i and i+1
    int64_t kstep;       // element offset of one K step
    int64_t k0;          // K index of the current step (uniform)
    int64_t K;
    int lds0;            // LDS element of this lane's 16 bytes, instruction 0
    int lds_wave0;       // LDS element where this wave's instruction 0 writes (uniform)
    int ldsstep;         // LDS elements between instructions
    int koff;            // KC: this lane's first k; RC: this wave's first k-row
    bool rows_full;      // all 128 panel rows exist (uniform)
    unsigned rmask;      // KC: bit i = row of instruction i exists; RC: bit0/1 = row pair exists
    double2 v[4];        // register path staging

    __device__ __forceinline__ void init(const double* __restrict__ base, int64_t ld, int64_t r0, int64_t R,
                                         int64_t K_, int tid) {
        const int w = tid >> 6, L = tid & 63;
        K = K_;
        k0 = 0;
        rows_full = r0 + BM <= R;
        if (KC) {
            const int r = 32 * w + (L >> 3);
            const int slot = L & 7;
            koff = 2 * (slot ^ kc_swizzle(r));
            p = base + (r0 + r) * ld + koff;
            istep = 8 * ld;
            kstep = BK;
            lds0 = r * BK + 2 * slot;
            lds_wave0 = 32 * w * BK;
            ldsstep = 8 * BK;
            rmask = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) rmask |= (r0 + r + 8 * i < R) ? (1u << i) : 0u;
        } else {
            koff = 4 * w;
            p = base + int64_t(koff) * ld + r0 + 2 * L;
            istep = ld;
            kstep = int64_t(BK) * ld;
            lds0 = koff * SR + 2 * L;
            lds_wave0 = koff * SR;
            ldsstep = SR;
            rmask = (r0 + 2 * L < R ? 1u : 0u) | (r0 + 2 * L + 1 < R ? 2u : 0u);
        }
    }

    // Interior step of a 16-byte aligned operand: HBM -> LDS directly.
    __device__ __forceinline__ bool direct_ok() const { return VEC && rows_full && k0 + BK <= K; }

    __device__ __forceinline__ void issue_direct(double* __restrict__ P) const {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            __builtin_amdgcn_global_load_lds((gptr_t)(p + i * istep),
                                             (lptr_t)(P + lds_wave0 + i * ldsstep), 16, 0, 0);
    }

    // Predicated path: registers, then store().
    __device__ __forceinline__ void load_regs() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double* q = p + i * istep;
            bool ok0, ok1;
            if (KC) {
                const bool rok = (rmask >> i) & 1u;
                const int64_t k = k0 + koff;
                ok0 = rok && (k < K);
                ok1 = rok && (k + 1 < K);
            } else {
                const bool kok = k0 + koff + i < K;
                ok0 = kok && (rmask & 1u);
                ok1 = kok && (rmask & 2u);
            }
            double2 t;
            t.x = ok0 ? q[0] : 0.0;
            t.y = ok1 ? q[1] : 0.0;
            v[i] = t;
        }
    }

    __device__ __forceinline__ void store(double* __restrict__ P) const {
#pragma unroll
        for (int i = 0; i < 4; ++i) *reinterpret_cast<double2*>(P + lds0 + i * ldsstep) = v[i];
    }

    __device__ __forceinline__ void advance() {
        p += kstep;
        k0 += BK;
    }
};

__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Instruction order of one half step: MFMA, then up to one global load and up
// to one fragment read, repeated; the last MFMA takes what is left (issued as
// one burst after the barrier, the loads and reads held every wave's first
// MFMA back).
template <int NV, int NR, int NM>
__device__ __forceinline__ void gm_interleave() {
    if constexpr (NM > 0) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        constexpr int V = NM == 1 ? NV : (NV > 0 ? 1 : 0);
        constexpr int R = NM == 1 ? NR : (NR > 0 ? 1 : 0);
        if constexpr (V > 0) __builtin_amdgcn_sched_group_barrier(0x010, V, 0);  // VMEM
        if constexpr (R > 0) __builtin_amdgcn_sched_group_barrier(0x100, R, 0);  // DS read
        gm_interleave<NV - V, NR - R, NM - 1>();
    }
}

// Workgroup id -> (tile_m, tile_n).  Workgroups are dealt round-robin to the 8
// XCDs; remap so that each XCD receives a contiguous run of logical tiles, then
// walk that run in groups of GROUP_M tile-rows so the workgroups resident on
// one XCD share A and B panels in that XCD's L2.
// upper: only the tiles with tm <= tn of a square tile grid (a symmetric
// product), numbered row by row.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int upper, int& tm, int& tn) {
    constexpr int NXCD = 8;
    constexpr int GROUP_M = 8;
    const int total = upper ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    const int xcd = bid % NXCD;
    const int q = total / NXCD, rem = total % NXCD;
    const int logical = xcd * q + (xcd < rem ? xcd : rem) + bid / NXCD;
    if (upper) {
        int r = logical, row = 0;
        while (r >= tiles_m - row) {
            r -= tiles_m - row;
            ++row;
        }
        tm = row;
        tn = row + r;
        return;
    }
    const int per_group = GROUP_M * tiles_n;
    const int group = logical / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = (tiles_m - first_m) < GROUP_M ? (tiles_m - first_m) : GROUP_M;
    const int in_group = logical - group * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

// TA: A stored K x M (MultT1).  TB: B stored N x K (MultT2).  Strides in elements.
template <bool TA, bool TB, bool VEC>
__global__ void __launch_bounds__(NTHREADS, 2)
dgemm_mfma_kernel(int64_t M, int64_t N, int64_t K, const double* __restrict__ A, int64_t lda,
                  const double* __restrict__ B, int64_t ldb, double* __restrict__ C, int64_t ldc,
                  int op, int tiles_m, int tiles_n, int64_t kchunk, int64_t cstride, int upper, int64_t bsa,
                  int64_t bsb, int64_t bsc) {
    __shared__ __attribute__((aligned(16))) double smem[2 * 2 * PANEL];
    if (gridDim.z > 1) {
        // strided batch: product blockIdx.z
        A += int64_t(blockIdx.z) * bsa;
        B += int64_t(blockIdx.z) * bsb;
        C += int64_t(blockIdx.z) * bsc;
    }
    if (gridDim.y > 1) {
        // split K: chunk blockIdx.y of kchunk (a multiple of BK) writes its
        // partial product to its own C slice
        const int64_t k0 = int64_t(blockIdx.y) * kchunk;
        A += TA ? k0 * lda : k0;
        B += TB ? k0 : k0 * ldb;
        K = K - k0 < kchunk ? K - k0 : kchunk;
        C += int64_t(blockIdx.y) * cstride;
    }
    constexpr bool AKC = !TA;   // A panel rows are M: K-contiguous unless A is stored transposed
    constexpr bool BKC = TB;    // B panel rows are N: K-contiguous only when B is stored transposed

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lg = lane >> 4;

    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, upper, tm, tn);
    const int64_t m0 = int64_t(tm) * BM;
    const int64_t n0 = int64_t(tn) * BN;

    PanelMover<AKC, VEC> la;
    PanelMover<BKC, VEC> lb;
    la.init(A, lda, m0, M, K, tid);
    lb.init(B, ldb, n0, N, K, tid);

    mrm_d4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = mrm_d4{0.0, 0.0, 0.0, 0.0};

    const int64_t nk = (K + BK - 1) / BK;
    auto stage = [&](int i) { return smem + i * 2 * PANEL; };
    // stage 0 <- K step 0, stage 1 <- K step 1
    if (la.direct_ok()) {
        la.issue_direct(stage(0));
    } else {
        la.load_regs();
        la.store(stage(0));
    }
    if (lb.direct_ok()) {
        lb.issue_direct(stage(0) + PANEL);
    } else {
        lb.load_regs();
        lb.store(stage(0) + PANEL);
    }
    if (nk > 1) {
        la.advance();
        lb.advance();
        if (la.direct_ok()) {
            la.issue_direct(stage(1));
        } else {
            la.load_regs();
            la.store(stage(1));
        }
        if (lb.direct_ok()) {
            lb.issue_direct(stage(1) + PANEL);
        } else {
            lb.load_regs();
            lb.store(stage(1) + PANEL);
        }
    }
    wait_vmem();
    __syncthreads();

    // fragment rows of this lane: A row wm*64 + 16p + li, B column wn*64 + 16c + li
    const int arow = wm * 64 + li;
    const int bcol = wn * 64 + li;
    double2 fa[2][4], fb[2][4];
    auto read_half = [&](const double* S, int m) {
#pragma unroll
        for (int p = 0; p < 4; ++p) fa[m][p] = frag_pair<AKC>(S, arow + 16 * p, m, lg);
#pragma unroll
        for (int c = 0; c < 4; ++c) fb[m][c] = frag_pair<BKC>(S + PANEL, bcol + 16 * c, m, lg);
    };
    auto mma_half = [&](int m) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[p][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[m][p].x, fb[m][c].x, acc[p][c], 0, 0, 0);
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[p][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[m][p].y, fb[m][c].y, acc[p][c], 0, 0, 0);
    };
    // fragment reads per half step: a 16-byte chunk per KC pair, two 8-byte
    // reads per RC pair
    constexpr int RD = (AKC ? 4 : 8) + (BKC ? 4 : 8);
    read_half(stage(0), 0);

    // K step t: the first half's 32 MFMAs run while the second half's
    // fragments are read; the barrier in the middle of the step publishes step
    // t+1 (its loads were issued one step earlier) and frees stage t%2, which
    // then receives step t+2 while the second half's MFMAs run and step t+1's
    // first-half fragments are read, so the matrix pipe has work on both sides
    // of the barrier.  `fast`: a tile whose panel moves are all direct up to
    // step t+2 (no branch, so the loads interleave with the MFMAs); otherwise
    // the predicated path stores its registers after the MFMAs.
    auto step = [&](int64_t t, auto fast) {
        constexpr bool F = decltype(fast)::value;
        const int cur = int(t & 1);
        read_half(stage(cur), 1);
        mma_half(0);
        if constexpr (F) gm_interleave<0, RD, 32>();
        wait_vmem();
        __syncthreads();
        bool a_regs = false, b_regs = false;
        if constexpr (F) {
            la.advance();
            lb.advance();
            la.issue_direct(stage(cur));
            lb.issue_direct(stage(cur) + PANEL);
        } else if (t + 2 < nk) {
            la.advance();
            lb.advance();
            if (la.direct_ok()) la.issue_direct(stage(cur));
            else a_regs = true;
            if (lb.direct_ok()) lb.issue_direct(stage(cur) + PANEL);
            else b_regs = true;
        }
        if (F || t + 1 < nk) read_half(stage(cur ^ 1), 0);
        mma_half(1);
        if constexpr (F) gm_interleave<8, RD, 32>();
        if (a_regs) {
            la.load_regs();
            la.store(stage(cur));
        }
        if (b_regs) {
            lb.load_regs();
            lb.store(stage(cur) + PANEL);
        }
    };
    // steps whose step t+2 moves directly: every step before the last partial
    // one of a tile with all rows present and aligned operands
    const int64_t nfull = K / BK;
    const bool tile_direct = VEC && la.rows_full && lb.rows_full;
    const int64_t nfast = tile_direct && nfull > 2 ? nfull - 2 : 0;
    int64_t t = 0;
    for (; t < nfast; ++t) step(t, std::true_type());
    for (; t < nk; ++t) step(t, std::false_type());

    // Epilogue: acc[p][c] element v of lane l is row 16p + l/16 + 4v, column
    // 16c + l%16 of the wave tile (16 lanes write 128 contiguous bytes of a
    // row).  For doAdd / doSub the 16 old values of a block row p (and of
    // their mirrors) are read together before any of them is stored: one
    // memory round trip per block row, four per tile (the compiler cannot
    // move a read of C past a store to C; one round trip per row measured 35%
    // of the LU's K = 256 trailing updates).
    const bool mirror = upper && tm != tn;  // the mirrored element of a symmetric product (M == N)
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int64_t r0 = m0 + wm * 64 + 16 * p + lg;
        double old[4][4], mold[4][4];
        if (op != MRM_DO_NONE) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int64_t row = r0 + 4 * v, col = n0 + wn * 64 + 16 * c + li;
                    const bool in = row < M && col < N;
                    old[v][c] = in ? C[row * ldc + col] : 0.0;
                    mold[v][c] = (in && mirror) ? C[col * ldc + row] : 0.0;
                }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = r0 + 4 * v;
            if (row >= M) continue;
            double* crow = C + row * ldc;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int64_t col = n0 + wn * 64 + 16 * c + li;
                if (col < N) {
                    const double x = acc[p][c][v];
                    crow[col] = op == MRM_DO_NONE ? x : (op == MRM_DO_ADD ? old[v][c] + x : old[v][c] - x);
                    if (mirror)
                        C[col * ldc + row] =
                            op == MRM_DO_NONE ? x : (op == MRM_DO_ADD ? mold[v][c] + x : mold[v][c] - x);
                }
            }
        }
    }
}

template <bool TA, bool TB>
void launch_variant(bool vec, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                    int64_t ldb, double* C, int64_t ldc, int op, int tiles_m, int tiles_n, int upper, hipStream_t s,
                    int splits = 1, int64_t kchunk = 0, int64_t cstride = 0, int batch = 1, int64_t bsa = 0,
                    int64_t bsb = 0, int64_t bsc = 0) {
    const unsigned ntiles = upper ? unsigned(tiles_m) * unsigned(tiles_m + 1) / 2 : unsigned(tiles_m) * unsigned(tiles_n);
    dim3 grid(ntiles, unsigned(splits), unsigned(batch)), block(NTHREADS);
    if (vec)
        hipLaunchKernelGGL((dgemm_mfma_kernel<TA, TB, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                           op, tiles_m, tiles_n, kchunk, cstride, upper, bsa, bsb, bsc);
    else
        hipLaunchKernelGGL((dgemm_mfma_kernel<TA, TB, false>), grid, block, 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                           op, tiles_m, tiles_n, kchunk, cstride, upper, bsa, bsb, bsc);
}

// C (op)= sum_c W[c] over `splits` partial products, in chunk order.
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const double* __restrict__ W, int64_t cstride, int splits,
                                                            double* __restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                            int op) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t i = idx / N, j = idx % N;
    double acc = W[idx];
    for (int c = 1; c < splits; ++c) acc += W[int64_t(c) * cstride + idx];
    double* cp = C + i * ldc + j;
    if (op == MRM_DO_NONE) *cp = acc;
    else if (op == MRM_DO_ADD) *cp = *cp + acc;
    else *cp = *cp - acc;
}

// Split-K when the output has too few tiles to fill the GPU and K is long
// (the covariance shape X^T X with X tall): up to 64 chunks of at least 512.
int splitk_count(int64_t tiles, int64_t K) {
    if (tiles >= 128 || K < 4096) return 1;
    // at most one wave of workgroups (two per CU): a partial second wave
    // would double the time
    int64_t p = 512 / tiles;
    p = p < K / 512 ? p : K / 512;
    return int(p < 64 ? p : 64);
}

// C (M x N) = 0: the empty product for doNone.
__global__ void zero_matrix_kernel(double* C, int64_t ldc, int64_t M, int64_t N) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < M * N) C[(idx / N) * ldc + idx % N] = 0.0;
}

}  // namespace

namespace {

// The split of a product: K chunks (1 = no split) and their chunk length.
struct Split {
    int used = 1;
    int64_t kchunk = 0;
    int upper = 0;
};

Split plan_split(int trans, int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B, int64_t ldb,
                 int want_splits) {
    Split sp;
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    // A^T A and A A^T are symmetric: compute the tiles on and above the
    // diagonal and mirror them (the mirrored sums are the same products in the
    // same order, so the result is the full product's bit for bit)
    sp.upper = (trans != MRM_NN && A == B && lda == ldb && M == N) ? 1 : 0;
    const int64_t ntiles = sp.upper ? tm * (tm + 1) / 2 : tm * tn;
    int splits = splitk_count(ntiles, K);
    if (want_splits > 0) {
        const int64_t most = (K + BK - 1) / BK;
        splits = int(want_splits < most ? want_splits : most);
    }
    if (splits > 1) {
        sp.kchunk = ((K + splits - 1) / splits + BK - 1) / BK * BK;
        sp.used = int((K + sp.kchunk - 1) / sp.kchunk);
    }
    return sp;
}

// C (M x N) = / += / -= op(A) * op(B), K the inner dimension.  Strides in bytes.
// want_splits > 0 asks for that many K chunks (at least BK each), 0 lets the
// shape decide.  The partial products of a split (used * M * N doubles) go to
// work; with checked, work_bytes is what work holds and a split that needs
// more is an error (nothing is allocated); otherwise work, when given, holds
// enough, and without it they come from a stream-ordered allocation.
int launch_dgemm(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                 int64_t ldb, double* C, int64_t ldc, int op, int want_splits, double* work, size_t work_bytes,
                 bool checked, hipStream_t stream) {
    if (M < 0 || N < 0 || K < 0 || op < 0 || op > 2 || trans < 0 || trans > 2) {
        mrm_set_error(MRM_BAD_ARG, "dgemm: bad argument (M=%lld N=%lld K=%lld op=%d trans=%d)", (long long)M,
                      (long long)N, (long long)K, op, trans);
        return MRM_BAD_ARG;
    }
    if (M == 0 || N == 0) return MRM_OK;
    if ((lda | ldb | ldc) & 7) {
        mrm_set_error(MRM_BAD_ARG, "dgemm: line widths must be multiples of 8 bytes");
        return MRM_BAD_ARG;
    }
    const int64_t ea = lda / 8, eb = ldb / 8, ec = ldc / 8;
    if (K == 0) {
        // Empty product: dest = 0 for doNone, unchanged for doAdd / doSub.
        if (op == MRM_DO_NONE) {
            const int64_t n = M * N;
            hipLaunchKernelGGL(zero_matrix_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, C, ec, M,
                               N);
            MRM_HIP_TRY(hipGetLastError());
        }
        return MRM_OK;
    }
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    if (tm * tn > 0x7fffffffLL) {
        mrm_set_error(MRM_BAD_DIMS, "dgemm: problem too large for one launch");
        return MRM_BAD_DIMS;
    }
    const bool vec = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0 &&
                     ((lda | ldb) & 15) == 0;
    const Split sp = plan_split(trans, M, N, K, A, lda, B, ldb, want_splits);
    const int upper = sp.upper;
    if (sp.used > 1) {
        const int64_t kchunk = sp.kchunk;
        const int used = sp.used;
        const int64_t cstride = M * N;
        const size_t need = size_t(used) * size_t(cstride) * 8;
        if (checked && (!work || need > work_bytes)) {
            mrm_set_error(MRM_BAD_DIMS, "dgemm: the split product needs %zu bytes of workspace, %zu given", need,
                          work ? work_bytes : size_t(0));
            return MRM_BAD_DIMS;
        }
        void* W = work;
        if (!W) MRM_HIP_TRY(hipMallocAsync(&W, need, stream));
        double* w = static_cast<double*>(W);
        switch (trans) {
            case MRM_NN: launch_variant<false, false>(vec, M, N, K, A, ea, B, eb, w, N, MRM_DO_NONE, int(tm), int(tn), upper, stream, used, kchunk, cstride); break;
            case MRM_TN: launch_variant<true, false>(vec, M, N, K, A, ea, B, eb, w, N, MRM_DO_NONE, int(tm), int(tn), upper, stream, used, kchunk, cstride); break;
            default:     launch_variant<false, true>(vec, M, N, K, A, ea, B, eb, w, N, MRM_DO_NONE, int(tm), int(tn), upper, stream, used, kchunk, cstride); break;
        }
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3(unsigned((M * N + 255) / 256)), dim3(256), 0, stream, w, cstride,
                           used, C, ec, M, N, op);
        MRM_HIP_TRY(hipGetLastError());
        if (!work) MRM_HIP_TRY(hipFreeAsync(W, stream));
        return MRM_OK;
    }
    switch (trans) {
        case MRM_NN: launch_variant<false, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), upper, stream); break;
        case MRM_TN: launch_variant<true, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), upper, stream); break;
        default:     launch_variant<false, true>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), upper, stream); break;
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

}  // namespace

int mrm_launch_dgemm_batched(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, int64_t sa,
                             const double* B, int64_t ldb, int64_t sb, double* C, int64_t ldc, int64_t sc, int batch,
                             int op, hipStream_t stream) {
    if (M < 0 || N < 0 || K < 0 || op < 0 || op > 2 || trans < 0 || trans > 2 || batch < 0 || batch > 65535 ||
        ((lda | ldb | ldc) & 7)) {
        mrm_set_error(MRM_BAD_ARG, "dgemm batched: bad argument");
        return MRM_BAD_ARG;
    }
    if (M == 0 || N == 0 || batch == 0) return MRM_OK;
    if (K == 0) {
        for (int b = 0; b < batch; ++b) {
            const int rc = launch_dgemm(trans, M, N, 0, A, lda, B, ldb, C + int64_t(b) * sc, ldc, op, 1, nullptr, 0,
                                        false, stream);
            if (rc) return rc;
        }
        return MRM_OK;
    }
    const int64_t ea = lda / 8, eb = ldb / 8, ec = ldc / 8;
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    const bool vec = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0 &&
                     ((lda | ldb) & 15) == 0 && ((sa | sb) & 1) == 0;
    switch (trans) {
        case MRM_NN: launch_variant<false, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), 0, stream, 1, 0, 0, batch, sa, sb, sc); break;
        case MRM_TN: launch_variant<true, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), 0, stream, 1, 0, 0, batch, sa, sb, sc); break;
        default:     launch_variant<false, true>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), 0, stream, 1, 0, 0, batch, sa, sb, sc); break;
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}

int mrm_launch_dgemm(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double* C, int64_t ldc, int op, hipStream_t stream) {
    return launch_dgemm(trans, M, N, K, A, lda, B, ldb, C, ldc, op, 0, nullptr, 0, false, stream);
}

int mrm_launch_dgemm_splitk(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, int op, int splits,
                            double* work, hipStream_t stream) {
    return launch_dgemm(trans, M, N, K, A, lda, B, ldb, C, ldc, op, splits < 1 ? 1 : splits, work, 0, false,
                        stream);
}

int mrm_launch_dgemm_ws(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double* C, int64_t ldc, int op, void* work, size_t work_bytes,
                        hipStream_t stream) {
    return launch_dgemm(trans, M, N, K, A, lda, B, ldb, C, ldc, op, 0, static_cast<double*>(work), work_bytes, true,
                        stream);
}

size_t mrm_dgemm_work_bytes(int trans, int64_t M, int64_t N, int64_t K, bool same_operand) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    // the operand pointers only matter for the symmetric test
    const double one = 0.0, two = 0.0;
    const void* A = &one;
    const void* B = same_operand ? static_cast<const void*>(&one) : static_cast<const void*>(&two);
    const Split sp = plan_split(trans, M, N, K, A, 8, B, 8, 0);
    return sp.used > 1 ? size_t(sp.used) * size_t(M) * size_t(N) * 8 : 0;
}
