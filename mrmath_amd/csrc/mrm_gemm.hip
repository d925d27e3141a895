// mrm_gemm.hip -- FP64 GEMM on the gfx950 matrix cores (v_mfma_f64_4x4x4_4b_f64).
//
// Replaces the reference's blocked FMA product (BlockMatrixMultiplication[T1/T2],
// src/BlockedMult.pas:350-724, and its micro-kernel FMAMatrixMultTransposed,
// src/FMAMatrixOperations.pas:55-92).  One kernel family covers NN, TN (MultT1)
// and NT (MultT2): the operand layout only changes how a 128x16 panel is read
// from HBM; inside LDS both panels are always stored K-contiguous, so the
// tile transposes the reference performs per K-block (src/BlockedMult.pas:435,
// 560-561) disappear into the load pattern.  The destination operation
// (doNone / doAdd / doSub, src/BlockedMult.pas:459-463) is fused into the
// epilogue.
//
// Workgroup: 256 threads = 4 waves in a 2x2 arrangement; workgroup tile
// BM x BN = 128 x 128, K step BK = 16; each wave owns a 64x64 block of C held
// as 4x4 MFMA accumulators (16 x 4 doubles per lane).  LDS holds two stages of
// both panels (72 KiB); the next K-panel is fetched into registers while the
// current one is consumed, so there is one barrier per K step.
//
// K ordering inside a step: lane l (i = l%16, g = l/16) owns the 4 doubles
// k = 4g..4g+3 of the K step, and MFMA round kk (0..3) consumes element kk of
// both operands.  A and B use the same mapping, so every product A[m,k]*B[k,n]
// is formed exactly once per step.
//
// Matrix instruction (template parameter MF).  On gfx950 the four-block
// v_mfma_f64_4x4x4_4b_f64 runs at the full FP64 rate while
// v_mfma_f64_16x16x4_f64 does not (pure-MFMA loops: ~70 vs ~48 TFLOP/s), so MF4
// is the default.  Measured lane layouts (l = lane):
//   16x16x4: A[l%16][l/16], B[l/16][l%16]; accumulator element v of lane l is
//            C[l/16 + 4v][l%16].
//   4x4x4_4b (four independent 4x4x4 blocks b): lane l = 16k + 4b + j supplies
//            A_b[j][k] and B_b[k][j]; D_b[i][n] lands in lane 16i + 4b + n.
//            The CBSZ/ABID/BLGP fields of FP64 MFMAs are negation modifiers on
//            gfx950, so there is no operand broadcast: each block needs its own
//            copy of its B quad.
// The MF4 wave tile (64x64) is built from 16x16 sub-tiles p (rows) x q
// (columns): the A fragment of sub-row p puts row 16p + l%16 (= 16p + 4b + j)
// in lane l, and the B fragment for column quad s of sub-column q puts column
// 16q + 4s + l%4 in every block; block b then yields C[16p + 4b + i][16q + 4s + n]
// in lane 16i + 4b + n.  Four quads s cover the 16x16 sub-tile.

#include <cstdlib>

#include "mrm_common.h"

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int NTHREADS = 256;
constexpr int SK = BK + 2;               // LDS row stride in doubles (144 B): spreads rows over banks
constexpr int PANEL = BM * SK;           // doubles per panel stage (BM == BN)

// Panel of 128 rows ("r", the M or N index) x 16 k.  KC: element (r,k) lives at
// base[r*ld + k] (K-contiguous rows); RC: at base[k*ld + r].  Each thread moves
// four 2-double vectors per K step.  The per-thread pointer is computed once;
// the four vectors sit at uniform offsets from it and every K step advances it
// by a uniform amount, so the loop carries one 64-bit address per panel.
//   KC: thread -> row rr = tid/8 (+32j), k pair kc = 2*(tid%8)
//   RC: thread -> k  kr = tid/64 (+4j),  row pair rc = 2*(tid%64)
template <bool KC, bool VEC>
struct PanelLoader {
    const double* p;     // this thread's element for j = 0 at the current K step
    int64_t jstep;       // element offset between vector j and j+1
    int64_t kstep;       // element offset of one K step
    int64_t kpos;        // K index of this thread's first element (KC: k pair; RC: k row of j=0)
    int64_t K;
    unsigned rmask;      // KC: bit j = row of vector j in range; RC: bit0/bit1 = row pair in range
    double2 v[4];

    __device__ __forceinline__ void init(const double* __restrict__ base, int64_t ld, int64_t r0, int64_t R,
                                         int64_t K_, int tid) {
        K = K_;
        if (KC) {
            const int64_t rr = r0 + (tid >> 3);
            kpos = (tid & 7) * 2;
            p = base + rr * ld + kpos;
            jstep = 32 * ld;
            kstep = BK;
            rmask = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) rmask |= (rr + 32 * j < R) ? (1u << j) : 0u;
        } else {
            const int64_t rc = r0 + (tid & 63) * 2;
            kpos = tid >> 6;
            p = base + kpos * ld + rc;
            jstep = 4 * ld;
            kstep = int64_t(BK) * ld;
            rmask = (rc < R ? 1u : 0u) | (rc + 1 < R ? 2u : 0u);
        }
    }

    __device__ __forceinline__ void load() {
        const bool kfull = kpos + BK <= K;   // every K index of this step is in range
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double* q = p + j * jstep;
            bool ok0, ok1;
            if (KC) {
                const bool rok = (rmask >> j) & 1u;
                ok0 = rok && (kfull || kpos < K);
                ok1 = rok && (kfull || kpos + 1 < K);
            } else {
                const bool kok = kfull || (kpos + 4 * j < K);
                ok0 = kok && (rmask & 1u);
                ok1 = kok && (rmask & 2u);
            }
            if (VEC && ok0 && ok1) {
                v[j] = *reinterpret_cast<const double2*>(q);
            } else {
                double2 t;
                t.x = ok0 ? q[0] : 0.0;
                t.y = ok1 ? q[1] : 0.0;
                v[j] = t;
            }
        }
    }

    __device__ __forceinline__ void advance() {
        p += kstep;
        kpos += BK;
    }

    __device__ __forceinline__ void store(double* __restrict__ P, int tid) const {
        if (KC) {
            const int r = tid >> 3, k = (tid & 7) * 2;
#pragma unroll
            for (int j = 0; j < 4; ++j) *reinterpret_cast<double2*>(P + (r + 32 * j) * SK + k) = v[j];
        } else {
            const int k = tid >> 6, r = (tid & 63) * 2;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                P[r * SK + k + 4 * j] = v[j].x;
                P[(r + 1) * SK + k + 4 * j] = v[j].y;
            }
        }
    }
};

// Workgroup id -> (tile_m, tile_n).  Workgroups are dealt round-robin to the 8
// XCDs; remap so that each XCD receives a contiguous run of logical tiles, then
// walk that run in groups of GROUP_M tile-rows so the workgroups resident on
// one XCD share A and B panels in that XCD's L2.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int NXCD = 8;
    constexpr int GROUP_M = 8;
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

enum { MF16 = 0, MF4 = 1 };

// TA: A stored K x M (MultT1).  TB: B stored N x K (MultT2).  Strides in elements.
template <bool TA, bool TB, bool VEC, int MF>
__global__ void __launch_bounds__(NTHREADS, 2)
dgemm_mfma_kernel(int64_t M, int64_t N, int64_t K, const double* __restrict__ A, int64_t lda,
                  const double* __restrict__ B, int64_t ldb, double* __restrict__ C, int64_t ldc,
                  int op, int tiles_m, int tiles_n) {
    __shared__ __attribute__((aligned(16))) double smem[2 * 2 * PANEL];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lg = lane >> 4;

    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t m0 = int64_t(tm) * BM;
    const int64_t n0 = int64_t(tn) * BN;

    // A panel rows are M: K-contiguous unless A is stored transposed.
    PanelLoader<!TA, VEC> la;
    // B panel rows are N: K-contiguous only when B is stored transposed.
    PanelLoader<TB, VEC> lb;
    la.init(A, lda, m0, M, K, tid);
    lb.init(B, ldb, n0, N, K, tid);

    // MF16: acc[p][q] is one 16x16 sub-tile (4 doubles per lane).
    // MF4:  acc[p][q] element s is column quad s of sub-tile (p, q).
    mrm_d4 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = mrm_d4{0.0, 0.0, 0.0, 0.0};

    const int64_t nk = (K + BK - 1) / BK;
    la.load();
    lb.load();
    la.store(smem, tid);
    lb.store(smem + PANEL, tid);
    __syncthreads();

    const int a_off = (wm * 64 + li) * SK + 4 * lg;
    const int b_off = (wn * 64 + (MF == MF4 ? (lane & 3) : li)) * SK + 4 * lg;

    for (int64_t t = 0; t < nk; ++t) {
        const int cur = int(t & 1);
        const bool more = (t + 1) < nk;
        if (more) {
            la.advance();
            lb.advance();
            la.load();
            lb.load();
        }
        const double* As = smem + cur * 2 * PANEL;
        const double* Bs = As + PANEL;
        if (MF == MF16) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                double2 a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = *reinterpret_cast<const double2*>(As + a_off + r * 16 * SK + 2 * h);
#pragma unroll
                for (int c = 0; c < 4; ++c) b[c] = *reinterpret_cast<const double2*>(Bs + b_off + c * 16 * SK + 2 * h);
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r].x, b[c].x, acc[r][c], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r].y, b[c].y, acc[r][c], 0, 0, 0);
            }
        } else {
#pragma unroll 1
            for (int kk = 0; kk < 4; ++kk) {
                double a[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = As[a_off + r * 16 * SK + kk];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    double b[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) b[q] = Bs[b_off + (c * 16 + 4 * q) * SK + kk];
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[r][c][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[r], b[q], acc[r][c][q], 0, 0, 0);
                }
            }
        }
        if (more) {
            double* Ns = smem + (cur ^ 1) * 2 * PANEL;
            la.store(Ns, tid);
            lb.store(Ns + PANEL, tid);
        }
        __syncthreads();
    }

    // Epilogue.  MF16: element v of lane l is row l/16 + 4v, column l%16 of the
    // sub-tile.  MF4: element s of lane l = 16i + 4b + n is row 4b + i, column 4s + n.
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int64_t row, col;
                if (MF == MF16) {
                    row = m0 + wm * 64 + r * 16 + lg + 4 * v;
                    col = n0 + wn * 64 + c * 16 + li;
                } else {
                    row = m0 + wm * 64 + r * 16 + 4 * ((lane >> 2) & 3) + lg;
                    col = n0 + wn * 64 + c * 16 + 4 * v + (lane & 3);
                }
                if (row < M && col < N) {
                    double* p = C + row * ldc + col;
                    const double x = acc[r][c][v];
                    if (op == MRM_DO_NONE) *p = x;
                    else if (op == MRM_DO_ADD) *p = *p + x;
                    else *p = *p - x;
                }
            }
        }
    }
}

template <bool TA, bool TB, int MF>
void launch_mf(bool vec, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
               int64_t ldb, double* C, int64_t ldc, int op, int tiles_m, int tiles_n, hipStream_t s) {
    dim3 grid(unsigned(tiles_m) * unsigned(tiles_n)), block(NTHREADS);
    if (vec)
        hipLaunchKernelGGL((dgemm_mfma_kernel<TA, TB, true, MF>), grid, block, 0, s, M, N, K, A, lda, B, ldb, C,
                           ldc, op, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((dgemm_mfma_kernel<TA, TB, false, MF>), grid, block, 0, s, M, N, K, A, lda, B, ldb, C,
                           ldc, op, tiles_m, tiles_n);
}

// MRM_DGEMM_MFMA=16 selects the 16x16x4 instruction (for comparison runs).
int mfma_scheme() {
    static int scheme = [] {
        const char* e = getenv("MRM_DGEMM_MFMA");
        return (e && atoi(e) == 16) ? int(MF16) : int(MF4);
    }();
    return scheme;
}

template <bool TA, bool TB>
void launch_variant(bool vec, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                    int64_t ldb, double* C, int64_t ldc, int op, int tiles_m, int tiles_n, hipStream_t s) {
    if (mfma_scheme() == MF16)
        launch_mf<TA, TB, MF16>(vec, M, N, K, A, lda, B, ldb, C, ldc, op, tiles_m, tiles_n, s);
    else
        launch_mf<TA, TB, MF4>(vec, M, N, K, A, lda, B, ldb, C, ldc, op, tiles_m, tiles_n, s);
}

// C (M x N) = / += / -= op(A) * op(B), K the inner dimension.  Strides in bytes.
__global__ void scale_zero_kernel(double* C, int64_t ldc, int64_t M, int64_t N) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < M * N) C[(idx / N) * ldc + idx % N] = 0.0;
}

}  // namespace

int mrm_launch_dgemm(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double* C, int64_t ldc, int op, hipStream_t stream) {
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
            hipLaunchKernelGGL(scale_zero_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, C, ec, M,
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
    switch (trans) {
        case MRM_NN: launch_variant<false, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), stream); break;
        case MRM_TN: launch_variant<true, false>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), stream); break;
        default:     launch_variant<false, true>(vec, M, N, K, A, ea, B, eb, C, ec, op, int(tm), int(tn), stream); break;
    }
    MRM_HIP_TRY(hipGetLastError());
    return MRM_OK;
}
