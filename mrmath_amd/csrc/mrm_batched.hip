// mrm_batched.hip -- thousands of small, uneven FP64 GEMMs and small linear
// solves in persistent launches replayed from a hipGraph.
//
// The reference has no batched entry: its callers loop over MatrixMult /
// MatrixLinEQSolve per problem (src/MatrixASMStubSwitch.pas:769-776,
// src/LinAlgLU.pas:481-592), and the small-size path of its multiply is
// FMAMatrixMult (src/FMAMatrixOperations.pas:94-143).  Here a plan gathers the
// problem list once:
//   GEMM: every problem is cut into T x T output tiles, all tiles of all
//     problems form one list, and a persistent grid of workgroups takes tiles
//     from an atomic counter (largest K first).  T in {32, 64, 128} is picked
//     by timing the actual batch (the autotuner) unless the caller fixes it.
//     FP64 on the FP64 matrix cores, or FP32 pieces with error compensation
//     on the FP32 matrix cores and FP64 accumulation per K chunk.
//   LU solve: systems are grouped by size class (n <= 32, 64, 96, 128); per class
//     one persistent launch where each wave (up to 64 unknowns) or each
//     workgroup of eight waves (above) takes whole systems from a counter,
//     holds A in registers, factors it with partial pivoting (the reference's
//     rule: first row of maximal |a|, singular at |pivot| <= 10/MaxDouble) and
//     substitutes (mrm_batched_lu_wave.hip, mrm_batched_lu_block.h).  FP64,
//     or FP32 factors with FP64 residuals and refinement (x += (LU)^-1 (b - A x),
//     A read again in FP64 from memory).
// The counter reset and the launches are captured once into a hipGraph; a run
// is one hipGraphLaunch.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "mrm_batched_lu.h"
#include "mrm_common.h"
#include "mrm_host.h"

namespace {

constexpr int BG_THREADS = 256;
constexpr int BG_BK = 16;

struct GemmP {  // element strides
    int64_t M, N, K, lda, ldb, ldc;
    const double* A;
    const double* B;
    double* C;
};

struct Tile {
    int p, m0, n0, pad;
};

typedef float mrm_f16 __attribute__((ext_vector_type(16)));

__global__ void zero_counters_kernel(unsigned* c, int n) {
    if (int(threadIdx.x) < n) c[threadIdx.x] = 0;
}

// One T x T tile of C = A * B per iteration of the persistent loop.  Four
// waves in 2 x 2, each (T/2) x (T/2) as (T/32)^2 blocks of 16 x 16 on
// v_mfma_f64_16x16x4f64 (lane l supplies A[l%16][l/16] and B[l/16][l%16];
// result register v of lane l is row l/16 + 4v, column l%16).  K in chunks of
// 16 staged in LDS as [k][row], the next chunk prefetched into registers.
template <int T>
__global__ void __launch_bounds__(BG_THREADS) bg_gemm_kernel(const GemmP* __restrict__ probs,
                                                             const Tile* __restrict__ tiles, int ntiles,
                                                             unsigned* __restrict__ counter) {
    constexpr int BP = T / 32;               // 16x16 blocks per wave per dimension
    constexpr int EA = T * BG_BK / BG_THREADS;  // elements per thread per chunk
    constexpr int LDT = T + 2;
    __shared__ double As[2][BG_BK][LDT];  // two stages: one barrier per K chunk
    __shared__ double Bs[2][BG_BK][LDT];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    for (;;) {
        if (tid == 0) s_item = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= ntiles) return;
        const Tile tl = tiles[item];
        const GemmP pr = probs[tl.p];
        const int64_t m0 = tl.m0, n0 = tl.n0;
        double ra[EA], rb[EA];
        // element tid + 256 i of a chunk: A row ar + (256/BK) i, column ak
        // (k fastest, row-major A); B k-row bk + (256/T) i, column bc
        static_assert(BG_THREADS % BG_BK == 0 && BG_THREADS % T == 0, "chunk layout");
        const int ak = tid % BG_BK, ar = tid / BG_BK;
        const int bc = tid % T, bk = tid / T;
        const double* pa = pr.A + (m0 + ar) * pr.lda + ak;
        const double* pb = pr.B + int64_t(bk) * pr.ldb + n0 + bc;
        const bool bcol_ok = n0 + bc < pr.N;
        auto fetch = [&](int64_t k0) {
            const bool ak_ok = k0 + ak < pr.K;
#pragma unroll
            for (int i = 0; i < EA; ++i) {
                const int rr = ar + (BG_THREADS / BG_BK) * i;
                ra[i] = (ak_ok && m0 + rr < pr.M) ? pa[int64_t((BG_THREADS / BG_BK) * i) * pr.lda + k0] : 0.0;
                const int kr = bk + (BG_THREADS / T) * i;
                rb[i] = (bcol_ok && k0 + kr < pr.K) ? pb[(k0 + (BG_THREADS / T) * i) * pr.ldb] : 0.0;
            }
        };
        auto stash = [&](int st) {
#pragma unroll
            for (int i = 0; i < EA; ++i) {
                const int idx = tid + BG_THREADS * i;
                As[st][idx % BG_BK][idx / BG_BK] = ra[i];
                Bs[st][idx / T][idx % T] = rb[i];
            }
        };
        mrm_d4 acc[BP][BP];
#pragma unroll
        for (int p = 0; p < BP; ++p)
#pragma unroll
            for (int c = 0; c < BP; ++c) acc[p][c] = mrm_d4{0.0, 0.0, 0.0, 0.0};
        fetch(0);
        stash(0);
        __syncthreads();
        int st = 0;
        for (int64_t k0 = 0; k0 < pr.K; k0 += BG_BK) {
            const bool more = k0 + BG_BK < pr.K;
            if (more) fetch(k0 + BG_BK);
#pragma unroll
            for (int kk = 0; kk < BG_BK / 4; ++kk) {
                const int k = 4 * kk + (lane >> 4);
                double a[BP], b[BP];
#pragma unroll
                for (int p = 0; p < BP; ++p) a[p] = As[st][k][wm * (T / 2) + 16 * p + (lane & 15)];
#pragma unroll
                for (int c = 0; c < BP; ++c) b[c] = Bs[st][k][wn * (T / 2) + 16 * c + (lane & 15)];
#pragma unroll
                for (int p = 0; p < BP; ++p)
#pragma unroll
                    for (int c = 0; c < BP; ++c)
                        acc[p][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p], b[c], acc[p][c], 0, 0, 0);
            }
            // the other stage was last read in the previous chunk, before the
            // barrier that ended it
            if (more) stash(st ^ 1);
            __syncthreads();
            st ^= 1;
        }
#pragma unroll
        for (int p = 0; p < BP; ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = m0 + wm * (T / 2) + 16 * p + (lane >> 4) + 4 * v;
                if (row >= pr.M) continue;
#pragma unroll
                for (int c = 0; c < BP; ++c) {
                    const int64_t col = n0 + wn * (T / 2) + 16 * c + (lane & 15);
                    if (col < pr.N) pr.C[row * pr.ldc + col] = acc[p][c][v];
                }
            }
    }
}

// The same tile loop on the FP32 matrix cores with first-order error
// compensation (the split of Ootomo and Yokota, "Recovering single precision
// accuracy from Tensor Cores while surpassing the FP32 theoretical peak
// performance", IJHPCA 2022, here with FP32 pieces of FP64 operands):
// A = Ah + Al, B = Bh + Bl with Ah = fp32(A), Al = fp32(A - Ah), and
// C = Ah Bh + (Ah Bl + Al Bh) with the neglected Al Bl below 2^-48 |A||B|.
// Each K chunk of 16 accumulates the main and the correction products in
// FP32 on v_mfma_f32_32x32x2_f32 (lane l supplies A[l%32][l/32] and
// B[l/32][l%32]; result register v of lane l is row 8(v/4) + 4(l/32) + v%4,
// column l%32), and the chunk sums are added into FP64 accumulators, so the
// error is that of FP32 sums over 16 products, not over K.  T in {64, 128}:
// four waves of (T/2) x (T/2) as (T/64)^2 blocks of 32 x 32.
template <int T>
__global__ void __launch_bounds__(BG_THREADS) bg_gemm_f32x3_kernel(const GemmP* __restrict__ probs,
                                                                   const Tile* __restrict__ tiles, int ntiles,
                                                                   unsigned* __restrict__ counter) {
    static_assert(T == 64 || T == 128, "32 x 32 blocks per wave");
    constexpr int BP = T / 64;
    constexpr int EA = T * BG_BK / BG_THREADS;
    constexpr int LDT = T + 4;
    __shared__ float Ah[2][BG_BK][LDT], Al[2][BG_BK][LDT];
    __shared__ float Bh[2][BG_BK][LDT], Bl[2][BG_BK][LDT];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    for (;;) {
        if (tid == 0) s_item = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= ntiles) return;
        const Tile tl = tiles[item];
        const GemmP pr = probs[tl.p];
        const int64_t m0 = tl.m0, n0 = tl.n0;
        double ra[EA], rb[EA];
        const int ak = tid % BG_BK, ar = tid / BG_BK;
        const int bc = tid % T, bk = tid / T;
        const double* pa = pr.A + (m0 + ar) * pr.lda + ak;
        const double* pb = pr.B + int64_t(bk) * pr.ldb + n0 + bc;
        const bool bcol_ok = n0 + bc < pr.N;
        auto fetch = [&](int64_t k0) {
            const bool ak_ok = k0 + ak < pr.K;
#pragma unroll
            for (int i = 0; i < EA; ++i) {
                const int rr = ar + (BG_THREADS / BG_BK) * i;
                ra[i] = (ak_ok && m0 + rr < pr.M) ? pa[int64_t((BG_THREADS / BG_BK) * i) * pr.lda + k0] : 0.0;
                const int kr = bk + (BG_THREADS / T) * i;
                rb[i] = (bcol_ok && k0 + kr < pr.K) ? pb[(k0 + (BG_THREADS / T) * i) * pr.ldb] : 0.0;
            }
        };
        auto stash = [&](int st) {
#pragma unroll
            for (int i = 0; i < EA; ++i) {
                const int idx = tid + BG_THREADS * i;
                const float ah = float(ra[i]), bh = float(rb[i]);
                Ah[st][idx % BG_BK][idx / BG_BK] = ah;
                Al[st][idx % BG_BK][idx / BG_BK] = float(ra[i] - double(ah));
                Bh[st][idx / T][idx % T] = bh;
                Bl[st][idx / T][idx % T] = float(rb[i] - double(bh));
            }
        };
        double acc[BP][BP][16];
#pragma unroll
        for (int p = 0; p < BP; ++p)
#pragma unroll
            for (int c = 0; c < BP; ++c)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[p][c][v] = 0.0;
        fetch(0);
        stash(0);
        __syncthreads();
        int st = 0;
        const int fr = lane & 31, fk = lane >> 5;
        for (int64_t k0 = 0; k0 < pr.K; k0 += BG_BK) {
            const bool more = k0 + BG_BK < pr.K;
            if (more) fetch(k0 + BG_BK);
            mrm_f16 hh[BP][BP], hc[BP][BP];
#pragma unroll
            for (int p = 0; p < BP; ++p)
#pragma unroll
                for (int c = 0; c < BP; ++c) {
                    hh[p][c] = mrm_f16{};
                    hc[p][c] = mrm_f16{};
                }
#pragma unroll
            for (int kk = 0; kk < BG_BK / 2; ++kk) {
                const int k = 2 * kk + fk;
                float a_h[BP], a_l[BP], b_h[BP], b_l[BP];
#pragma unroll
                for (int p = 0; p < BP; ++p) {
                    a_h[p] = Ah[st][k][wm * (T / 2) + 32 * p + fr];
                    a_l[p] = Al[st][k][wm * (T / 2) + 32 * p + fr];
                }
#pragma unroll
                for (int c = 0; c < BP; ++c) {
                    b_h[c] = Bh[st][k][wn * (T / 2) + 32 * c + fr];
                    b_l[c] = Bl[st][k][wn * (T / 2) + 32 * c + fr];
                }
#pragma unroll
                for (int p = 0; p < BP; ++p)
#pragma unroll
                    for (int c = 0; c < BP; ++c) {
                        hh[p][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_h[p], b_h[c], hh[p][c], 0, 0, 0);
                        hc[p][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_h[p], b_l[c], hc[p][c], 0, 0, 0);
                        hc[p][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_l[p], b_h[c], hc[p][c], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int p = 0; p < BP; ++p)
#pragma unroll
                for (int c = 0; c < BP; ++c)
#pragma unroll
                    for (int v = 0; v < 16; ++v) acc[p][c][v] += double(hh[p][c][v]) + double(hc[p][c][v]);
            if (more) stash(st ^ 1);
            __syncthreads();
            st ^= 1;
        }
#pragma unroll
        for (int p = 0; p < BP; ++p)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t row = m0 + wm * (T / 2) + 32 * p + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row >= pr.M) continue;
#pragma unroll
                for (int c = 0; c < BP; ++c) {
                    const int64_t col = n0 + wn * (T / 2) + 32 * c + (lane & 31);
                    if (col < pr.N) pr.C[row * pr.ldc + col] = acc[p][c][v];
                }
            }
    }
}

// Batched small LU solves: the kernels of mrm_batched_lu.h.
using LuP = mrm::BatchLuP;

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
struct Launch {
    const void* fn;
    int grid;
    int block;
    size_t lds;
    std::vector<void*> args;
};

struct Plan {
    int kind = 0;          // 1 GEMM, 2 LU
    int tile = 0;          // GEMM tile chosen
    int mode = 0;          // precision: MRM_SOLVE_FP64 or MRM_SOLVE_FP32_IR
    void* dev = nullptr;   // descriptors, tiles, counters (one allocation)
    unsigned* counters = nullptr;
    int ncounters = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // argument storage referenced by the captured launches
    GemmP* probs = nullptr;
    Tile* tiles = nullptr;
    int ntiles = 0;
    static constexpr int NCLS = 4;
    LuP* sys[NCLS] = {nullptr, nullptr, nullptr, nullptr};
    int counts[NCLS] = {0, 0, 0, 0};
    int nmax[NCLS] = {32, 64, 96, 128};
    int* status = nullptr;
    int refine = 2;
};

int resident_grid(const void* fn, int block, size_t lds) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, lds) != hipSuccess || per < 1) per = 1;
    return cus * per;
}

template <int T>
const void* gemm_fn() {
    return reinterpret_cast<const void*>(&bg_gemm_kernel<T>);
}

const void* gemm_fn_for(int t, int mode) {
    if (mode == MRM_SOLVE_FP32_IR)
        return t == 64 ? reinterpret_cast<const void*>(&bg_gemm_f32x3_kernel<64>)
                       : reinterpret_cast<const void*>(&bg_gemm_f32x3_kernel<128>);
    return t == 32 ? gemm_fn<32>() : t == 64 ? gemm_fn<64>() : gemm_fn<128>();
}

// Tile list for tile size t, problems with the longest K first.
std::vector<Tile> make_tiles(const std::vector<GemmP>& probs, int t) {
    std::vector<int> order(probs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return probs[a].K > probs[b].K; });
    std::vector<Tile> tiles;
    for (int p : order) {
        if (probs[p].M == 0 || probs[p].N == 0) continue;
        for (int64_t m = 0; m < probs[p].M; m += t)
            for (int64_t n = 0; n < probs[p].N; n += t) tiles.push_back(Tile{p, int(m), int(n), 0});
    }
    return tiles;
}

int launch_gemm(Plan& pl, int t, hipStream_t s) {
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, pl.counters, 1);
    const void* fn = gemm_fn_for(t, pl.mode);
    const int grid = resident_grid(fn, BG_THREADS, 0);
    void* args[] = {&pl.probs, &pl.tiles, &pl.ntiles, &pl.counters};
    MRM_HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(BG_THREADS), args, 0, s));
    return MRM_OK;
}

// Systems up to 64 unknowns go one per wave (four per workgroup), larger
// ones one per workgroup of eight waves (mrm_batched_lu_*.hip).
const void* lu_fn(int mode, int cls) {
    static const int nmax[Plan::NCLS] = {32, 64, 96, 128};
    if (cls < 2) return mrm::batched_lu_wave_fn(mode, nmax[cls]);
    if (mode == 0)
        return cls == 2 ? mrm::batched_lu_block_kernel<0, 96>() : mrm::batched_lu_block_kernel<0, 128>();
    return cls == 2 ? mrm::batched_lu_block_kernel<1, 96>() : mrm::batched_lu_block_kernel<1, 128>();
}

// The solve plan's graph is built node by node: the counter reset, then the
// size classes' persistent kernels as independent branches, so the workgroups
// of one class take over the CUs that another class's tail leaves idle.
int build_lu_graph(Plan& pl) {
    hipGraph_t g = nullptr;
    MRM_HIP_TRY(hipGraphCreate(&g, 0));
    auto fail = [&](hipError_t e) {
        (void)hipGraphDestroy(g);
        mrm_set_error(MRM_HIP_ERROR, "batched solve: graph construction failed: %s", hipGetErrorString(e));
        return MRM_HIP_ERROR;
    };
    unsigned* ctrs = pl.counters;
    int nc = Plan::NCLS;
    void* zargs[] = {&ctrs, &nc};
    hipKernelNodeParams zp{};
    zp.func = reinterpret_cast<void*>(zero_counters_kernel);
    zp.gridDim = dim3(1);
    zp.blockDim = dim3(64);
    zp.sharedMemBytes = 0;
    zp.kernelParams = zargs;
    zp.extra = nullptr;
    hipGraphNode_t zero = nullptr;
    hipError_t e = hipGraphAddKernelNode(&zero, g, nullptr, 0, &zp);
    if (e != hipSuccess) return fail(e);
    for (int c = 0; c < Plan::NCLS; ++c) {
        if (pl.counts[c] == 0) continue;
        const void* fn = lu_fn(pl.mode, c);
        const int block = c < 2 ? mrm::BL_THREADS : mrm::BL_BLOCK_THREADS;
        unsigned* ctr = pl.counters + c;
        void* args[] = {&pl.sys[c], &pl.counts[c], &ctr, &pl.status, &pl.refine};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void*>(fn);
        kp.gridDim = dim3(unsigned(resident_grid(fn, block, 0)));
        kp.blockDim = dim3(unsigned(block));
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        if ((e = hipGraphAddKernelNode(&node, g, &zero, 1, &kp)) != hipSuccess) return fail(e);
    }
    pl.graph = g;
    MRM_HIP_TRY(hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0));
    return MRM_OK;
}

int capture(Plan& pl, hipStream_t s) {
    if (pl.kind == 2) return build_lu_graph(pl);
    MRM_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = launch_gemm(pl, pl.tile, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MRM_HIP_TRY(e);
    pl.graph = g;
    MRM_HIP_TRY(hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0));
    return MRM_OK;
}

void destroy(Plan* pl) {
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->dev) (void)hipFree(pl->dev);
    delete pl;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int32_t mrm_batched_dgemm_plan(const mrm_gemm_desc* problems, int64_t count, int32_t tile, void** plan) {
    return mrm_batched_dgemm_plan_ex(problems, count, tile, MRM_SOLVE_FP64, plan);
}

int32_t mrm_batched_dgemm_plan_ex(const mrm_gemm_desc* problems, int64_t count, int32_t tile, int32_t mode,
                                  void** plan) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    const bool f32 = mode == MRM_SOLVE_FP32_IR;
    if (!plan || count < 0 || count > (1 << 24) || (count > 0 && !problems) ||
        (mode != MRM_SOLVE_FP64 && !f32) || (tile != 0 && tile != 64 && tile != 128 && (f32 || tile != 32))) {
        mrm_set_error(MRM_BAD_ARG, "batched gemm: bad plan arguments (tiles 32/64/128 in FP64, 64/128 in FP32)");
        return MRM_BAD_ARG;
    }
    *plan = nullptr;
    std::vector<GemmP> probs(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) {
        const mrm_gemm_desc& d = problems[i];
        if (d.M < 0 || d.N < 0 || d.K < 0 || d.M > (1 << 20) || d.N > (1 << 20) || ((d.lda | d.ldb | d.ldc) & 7) ||
            (d.M > 0 && d.K > 0 && (d.lda < d.K * 8 || !d.A)) || (d.K > 0 && d.N > 0 && (d.ldb < d.N * 8 || !d.B)) ||
            (d.M > 0 && d.N > 0 && (d.ldc < d.N * 8 || !d.C))) {
            mrm_set_error(MRM_BAD_DIMS, "batched gemm: problem %lld has bad sizes, strides or pointers", (long long)i);
            return MRM_BAD_DIMS;
        }
        probs[size_t(i)] = GemmP{d.M, d.N, d.K, d.lda / 8, d.ldb / 8, d.ldc / 8, d.A, d.B, d.C};
    }
    Plan* pl = new Plan;
    pl->kind = 1;
    pl->mode = mode;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    // the largest tile list (T = 32) bounds the device storage
    const std::vector<Tile> t32 = make_tiles(probs, 32);
    const size_t bytes = align256(probs.size() * sizeof(GemmP) + 1) + align256(t32.size() * sizeof(Tile) + 1) + 256;
    if (hipMalloc(&pl->dev, bytes) != hipSuccess) {
        delete pl;
        mrm_set_error(MRM_NO_MEMORY, "batched gemm: cannot allocate the plan");
        return MRM_NO_MEMORY;
    }
    char* base = static_cast<char*>(pl->dev);
    pl->probs = reinterpret_cast<GemmP*>(base);
    pl->tiles = reinterpret_cast<Tile*>(base + align256(probs.size() * sizeof(GemmP) + 1));
    pl->counters = reinterpret_cast<unsigned*>(base + bytes - 256);
    pl->ncounters = 1;
    if (!probs.empty()) rc = hipMemcpyAsync(pl->probs, probs.data(), probs.size() * sizeof(GemmP),
                                            hipMemcpyHostToDevice, s) == hipSuccess ? MRM_OK : MRM_HIP_ERROR;
    auto use = [&](int t) -> int {
        const std::vector<Tile> tl = t == 32 ? t32 : make_tiles(probs, t);
        pl->ntiles = int(tl.size());
        if (!tl.empty() && hipMemcpyAsync(pl->tiles, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice,
                                          s) != hipSuccess) {
            mrm_set_error(MRM_HIP_ERROR, "batched gemm: tile upload failed");
            return MRM_HIP_ERROR;
        }
        return MRM_OK;
    };
    int best = tile;
    if (!rc && tile == 0) {
        // autotune: time each tile size on the batch itself (C = A*B is
        // idempotent), best of two runs each
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        float best_ms = 1e30f;
        for (int t : {32, 64, 128}) {
            if (f32 && t == 32) continue;
            if ((rc = use(t))) break;
            float tbest = 1e30f;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                (void)hipEventRecord(e0, s);
                rc = launch_gemm(*pl, t, s);
                (void)hipEventRecord(e1, s);
                if (!rc && hipEventSynchronize(e1) == hipSuccess) {
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    tbest = std::min(tbest, ms);
                }
            }
            if (tbest < best_ms) {
                best_ms = tbest;
                best = t;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (!rc) rc = use(best);
    pl->tile = best;
    if (!rc) rc = capture(*pl, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MRM_HIP_ERROR;
    if (rc) {
        destroy(pl);
        return rc;
    }
    *plan = pl;
    return MRM_OK;
}

int32_t mrm_batched_lu_plan(const mrm_lu_desc* systems, int64_t count, int32_t mode, int32_t* status, void** plan) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if (!plan || count < 0 || count > (1 << 24) || (count > 0 && !systems) ||
        (mode != MRM_SOLVE_FP64 && mode != MRM_SOLVE_FP32_IR)) {
        mrm_set_error(MRM_BAD_ARG, "batched solve: bad plan arguments (modes: FP64, FP32_IR)");
        return MRM_BAD_ARG;
    }
    *plan = nullptr;
    std::vector<LuP> cls[Plan::NCLS];
    for (int64_t i = 0; i < count; ++i) {
        const mrm_lu_desc& d = systems[i];
        if (d.n < 0 || d.n > 128 || (d.lda & 7) || (d.n > 0 && (d.lda < d.n * 8 || !d.A || !d.b || !d.x))) {
            mrm_set_error(MRM_BAD_DIMS, "batched solve: system %lld has n outside 0..128 or a bad layout",
                          (long long)i);
            return MRM_BAD_DIMS;
        }
        const LuP p{d.n, d.lda / 8, d.A, d.b, d.x, int(i)};
        cls[d.n <= 32 ? 0 : d.n <= 64 ? 1 : d.n <= 96 ? 2 : 3].push_back(p);
    }
    Plan* pl = new Plan;
    pl->kind = 2;
    pl->mode = mode;
    pl->status = status;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    size_t bytes = 256;
    for (auto& c : cls) {
        // largest systems first inside a class
        std::stable_sort(c.begin(), c.end(), [](const LuP& a, const LuP& b) { return a.n > b.n; });
        bytes += align256(c.size() * sizeof(LuP) + 1);
    }
    if (hipMalloc(&pl->dev, bytes) != hipSuccess) {
        delete pl;
        mrm_set_error(MRM_NO_MEMORY, "batched solve: cannot allocate the plan");
        return MRM_NO_MEMORY;
    }
    char* base = static_cast<char*>(pl->dev);
    pl->counters = reinterpret_cast<unsigned*>(base);
    pl->ncounters = Plan::NCLS;
    size_t off = 256;
    for (int c = 0; c < Plan::NCLS; ++c) {
        pl->sys[c] = reinterpret_cast<LuP*>(base + off);
        pl->counts[c] = int(cls[c].size());
        if (!cls[c].empty() &&
            hipMemcpyAsync(pl->sys[c], cls[c].data(), cls[c].size() * sizeof(LuP), hipMemcpyHostToDevice, s) !=
                hipSuccess)
            rc = MRM_HIP_ERROR;
        off += align256(cls[c].size() * sizeof(LuP) + 1);
    }
    pl->tile = 64;
    if (!rc) rc = capture(*pl, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MRM_HIP_ERROR;
    if (rc) {
        if (rc == MRM_HIP_ERROR) mrm_set_error(MRM_HIP_ERROR, "batched solve: plan set-up failed");
        destroy(pl);
        return rc;
    }
    *plan = pl;
    return MRM_OK;
}

int32_t mrm_batched_run(void* plan, void* stream) {
    if (!plan) {
        mrm_set_error(MRM_BAD_ARG, "null plan");
        return MRM_BAD_ARG;
    }
    Plan* pl = static_cast<Plan*>(plan);
    MRM_HIP_TRY(hipGraphLaunch(pl->exec, mrm_resolve_stream(stream)));
    return MRM_OK;
}

int32_t mrm_batched_plan_tile(void* plan) { return plan ? static_cast<Plan*>(plan)->tile : 0; }

void mrm_batched_destroy(void* plan) {
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    destroy(static_cast<Plan*>(plan));
}

}  // extern "C"
