// mrm_gmres.hip -- GMRES correction solves preconditioned by the low-precision
// LU factors (the inner solver of GMRES-based iterative refinement, Carson and
// Higham, "A new analysis of iterative refinement and its application to
// accurate solution of ill-conditioned sparse linear systems", SISC 2017, and
// "Accelerating the solution of linear systems by iterative refinement in
// three precisions", SISC 2018).
//
// Plain refinement x += (LU)^-1 P r needs kappa(A) * u_f < 1 for the precision
// u_f the factors were computed in; left-preconditioned GMRES on
// (LU)^-1 P A d = (LU)^-1 P r, run in FP64 with every product by A in FP64,
// converges while kappa(A) u_f is well above 1, because the preconditioned
// matrix stays close to the identity in the directions that matter.  The
// reference refines with the plain step only (MatrixLinEQSolve,
// src/LinAlgLU.pas:542-586).
//
// Every right-hand side gets its own Krylov space; the nrhs spaces advance in
// lockstep so each step is one FP64 GEMM A * V_k (n x n times n x nrhs), one
// multi-column LU solve and column-wise Gram-Schmidt (classical, twice) on the
// GPU, while the small Hessenberg least-squares problems (Givens rotations)
// run on the host from the column dot products.  Column sums are
// deterministic: partial sums per row block, added in block order.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mrm_common.h"
#include "mrm_gmres.h"
#include "mrm_lu.h"

namespace {

constexpr int GM_ROWS = 256;  // rows per partial-sum block
constexpr int GM_PROBE = 10;  // Krylov steps before a correction that gains less than 10x is given up

// part[b][i][c] = sum over the rows r of block b of V_i[r][c] * W[r][c], i < nv.
// Matrices n x nrhs with row stride nrhs; V_i = V + i * vstride.  Thread t:
// column lane t % 64, row lane t / 64.
__global__ void __launch_bounds__(256) gm_dots_kernel(const double* __restrict__ V, int64_t vstride,
                                                      const double* __restrict__ W, int n, int nrhs, int nv,
                                                      double* __restrict__ part) {
    __shared__ double sh[4][64];
    const int cl = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const int64_t r0 = int64_t(blockIdx.x) * GM_ROWS;
    const int64_t r1 = r0 + GM_ROWS < n ? r0 + GM_ROWS : n;
    for (int c0 = 0; c0 < nrhs; c0 += 64) {
        const int c = c0 + cl;
        for (int i = 0; i < nv; ++i) {
            double acc = 0.0;
            if (c < nrhs) {
                const double* vi = V + i * vstride;
                for (int64_t r = r0 + rl; r < r1; r += 4) acc = fma(vi[r * nrhs + c], W[r * nrhs + c], acc);
            }
            sh[rl][cl] = acc;
            __syncthreads();
            if (rl == 0 && c < nrhs)
                part[(int64_t(blockIdx.x) * nv + i) * nrhs + c] = ((sh[0][cl] + sh[1][cl]) + sh[2][cl]) + sh[3][cl];
            __syncthreads();
        }
    }
}

// out[i][c] (+)= sum_b part[b][i][c], blocks in order
__global__ void __launch_bounds__(256) gm_sum_kernel(const double* __restrict__ part, int nblocks, int nv, int nrhs,
                                                     double* __restrict__ out, int accumulate) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nv * nrhs) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[int64_t(b) * nv * nrhs + idx];
    out[idx] = accumulate ? out[idx] + s : s;
}

// W[r][c] -= sum_{i < nv} V_i[r][c] * h[i][c]
__global__ void __launch_bounds__(256) gm_orth_kernel(const double* __restrict__ V, int64_t vstride,
                                                      double* __restrict__ W, int64_t elems, int nrhs, int nv,
                                                      const double* __restrict__ h) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= elems) return;
    const int c = int(e % nrhs);
    double w = W[e];
    for (int i = 0; i < nv; ++i) w = fma(-V[i * vstride + e], h[i * nrhs + c], w);
    W[e] = w;
}

// W[r][c] *= 1 / sqrt(ss[c]) (0 where ss[c] is 0 or not finite: a finished or
// broken-down column keeps finite, unused basis vectors)
__global__ void __launch_bounds__(256) gm_scale_kernel(double* __restrict__ W, int64_t elems, int nrhs,
                                                       const double* __restrict__ ss) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= elems) return;
    const double q = ss[e % nrhs];
    W[e] = (q > 0.0 && q < INFINITY) ? W[e] / sqrt(q) : 0.0;
}

// D[r][c] = sum_{i < nv} V_i[r][c] * y[i][c]
__global__ void __launch_bounds__(256) gm_combine_kernel(const double* __restrict__ V, int64_t vstride,
                                                         double* __restrict__ D, int64_t elems, int nrhs, int nv,
                                                         const double* __restrict__ y) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= elems) return;
    const int c = int(e % nrhs);
    double d = 0.0;
    for (int i = 0; i < nv; ++i) d = fma(V[i * vstride + e], y[i * nrhs + c], d);
    D[e] = d;
}

struct GmWork {
    double* V = nullptr;     // (m+1) basis blocks, then one product block
    size_t V_elems = 0;
    double* part = nullptr;  // partial column sums
    size_t part_elems = 0;
    double* small = nullptr; // h (m+1 x nrhs), h2, ss (nrhs), y (m+1 x nrhs)
    size_t small_elems = 0;
};

GmWork& gm_work() {
    static GmWork w;
    return w;
}

int grow(double*& p, size_t& have, size_t need) {
    if (need <= have) return MRM_OK;
    if (p) MRM_HIP_TRY(hipFree(p));
    p = nullptr;
    have = 0;
    if (hipMalloc(reinterpret_cast<void**>(&p), need * 8) != hipSuccess) {
        mrm_set_error(MRM_NO_MEMORY, "gmres: cannot allocate %zu bytes", need * 8);
        return MRM_NO_MEMORY;
    }
    have = need;
    return MRM_OK;
}

unsigned blocks_for(int64_t elems) { return unsigned((elems + 255) / 256); }

}  // namespace

namespace mrm {

void gmres_release_workspace() {
    GmWork& w = gm_work();
    (void)hipFree(w.V);
    (void)hipFree(w.part);
    (void)hipFree(w.small);
    w = GmWork();
}

int gmres_correction(const double* A, int64_t lda, int n, const double* LU, const int* perm, const double* dinv,
                     const double* R, int64_t ldr, double* D, int nrhs, int m, double tol, int* steps,
                     hipStream_t s) {
    if (steps) *steps = 0;
    if (m < 1 || m > MRM_GMRES_MAX_DIM) {
        mrm_set_error(MRM_BAD_ARG, "gmres: Krylov dimension %d outside 1..%d", m, MRM_GMRES_MAX_DIM);
        return MRM_BAD_ARG;
    }
    GmWork& w = gm_work();
    const int64_t blk = int64_t(n) * nrhs;       // elements of one n x nrhs block
    const int nblocks = (n + GM_ROWS - 1) / GM_ROWS;
    int rc;
    if ((rc = grow(w.V, w.V_elems, size_t(m + 2) * size_t(blk)))) return rc;
    if ((rc = grow(w.part, w.part_elems, size_t(nblocks) * size_t(m + 1) * size_t(nrhs)))) return rc;
    if ((rc = grow(w.small, w.small_elems, size_t(2 * (m + 1) + 1 + (m + 1)) * size_t(nrhs)))) return rc;
    double* V = w.V;
    double* P = w.V + int64_t(m + 1) * blk;       // A * V_k
    double* h = w.small;
    double* h2 = h + (m + 1) * nrhs;
    double* ss = h2 + (m + 1) * nrhs;
    double* y = ss + nrhs;
    const unsigned eb = blocks_for(blk);

    // column dot products of V_0..V_{nv-1} with W into out (accumulate: +=)
    auto dots = [&](const double* W, int nv, double* out, int accumulate) {
        hipLaunchKernelGGL(gm_dots_kernel, dim3(unsigned(nblocks)), dim3(256), 0, s, V, blk, W, n, nrhs, nv,
                           w.part);
        hipLaunchKernelGGL(gm_sum_kernel, dim3(blocks_for(int64_t(nv) * nrhs)), dim3(256), 0, s, w.part, nblocks,
                           nv, nrhs, out, accumulate);
    };
    auto sumsq = [&](const double* W) {
        hipLaunchKernelGGL(gm_dots_kernel, dim3(unsigned(nblocks)), dim3(256), 0, s, W, blk, W, n, nrhs, 1, w.part);
        hipLaunchKernelGGL(gm_sum_kernel, dim3(blocks_for(nrhs)), dim3(256), 0, s, w.part, nblocks, 1, nrhs, ss, 0);
    };

    // host side of every column: Givens-reduced Hessenberg columns, rotations,
    // right-hand side g, the Krylov dimension it stopped at
    struct Col {
        std::vector<double> Rm;  // (m) x (m) upper triangle, column-major by step
        std::vector<double> cs, sn, g;
        double beta = 0.0;
        int k = 0;               // basis vectors in use
        bool active = true;
    };
    std::vector<Col> col(static_cast<size_t>(nrhs));
    for (Col& c : col) {
        c.Rm.assign(size_t(m) * size_t(m), 0.0);
        c.cs.assign(size_t(m), 0.0);
        c.sn.assign(size_t(m), 0.0);
        c.g.assign(size_t(m + 1), 0.0);
    }
    std::vector<double> hh(size_t(m + 1) * size_t(nrhs)), hs(static_cast<size_t>(nrhs));

    // V_0 = (LU)^-1 P r, beta = ||V_0||_2, V_0 /= beta
    if ((rc = lu_solve<double>(LU, n, n, perm, R, ldr, V, nrhs, nrhs, s, dinv))) return rc;
    sumsq(V);
    MRM_HIP_TRY(hipMemcpyAsync(hs.data(), ss, size_t(nrhs) * 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(gm_scale_kernel, dim3(eb), dim3(256), 0, s, V, blk, nrhs, ss);
    MRM_HIP_TRY(hipStreamSynchronize(s));
    int active = 0;
    double worst = 1.0;  // largest relative preconditioned residual of an active column
    for (int c = 0; c < nrhs; ++c) {
        Col& cc = col[size_t(c)];
        cc.beta = std::sqrt(hs[size_t(c)]);
        cc.g[0] = cc.beta;
        cc.active = cc.beta > 0.0 && std::isfinite(cc.beta);
        if (!std::isfinite(cc.beta)) cc.k = -1;  // NaN / inf residual: D column made NaN below
        active += cc.active ? 1 : 0;
    }

    int k = 0;
    for (; k < m && active > 0; ++k) {
        // after GM_PROBE steps without a tenfold reduction of every column's
        // residual the factors are too poor a preconditioner to finish within
        // the dimension: stop and return the best correction of the space so far
        if (k >= GM_PROBE && worst > 0.1) break;
        double* Vk = V + int64_t(k) * blk;
        double* Vn = V + int64_t(k + 1) * blk;
        // Vn = (LU)^-1 P (A V_k)
        if ((rc = mrm_launch_dgemm(MRM_NN, n, nrhs, n, A, lda * 8, Vk, int64_t(nrhs) * 8, P, int64_t(nrhs) * 8,
                                   MRM_DO_NONE, s)))
            return rc;
        if ((rc = lu_solve<double>(LU, n, n, perm, P, nrhs, Vn, nrhs, nrhs, s, dinv))) return rc;
        // classical Gram-Schmidt twice against V_0..V_k
        dots(Vn, k + 1, h, 0);
        hipLaunchKernelGGL(gm_orth_kernel, dim3(eb), dim3(256), 0, s, V, blk, Vn, blk, nrhs, k + 1, h);
        dots(Vn, k + 1, h2, 0);
        hipLaunchKernelGGL(gm_orth_kernel, dim3(eb), dim3(256), 0, s, V, blk, Vn, blk, nrhs, k + 1, h2);
        hipLaunchKernelGGL(gm_sum_kernel, dim3(blocks_for(int64_t(k + 1) * nrhs)), dim3(256), 0, s, h2, 1, k + 1,
                           nrhs, h, 1);
        sumsq(Vn);
        MRM_HIP_TRY(hipMemcpyAsync(hh.data(), h, size_t(k + 1) * size_t(nrhs) * 8, hipMemcpyDeviceToHost, s));
        MRM_HIP_TRY(hipMemcpyAsync(hs.data(), ss, size_t(nrhs) * 8, hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(gm_scale_kernel, dim3(eb), dim3(256), 0, s, Vn, blk, nrhs, ss);
        MRM_HIP_TRY(hipStreamSynchronize(s));
        if (steps) *steps = k + 1;
        worst = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            Col& cc = col[size_t(c)];
            if (!cc.active) continue;
            std::vector<double> hc(size_t(k + 2));
            for (int i = 0; i <= k; ++i) hc[size_t(i)] = hh[size_t(i) * size_t(nrhs) + size_t(c)];
            hc[size_t(k + 1)] = std::sqrt(hs[size_t(c)]);
            for (int j = 0; j < k; ++j) {
                const double a = hc[size_t(j)], b = hc[size_t(j + 1)];
                hc[size_t(j)] = cc.cs[size_t(j)] * a + cc.sn[size_t(j)] * b;
                hc[size_t(j + 1)] = -cc.sn[size_t(j)] * a + cc.cs[size_t(j)] * b;
            }
            const double a = hc[size_t(k)], b = hc[size_t(k + 1)];
            const double rr = std::hypot(a, b);
            const double cs = rr > 0.0 ? a / rr : 1.0, sn = rr > 0.0 ? b / rr : 0.0;
            cc.cs[size_t(k)] = cs;
            cc.sn[size_t(k)] = sn;
            hc[size_t(k)] = rr;
            for (int i = 0; i <= k; ++i) cc.Rm[size_t(k) * size_t(m) + size_t(i)] = hc[size_t(i)];
            cc.g[size_t(k + 1)] = -sn * cc.g[size_t(k)];
            cc.g[size_t(k)] = cs * cc.g[size_t(k)];
            cc.k = k + 1;
            const double res = std::fabs(cc.g[size_t(k + 1)]);
            worst = std::fmax(worst, res / cc.beta);
            // converged, broken down (the exact correction is in the space),
            // or no longer finite
            if (!(res > tol * cc.beta) || !(hc[size_t(k + 1)] >= 0.0) || !(rr > 0.0)) {
                cc.active = false;
                --active;
            }
        }
    }

    // y_c = R_c^-1 g_c on the host, D = sum_i V_i y_i
    std::vector<double> yh(size_t(m + 1) * size_t(nrhs), 0.0);
    for (int c = 0; c < nrhs; ++c) {
        Col& cc = col[size_t(c)];
        if (cc.k < 0) {
            yh[size_t(c)] = NAN;
            continue;
        }
        for (int i = cc.k - 1; i >= 0; --i) {
            double t = cc.g[size_t(i)];
            for (int j = i + 1; j < cc.k; ++j) t -= cc.Rm[size_t(j) * size_t(m) + size_t(i)] * yh[size_t(j) * size_t(nrhs) + size_t(c)];
            const double d = cc.Rm[size_t(i) * size_t(m) + size_t(i)];
            yh[size_t(i) * size_t(nrhs) + size_t(c)] = d != 0.0 ? t / d : 0.0;
        }
    }
    const int nv = k > 0 ? k : 1;
    MRM_HIP_TRY(hipMemcpyAsync(y, yh.data(), size_t(nv) * size_t(nrhs) * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gm_combine_kernel, dim3(eb), dim3(256), 0, s, V, blk, D, blk, nrhs, nv, y);
    MRM_HIP_TRY(hipGetLastError());
    // yh is read by the copy above: finish before it goes out of scope
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

}  // namespace mrm
