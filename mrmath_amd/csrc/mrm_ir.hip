// mrm_ir.hip -- linear solve with a low-precision LU and FP64 iterative
// refinement (LU-IR), and its plain FP64 counterpart.
//
// The reference solves A X = B by LU, per-column back substitution and
// NumRefinments fixed FP64 refinement steps (MatrixLinEQSolve,
// src/LinAlgLU.pas:481-592).  Here the factorisation may run its trailing
// updates on the FP32 or BF16 matrix cores (panel, pivot search and storage
// stay FP64), and refinement runs until the normwise backward error
//   max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
// reaches the tolerance, every residual formed in FP64 from the original A:
//   x = (LU)^-1 P b;  repeat r = b - A x;  x += (LU)^-1 P r.
// When a plain step gains less than tenfold (kappa(A) u_f near or above 1 for
// the factors' precision u_f), the corrections d ~= A^-1 r come from GMRES
// preconditioned by the same factors instead (GMRES-IR, mrm_gmres.hip); if a
// GMRES step also gains less than tenfold, the factorisation is redone in FP64.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mrm_common.h"
#include "mrm_gmres.h"
#include "mrm_host.h"
#include "mrm_ir.h"
#include "mrm_lu.h"

namespace {

// A plain refinement step of low-precision factors must shrink the backward
// error at least this much, or the corrections switch to GMRES: each GMRES
// Krylov step costs one solve, like a plain step, but contracts much faster
// when kappa(A) u_f is near 1 (plain steps then gain less than a factor 10).
constexpr double IR_PLAIN_RATE = 0.1;

// The norms are combined with atomicMax on the bit patterns of non-negative
// doubles, which order like the values; a NaN is recorded as the quiet NaN
// 0x7ff8..., which orders above +inf, so it reaches the host.
constexpr unsigned long long NAN_BITS = 0x7ff8000000000000ull;

__device__ __forceinline__ void max_bits(unsigned long long* out, double v) {
    if (isnan(v)) atomicMax(out, NAN_BITS);
    else if (v > 0.0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(v)));
}

// out[c] = max over rows of |M(r, c)| (NaN if the column holds one); out zeroed.
__global__ void __launch_bounds__(256) colmax_kernel(const double* __restrict__ Mx, int64_t ld, int64_t rows,
                                                     int64_t cols, unsigned long long* __restrict__ out) {
    const int64_t r0 = int64_t(blockIdx.x) * 256;
    const int64_t r1 = r0 + 256 < rows ? r0 + 256 : rows;
    for (int64_t c = threadIdx.x % 64; c < cols; c += 64) {
        double m = 0.0;
        bool nan = false;
        for (int64_t r = r0 + threadIdx.x / 64; r < r1; r += 4) {
            const double a = fabs(Mx[r * ld + c]);
            nan |= isnan(a);
            m = fmax(m, a);
        }
        max_bits(out + c, nan ? NAN : m);
    }
}

// out = max_r sum_c |A(r, c)| (one wave per row; a NaN in the sum propagates)
__global__ void __launch_bounds__(256) norm_inf_kernel(const double* __restrict__ A, int64_t lda, int64_t n,
                                                       unsigned long long* __restrict__ out) {
    const int64_t r = int64_t(blockIdx.x) * 4 + threadIdx.x / 64;
    if (r >= n) return;
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int64_t c = lane; c < n; c += 64) s += fabs(A[r * lda + c]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) max_bits(out, s);
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Library-owned buffers of the refinement (grow-only; used under the host mutex).
struct IrWork {
    double* lu = nullptr;
    size_t lu_elems = 0;
    double* r = nullptr;
    size_t r_elems = 0;
    double* d = nullptr;
    size_t d_elems = 0;
    int* indx = nullptr;
    size_t indx_elems = 0;
    int* perm = nullptr;
    size_t perm_elems = 0;
    unsigned long long* norms = nullptr;  // [anorm][bn nrhs][rn nrhs][xn nrhs]
    size_t norm_elems = 0;
    double* dinv = nullptr;               // diagonal-block inverses of the factors
    size_t dinv_elems = 0;
};

template <typename P>
int grow(P*& p, size_t& have, size_t need, size_t elem) {
    if (need <= have) return MRM_OK;
    if (p) MRM_HIP_TRY(hipFree(p));
    p = nullptr;
    have = 0;
    if (hipMalloc(reinterpret_cast<void**>(&p), need * elem) != hipSuccess) {
        mrm_set_error(MRM_NO_MEMORY, "solve: cannot allocate %zu bytes", need * elem);
        return MRM_NO_MEMORY;
    }
    have = need;
    return MRM_OK;
}

IrWork& ir_work() {
    static IrWork w;
    return w;
}

// Factor a copy of A in `mode`, then solve and refine; returns MRM_OK,
// MRM_SINGULAR, MRM_NOT_CONVERGED or an error.
int factor_solve_refine(int mode, const double* A, int64_t lda, int n, const double* B, int64_t ldb, double* X,
                        int64_t ldx, int nrhs, int max_iter, double tol, mrm::IrStats& st, hipStream_t s) {
    IrWork& w = ir_work();
    mrm::LuWork* ws;
    int rc;
    if ((rc = mrm::lu_workspace(&ws))) return rc;
    if ((rc = mrm::lu_lowp_reserve(*ws, n, mode))) return rc;
    const int64_t ldr = nrhs;
    // the factorisation works on a copy of A, so a relaunch copies it again
    auto copy_a = [&]() -> int {
        MRM_HIP_TRY(hipMemcpy2DAsync(w.lu, size_t(n) * 8, A, size_t(lda) * 8, size_t(n) * 8, size_t(n),
                                     hipMemcpyDeviceToDevice, s));
        return MRM_OK;
    };
    if ((rc = copy_a())) return rc;
    if ((rc = mrm::lu_factor_sync(w.lu, n, n, w.indx, w.perm, *ws, s, mode, copy_a))) return rc;
    // the factors serve every refinement step: invert their diagonal blocks once
    double* dinv = mrm::lu_solve_inverse_elems(n) ? w.dinv : nullptr;
    if (dinv && (rc = mrm::lu_solve_inverses(w.lu, n, n, dinv, s))) return rc;
    if ((rc = mrm::lu_solve<double>(w.lu, n, n, w.perm, B, ldb, X, ldx, nrhs, s, dinv))) return rc;

    unsigned long long* anorm = w.norms;
    unsigned long long* bn = anorm + 1;
    unsigned long long* rn = bn + nrhs;
    unsigned long long* xn = rn + nrhs;
    const unsigned rowblocks = unsigned((n + 255) / 256);
    hipLaunchKernelGGL(zero_u64_kernel, dim3((1 + 3 * nrhs + 255) / 256), dim3(256), 0, s, w.norms, 1 + 3 * nrhs);
    hipLaunchKernelGGL(norm_inf_kernel, dim3(unsigned((n + 3) / 4)), dim3(256), 0, s, A, lda, int64_t(n), anorm);
    hipLaunchKernelGGL(colmax_kernel, dim3(rowblocks), dim3(256), 0, s, B, ldb, int64_t(n), int64_t(nrhs), bn);
    std::vector<unsigned long long> h(size_t(1 + 3 * nrhs));
    double prev = INFINITY;
    bool gmres = false;  // corrections from GMRES preconditioned by the factors
    for (int it = 0;; ++it) {
        // r = b - A x in FP64
        MRM_HIP_TRY(hipMemcpy2DAsync(w.r, size_t(ldr) * 8, B, size_t(ldb) * 8, size_t(nrhs) * 8, size_t(n),
                                     hipMemcpyDeviceToDevice, s));
        if ((rc = mrm_launch_dgemm(MRM_NN, n, nrhs, n, A, lda * 8, X, ldx * 8, w.r, ldr * 8, MRM_DO_SUB, s)))
            return rc;
        hipLaunchKernelGGL(zero_u64_kernel, dim3((2 * nrhs + 255) / 256), dim3(256), 0, s, rn, 2 * nrhs);
        hipLaunchKernelGGL(colmax_kernel, dim3(rowblocks), dim3(256), 0, s, w.r, ldr, int64_t(n), int64_t(nrhs), rn);
        hipLaunchKernelGGL(colmax_kernel, dim3(rowblocks), dim3(256), 0, s, X, ldx, int64_t(n), int64_t(nrhs), xn);
        MRM_HIP_TRY(hipMemcpyAsync(h.data(), w.norms, h.size() * 8, hipMemcpyDeviceToHost, s));
        MRM_HIP_TRY(hipStreamSynchronize(s));
        auto val = [&](size_t i) {
            double v;
            memcpy(&v, &h[i], 8);
            return v;
        };
        const double an = val(0);
        double berr = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            const double den = an * val(size_t(1 + 2 * nrhs + c)) + val(size_t(1 + c));
            const double rv = val(size_t(1 + nrhs + c));
            const double e = den > 0.0 ? rv / den : rv;
            if (!(e <= berr)) berr = e;  // a NaN stays: neither converged nor improving
        }
        st.iterations = it;
        st.backward_error = berr;
        if (berr <= tol) return MRM_OK;
        if (it >= max_iter || std::isnan(berr)) return MRM_NOT_CONVERGED;
        if (!(berr < IR_PLAIN_RATE * prev)) {
            // the step gained less than tenfold: after a plain step with
            // low-precision factors continue with GMRES corrections
            // preconditioned by them; after a GMRES step (GMRES did not reach
            // its tolerance: the factors are too poor a preconditioner, and
            // FP64 factors are the faster way) or a stall in FP64 (less than
            // twofold) the run ends
            if (gmres || mode == MRM_SOLVE_FP64) {
                if (gmres || !(berr < 0.5 * prev)) return MRM_NOT_CONVERGED;
            } else {
                gmres = true;
                st.gmres = 1;
            }
        }
        prev = berr;
        if (gmres) {
            int k = 0;
            rc = mrm::gmres_correction(A, lda, n, w.lu, w.perm, dinv, w.r, ldr, w.d, nrhs, mrm::IR_GMRES_DIM,
                                       mrm::IR_GMRES_TOL, &k, s);
            st.krylov_steps += k;
        } else {
            rc = mrm::lu_solve<double>(w.lu, n, n, w.perm, w.r, ldr, w.d, ldr, nrhs, s, dinv);
        }
        if (rc) return rc;
        mrm::add_inplace<double>(X, ldx, w.d, ldr, n, nrhs, s);
    }
}

}  // namespace

namespace mrm {

void ir_release_workspace() {
    IrWork& w = ir_work();
    (void)hipFree(w.lu);
    (void)hipFree(w.r);
    (void)hipFree(w.d);
    (void)hipFree(w.indx);
    (void)hipFree(w.perm);
    (void)hipFree(w.norms);
    (void)hipFree(w.dinv);
    w = IrWork();
}

int lineq_solve_ir(int mode, const double* A, int64_t lda, int n, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int nrhs, int max_iter, double tol, IrStats* stats, hipStream_t s) {
    IrStats st;
    if (mode != MRM_SOLVE_FP64 && mode != MRM_SOLVE_FP32_IR && mode != MRM_SOLVE_BF16_IR &&
        mode != MRM_SOLVE_BF16X3_IR) {
        mrm_set_error(MRM_BAD_ARG, "solve: unknown mode %d", mode);
        return MRM_BAD_ARG;
    }
    if (n == 0 || nrhs == 0) {
        if (stats) *stats = st;
        return MRM_OK;
    }
    IrWork& w = ir_work();
    int rc;
    const size_t rhs = size_t(n) * size_t(nrhs);
    if ((rc = grow(w.lu, w.lu_elems, size_t(n) * size_t(n), 8))) return rc;
    if ((rc = grow(w.r, w.r_elems, rhs, 8))) return rc;
    if ((rc = grow(w.d, w.d_elems, rhs, 8))) return rc;
    if ((rc = grow(w.indx, w.indx_elems, size_t(n), 4))) return rc;
    if ((rc = grow(w.perm, w.perm_elems, size_t(n), 4))) return rc;
    if ((rc = grow(w.norms, w.norm_elems, size_t(1 + 3 * nrhs), 8))) return rc;
    if ((rc = grow(w.dinv, w.dinv_elems, mrm::lu_solve_inverse_elems(n), 8))) return rc;
    rc = factor_solve_refine(mode, A, lda, n, B, ldb, X, ldx, nrhs, max_iter, tol, st, s);
    if ((rc == MRM_NOT_CONVERGED || rc == MRM_SINGULAR) && mode != MRM_SOLVE_FP64) {
        // the low-precision factors cannot reach the tolerance for this
        // conditioning (or hit an exactly zero pivot the FP64 updates might
        // not): redo the factorisation in FP64
        const IrStats first = st;
        st = IrStats();
        rc = factor_solve_refine(MRM_SOLVE_FP64, A, lda, n, B, ldb, X, ldx, nrhs, max_iter, tol, st, s);
        st.lowp_iterations = first.iterations;
        st.lowp_backward_error = first.backward_error;
        st.lowp_krylov_steps = first.krylov_steps;
        st.fell_back = 1;
    }
    if (stats) *stats = st;
    if (rc == MRM_NOT_CONVERGED)
        mrm_set_error(MRM_NOT_CONVERGED, "solve: backward error %.3e above the tolerance %.3e after %d steps",
                      st.backward_error, tol, st.iterations);
    return rc;
}

}  // namespace mrm
