// mrm_lu_api.hip -- C-ABI entries of the LU factorisation and linear solver.
//
// Host-pointer entries keep the reference's signatures and semantics
// (MatrixLUDecompInPlace, MatrixLUBackSubst, MatrixLinEQSolve,
// src/LinAlgLU.pas:283-592); device-pointer entries run on a caller's stream.

#include <atomic>
#include <cstring>
#include <vector>

#include "mrm_common.h"
#include "mrm_host.h"
#include "mrm_ir.h"
#include "mrm_lowp.h"
#include "mrm_lu.h"

namespace mrm {

namespace {
LuWork g_lu_ws;
bool g_lu_ready = false;
// deadline of plain panel launches in 100 MHz ticks (the test hook sets it)
std::atomic<uint64_t> g_plain_timeout{LU_PANEL_TIMEOUT_TICKS};
std::atomic<int64_t> g_relaunches{0};
}  // namespace

int lu_workspace(LuWork** out) {
    if (!g_lu_ready) {
        MRM_HIP_TRY(hipMalloc(&g_lu_ws.status, 256));
        MRM_HIP_TRY(hipMalloc(&g_lu_ws.xch, lu_xch_bytes()));
        g_lu_ready = true;
    }
    g_lu_ws.plain_timeout = g_plain_timeout.load();
    *out = &g_lu_ws;
    return MRM_OK;
}

void lu_release_workspace() {
    if (!g_lu_ready) return;
    (void)hipFree(g_lu_ws.status);
    (void)hipFree(g_lu_ws.xch);
    if (g_lu_ws.lowp) (void)hipFree(g_lu_ws.lowp);
    if (g_lu_ws.backup) (void)hipFree(g_lu_ws.backup);
    g_lu_ws = LuWork();
    g_lu_ready = false;
}

uint64_t lu_plain_panel_timeout() { return g_plain_timeout.load(); }
void lu_count_relaunch() { g_relaunches.fetch_add(1); }

}  // namespace mrm

namespace {

// Factor the staged copy dA on the device and read the status (synchronises
// s); a panel timeout is retried after restore() has put the original matrix
// back into dA.
template <typename Restore>
int factor_and_status(double* dA, int64_t lda, int n, int* dindx, int* dperm, hipStream_t s, Restore restore) {
    mrm::LuWork* ws;
    int rc = mrm::lu_workspace(&ws);
    if (rc) return rc;
    return mrm::lu_factor_sync(dA, lda, n, dindx, dperm, *ws, s, MRM_SOLVE_FP64, restore);
}

// Host-pointer entries: the original is the caller's host matrix.
int factor_host_copy(double* dA, int64_t n, const double* A, int64_t LineWidthA, int* dindx, int* dperm,
                     hipStream_t s) {
    auto upload = [&]() -> int {
        return mrm_copy_h2d(dA, size_t(n) * 8, A, size_t(LineWidthA), size_t(n) * 8, size_t(n), s);
    };
    int rc = upload();
    if (rc) return rc;
    return factor_and_status(dA, n, int(n), dindx, dperm, s, upload);
}

int check_square(int64_t n, int64_t lw) {
    if (n < 0 || n > 0x7fffffffLL) {
        mrm_set_error(MRM_BAD_DIMS, "Dimension error");
        return MRM_BAD_DIMS;
    }
    if (lw < n * 8 || (lw & 7)) {
        mrm_set_error(MRM_BAD_DIMS, "Dimension Error: LineWidth < width*8");
        return MRM_BAD_DIMS;
    }
    return MRM_OK;
}

// Row i of P*A is row perm[i] of A, from the reference's sequential indx.
std::vector<int> perm_from_indx(const int32_t* indx, int64_t n) {
    std::vector<int> p(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) p[size_t(i)] = int(i);
    for (int64_t i = 0; i < n; ++i) {
        const int r = indx[i];
        if (r != i) std::swap(p[size_t(i)], p[size_t(r)]);
    }
    return p;
}

}  // namespace

extern "C" {

int64_t mrm_lu_panel_test_deadline(int64_t ticks) {
    mrm::g_plain_timeout.store(ticks < 0 ? mrm::LU_PANEL_TIMEOUT_TICKS : uint64_t(ticks));
    return mrm::g_relaunches.load();
}

int32_t mrm_lu_lookahead_available(void) {
    if (mrm_ensure_init()) return 0;
    return mrm::lu_lookahead_usable() ? 1 : 0;
}

int32_t mrm_dev_lu_factor(double* A, int64_t lda, int64_t n, int32_t* indx, int32_t* perm, void* stream) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(n, lda))) return rc;
    if (n == 0) return MRM_OK;
    if (!A || !indx) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_resolve_stream(stream);
    mrm::LuWork* ws;
    if ((rc = mrm::lu_workspace(&ws))) return rc;
    // A is factored in place: keep a copy for the relaunch after a panel timeout
    const size_t elems = size_t(n) * size_t(n);
    if (ws->backup_elems < elems) {
        if (ws->backup) MRM_HIP_TRY(hipFree(ws->backup));
        ws->backup = nullptr;
        ws->backup_elems = 0;
        if (hipMalloc(&ws->backup, elems * 8) != hipSuccess) {
            mrm_set_error(MRM_NO_MEMORY, "lu: cannot allocate %zu bytes for the copy of A", elems * 8);
            return MRM_NO_MEMORY;
        }
        ws->backup_elems = elems;
    }
    double* backup = ws->backup;
    MRM_HIP_TRY(hipMemcpy2DAsync(backup, size_t(n) * 8, A, size_t(lda), size_t(n) * 8, size_t(n),
                                 hipMemcpyDeviceToDevice, s));
    auto restore = [&]() -> int {
        MRM_HIP_TRY(hipMemcpy2DAsync(A, size_t(lda), backup, size_t(n) * 8, size_t(n) * 8, size_t(n),
                                     hipMemcpyDeviceToDevice, s));
        return MRM_OK;
    };
    return factor_and_status(A, lda / 8, int(n), indx, perm, s, restore);
}

int32_t mrm_dev_lu_solve(const double* LU, int64_t lda, int64_t n, const int32_t* perm, const double* B, int64_t ldb,
                         double* X, int64_t ldx, int64_t nrhs, void* stream) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(n, lda))) return rc;
    if (nrhs < 0 || ldb < nrhs * 8 || ldx < nrhs * 8 || ((ldb | ldx) & 7)) {
        mrm_set_error(MRM_BAD_DIMS, "solve: bad right-hand-side layout");
        return MRM_BAD_DIMS;
    }
    if (n == 0 || nrhs == 0) return MRM_OK;
    return mrm::lu_solve<double>(LU, lda / 8, int(n), perm, B, ldb / 8, X, ldx / 8, int(nrhs),
                                 mrm_resolve_stream(stream));
}

// MatrixLUDecompInPlace (src/LinAlgLU.pas:283-295)
int32_t mrm_lu_decomp_inplace(double* A, int64_t LineWidthA, int64_t width, int32_t* indx) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthA))) return rc;
    if (width == 0) return MRM_OK;
    if (!A || !indx) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    const int64_t n = width;
    void *dA, *dI;
    if ((rc = mrm_staging(0, size_t(n * n) * 8, &dA))) return rc;
    if ((rc = mrm_staging(5, size_t(n) * 4, &dI))) return rc;
    rc = factor_host_copy(static_cast<double*>(dA), n, A, LineWidthA, static_cast<int*>(dI), nullptr, s);
    if (rc != MRM_OK && rc != MRM_SINGULAR) return rc;
    // the reference leaves A partially factored on leSingular too (rc keeps the status)
    if (int cr = mrm_copy_d2h(A, size_t(LineWidthA), dA, size_t(n) * 8, size_t(n) * 8, size_t(n), s)) return cr;
    MRM_HIP_TRY(hipMemcpyAsync(indx, dI, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return rc;
}

namespace {

__global__ void identity_kernel(double* X, int64_t n) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx < n * n) X[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
}

}  // namespace

// MatrixInverseInPlace (src/LinAlgLU.pas:386-427)
int32_t mrm_inverse_inplace(double* A, int64_t LineWidthA, int64_t width) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthA))) return rc;
    if (width == 0) return MRM_OK;
    if (!A) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    const int64_t n = width;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dLU, *dE, *dX, *dI, *dP;
    if ((rc = mrm_staging(0, size_t(n * n) * 8, &dLU))) return rc;
    if ((rc = mrm_staging(1, size_t(n * n) * 8, &dE))) return rc;
    if ((rc = mrm_staging(2, size_t(n * n) * 8, &dX))) return rc;
    if ((rc = mrm_staging(5, size_t(n) * 4, &dI))) return rc;
    if ((rc = mrm_staging(6, size_t(n) * 4, &dP))) return rc;
    rc = factor_host_copy(static_cast<double*>(dLU), n, A, LineWidthA, static_cast<int*>(dI),
                          static_cast<int*>(dP), s);
    if (rc) return rc;  // leSingular leaves A unchanged, as the reference does
    hipLaunchKernelGGL(identity_kernel, dim3(unsigned((n * n + 255) / 256)), dim3(256), 0, s,
                       static_cast<double*>(dE), n);
    if ((rc = mrm::lu_solve<double>(static_cast<const double*>(dLU), n, int(n), static_cast<const int*>(dP),
                                    static_cast<const double*>(dE), n, static_cast<double*>(dX), n, int(n), s)))
        return rc;
    if ((rc = mrm_copy_d2h(A, size_t(LineWidthA), dX, size_t(n) * 8, size_t(n) * 8, size_t(n), s))) return rc;
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

// MatrixDeterminant (src/LinAlgLU.pas:429-478)
int32_t mrm_determinant(const double* A, int64_t LineWidthA, int64_t width, double* det) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthA))) return rc;
    if (!det || (width > 0 && !A)) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    if (width == 0) {
        *det = 1.0;
        return MRM_OK;
    }
    const int64_t n = width;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dLU, *dI;
    if ((rc = mrm_staging(0, size_t(n * n) * 8, &dLU))) return rc;
    if ((rc = mrm_staging(5, size_t(n) * 4, &dI))) return rc;
    rc = factor_host_copy(static_cast<double*>(dLU), n, A, LineWidthA, static_cast<int*>(dI), nullptr, s);
    if (rc == MRM_SINGULAR) {
        *det = 0.0;
        return MRM_SINGULAR;
    }
    if (rc) return rc;
    std::vector<double> diag(static_cast<size_t>(n));
    std::vector<int32_t> indx(static_cast<size_t>(n));
    // the diagonal is a strided column: one double every (n+1) elements
    if ((rc = mrm_copy_d2h(diag.data(), 8, dLU, size_t(n + 1) * 8, 8, size_t(n), s))) return rc;
    MRM_HIP_TRY(hipMemcpyAsync(indx.data(), dI, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    MRM_HIP_TRY(hipStreamSynchronize(s));
    double d = 1.0;
    for (int64_t i = 0; i < n; ++i)
        if (indx[size_t(i)] != i) d = -d;
    for (int64_t i = 0; i < n; ++i) d = d * diag[size_t(i)];
    *det = d;
    return MRM_OK;
}

// MatrixLUBackSubst (src/LinAlgLU.pas:315-385): one right-hand side, stride LineWidthB.
int32_t mrm_lu_backsubst(const double* LUDecomp, int64_t LineWidthLU, int64_t width, const int32_t* indx, double* B,
                         int64_t LineWidthB) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthLU))) return rc;
    if (width == 0) return MRM_OK;
    if (!LUDecomp || !indx || !B || LineWidthB < 8 || (LineWidthB & 7)) {
        mrm_set_error(MRM_BAD_ARG, "bad argument");
        return MRM_BAD_ARG;
    }
    const int64_t n = width;
    std::vector<int> perm = perm_from_indx(indx, n);
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dLU, *dP, *dB, *dX;
    if ((rc = mrm_staging(1, size_t(n * n) * 8, &dLU))) return rc;
    if ((rc = mrm_staging(5, size_t(n) * 4, &dP))) return rc;
    if ((rc = mrm_staging(2, size_t(n) * 8, &dB))) return rc;
    if ((rc = mrm_staging(3, size_t(n) * 8, &dX))) return rc;
    if ((rc = mrm_copy_h2d(dLU, size_t(n) * 8, LUDecomp, size_t(LineWidthLU), size_t(n) * 8, size_t(n), s))) return rc;
    MRM_HIP_TRY(hipMemcpyAsync(dP, perm.data(), size_t(n) * 4, hipMemcpyHostToDevice, s));
    if ((rc = mrm_copy_h2d(dB, 8, B, size_t(LineWidthB), 8, size_t(n), s))) return rc;
    rc = mrm::lu_solve<double>(static_cast<const double*>(dLU), n, int(n), static_cast<const int*>(dP),
                               static_cast<const double*>(dB), 1, static_cast<double*>(dX), 1, 1, s);
    if (rc) return rc;
    if ((rc = mrm_copy_d2h(B, size_t(LineWidthB), dX, 8, 8, size_t(n), s))) return rc;
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

// MatrixLinEQSolve (src/LinAlgLU.pas:481-592): X = A^-1 B for width2 right-hand
// sides, then NumRefinments FP64 refinement steps r = A x - b, x -= LU^-1 r
// (applied to every right-hand side; the reference's residual loop is correct
// only for one).
int32_t mrm_lineq_solve(const double* A, int64_t LineWidthA, int64_t width, const double* B, int64_t LineWidthB,
                        double* X, int64_t LineWidthX, int64_t width2, int64_t NumRefinments) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthA))) return rc;
    if (width2 < 0 || LineWidthB < width2 * 8 || LineWidthX < width2 * 8 || ((LineWidthB | LineWidthX) & 7) ||
        NumRefinments < 0) {
        mrm_set_error(MRM_BAD_DIMS, "Dimension error");
        return MRM_BAD_DIMS;
    }
    if (width == 0 || width2 == 0) return MRM_OK;
    if (!A || !B || !X) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    const int64_t n = width, m = width2;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dA, *dLU, *dB, *dX, *dR, *dD, *dI, *dP;
    if ((rc = mrm_staging(0, size_t(n * n) * 8, &dA))) return rc;
    if ((rc = mrm_staging(1, size_t(n * n) * 8, &dLU))) return rc;
    if ((rc = mrm_staging(2, size_t(n * m) * 8, &dB))) return rc;
    if ((rc = mrm_staging(3, size_t(n * m) * 8, &dX))) return rc;
    if ((rc = mrm_staging(4, size_t(n * m) * 8, &dR))) return rc;
    if ((rc = mrm_staging(5, size_t(n) * 4, &dI))) return rc;
    if ((rc = mrm_staging(6, size_t(n) * 4, &dP))) return rc;
    if ((rc = mrm_staging(7, size_t(n * m) * 8, &dD))) return rc;
    if ((rc = mrm_copy_h2d(dA, size_t(n) * 8, A, size_t(LineWidthA), size_t(n) * 8, size_t(n), s))) return rc;
    auto copy_a = [&]() -> int {
        MRM_HIP_TRY(hipMemcpyAsync(dLU, dA, size_t(n * n) * 8, hipMemcpyDeviceToDevice, s));
        return MRM_OK;
    };
    if ((rc = copy_a())) return rc;
    if ((rc = mrm_copy_h2d(dB, size_t(m) * 8, B, size_t(LineWidthB), size_t(m) * 8, size_t(n), s))) return rc;
    double* a = static_cast<double*>(dA);
    double* lu = static_cast<double*>(dLU);
    double* b = static_cast<double*>(dB);
    double* x = static_cast<double*>(dX);
    double* r = static_cast<double*>(dR);
    double* d = static_cast<double*>(dD);
    int* perm = static_cast<int*>(dP);
    rc = factor_and_status(lu, n, int(n), static_cast<int*>(dI), perm, s, copy_a);
    if (rc) return rc;
    // the diagonal-block inverses of the factors serve the solve and every
    // refinement step
    void* dinv = nullptr;
    const size_t inv_elems = mrm::lu_solve_inverse_elems(int(n));
    if (inv_elems) {
        MRM_HIP_TRY(hipMallocAsync(&dinv, inv_elems * 8, s));
        rc = mrm::lu_solve_inverses(lu, n, int(n), static_cast<double*>(dinv), s);
    }
    const double* inv = static_cast<const double*>(dinv);
    if (!rc) rc = mrm::lu_solve<double>(lu, n, int(n), perm, b, m, x, m, int(m), s, inv);
    for (int64_t it = 0; it < NumRefinments && !rc; ++it) {
        // r = b - A x  (the negated residual of the reference), d = LU^-1 r, x += d
        if (hipMemcpyAsync(r, b, size_t(n * m) * 8, hipMemcpyDeviceToDevice, s) != hipSuccess) {
            mrm_set_error(MRM_HIP_ERROR, "lineq solve: residual copy failed");
            rc = MRM_HIP_ERROR;
            break;
        }
        if ((rc = mrm_launch_dgemm(MRM_NN, n, m, n, a, n * 8, x, m * 8, r, m * 8, MRM_DO_SUB, s))) break;
        if ((rc = mrm::lu_solve<double>(lu, n, int(n), perm, r, m, d, m, int(m), s, inv))) break;
        mrm::add_inplace<double>(x, m, d, m, int(n), int(m), s);
    }
    if (dinv) {
        const hipError_t e = hipFreeAsync(dinv, s);
        if (!rc && e != hipSuccess) MRM_HIP_TRY(e);
    }
    if (rc) return rc;
    if ((rc = mrm_copy_d2h(X, size_t(LineWidthX), x, size_t(m) * 8, size_t(m) * 8, size_t(n), s))) return rc;
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

int64_t mrm_lowp_workspace_bytes(int64_t M, int64_t N, int64_t K, int32_t mode) {
    if (M < 0 || N < 0 || K < 0 ||
        (mode != MRM_SOLVE_FP32_IR && mode != MRM_SOLVE_BF16_IR && mode != MRM_SOLVE_BF16X3_IR))
        return 0;
    return int64_t(mrm_lowp_workspace(M, N, K, mode));
}

int32_t mrm_dev_gemm_lowp(int32_t mode, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                          const double* B, int64_t ldb, double* C, int64_t ldc, void* work, int64_t work_bytes,
                          void* stream) {
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if (M < 0 || N < 0 || K < 0 || ((lda | ldb | ldc) & 7) || (M > 0 && K > 0 && lda < K * 8) ||
        (K > 0 && N > 0 && ldb < N * 8) || (M > 0 && N > 0 && ldc < N * 8)) {
        mrm_set_error(MRM_BAD_DIMS, "lowp update: bad dimensions or strides");
        return MRM_BAD_DIMS;
    }
    if (M > 0 && N > 0 && K > 0 && (!A || !B || !C)) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    return mrm_launch_lowp_update(mode, M, N, K, A, lda / 8, B, ldb / 8, C, ldc / 8, work,
                                  work_bytes < 0 ? 0 : size_t(work_bytes), mrm_resolve_stream(stream));
}

int32_t mrm_dev_lineq_solve_ir(int32_t mode, const double* A, int64_t lda, int64_t n, const double* B, int64_t ldb,
                               double* X, int64_t ldx, int64_t nrhs, int32_t max_iter, double tol,
                               int32_t* iterations, double* backward_error, int32_t* fell_back, void* stream) {
    mrm_solve_stats st;
    const int32_t rc = mrm_dev_lineq_solve_ir_stats(mode, A, lda, n, B, ldb, X, ldx, nrhs, max_iter, tol, &st, stream);
    if (iterations) *iterations = st.iterations;
    if (backward_error) *backward_error = st.backward_error;
    if (fell_back) *fell_back = st.fell_back;
    return rc;
}

int32_t mrm_dev_lineq_solve_ir_stats(int32_t mode, const double* A, int64_t lda, int64_t n, const double* B,
                                     int64_t ldb, double* X, int64_t ldx, int64_t nrhs, int32_t max_iter, double tol,
                                     mrm_solve_stats* stats, void* stream) {
    if (stats) *stats = mrm_solve_stats{};
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(n, lda))) return rc;
    if (nrhs < 0 || nrhs > 0x7fffffffLL || ldb < nrhs * 8 || ldx < nrhs * 8 || ((ldb | ldx) & 7) || max_iter < 0 ||
        !(tol >= 0.0)) {
        mrm_set_error(MRM_BAD_DIMS, "solve: bad right-hand-side layout, max_iter or tol");
        return MRM_BAD_DIMS;
    }
    if (n > 0 && nrhs > 0 && (!A || !B || !X)) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    mrm::IrStats st;
    rc = mrm::lineq_solve_ir(mode, A, lda / 8, int(n), B, ldb / 8, X, ldx / 8, int(nrhs), max_iter, tol, &st,
                             mrm_resolve_stream(stream));
    if (stats) {
        stats->iterations = st.iterations;
        stats->fell_back = st.fell_back;
        stats->gmres = st.gmres;
        stats->krylov_steps = st.krylov_steps;
        stats->backward_error = st.backward_error;
        stats->lowp_iterations = st.lowp_iterations;
        stats->lowp_krylov_steps = st.lowp_krylov_steps;
        stats->lowp_backward_error = st.lowp_backward_error;
    }
    return rc;
}

int32_t mrm_lineq_solve_ex(const double* A, int64_t LineWidthA, int64_t width, const double* B, int64_t LineWidthB,
                           double* X, int64_t LineWidthX, int64_t width2, int64_t NumRefinments, int32_t mode) {
    if (mode == MRM_SOLVE_FP64)
        return mrm_lineq_solve(A, LineWidthA, width, B, LineWidthB, X, LineWidthX, width2, NumRefinments);
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((rc = check_square(width, LineWidthA))) return rc;
    if (width2 < 0 || LineWidthB < width2 * 8 || LineWidthX < width2 * 8 || ((LineWidthB | LineWidthX) & 7) ||
        NumRefinments < 0) {
        mrm_set_error(MRM_BAD_DIMS, "Dimension error");
        return MRM_BAD_DIMS;
    }
    if (mode != MRM_SOLVE_FP32_IR && mode != MRM_SOLVE_BF16_IR && mode != MRM_SOLVE_BF16X3_IR) {
        mrm_set_error(MRM_BAD_ARG, "solve: unknown mode %d", mode);
        return MRM_BAD_ARG;
    }
    if (width == 0 || width2 == 0) return MRM_OK;
    if (!A || !B || !X) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    const int64_t n = width, m = width2;
    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dA, *dB, *dX;
    if ((rc = mrm_staging(0, size_t(n * n) * 8, &dA))) return rc;
    if ((rc = mrm_staging(2, size_t(n * m) * 8, &dB))) return rc;
    if ((rc = mrm_staging(3, size_t(n * m) * 8, &dX))) return rc;
    if ((rc = mrm_copy_h2d(dA, size_t(n) * 8, A, size_t(LineWidthA), size_t(n) * 8, size_t(n), s))) return rc;
    if ((rc = mrm_copy_h2d(dB, size_t(m) * 8, B, size_t(LineWidthB), size_t(m) * 8, size_t(n), s))) return rc;
    mrm::IrStats st;
    rc = mrm::lineq_solve_ir(mode, static_cast<const double*>(dA), n, int(n), static_cast<const double*>(dB), m,
                             static_cast<double*>(dX), m, int(m), NumRefinments ? int(NumRefinments) : 50, 1e-14,
                             &st, s);
    if (rc != MRM_OK && rc != MRM_NOT_CONVERGED) return rc;
    // rc keeps the status of the solve
    if (int cr = mrm_copy_d2h(X, size_t(LineWidthX), dX, size_t(m) * 8, size_t(m) * 8, size_t(n), s)) return cr;
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return rc;
}

}  // extern "C"
