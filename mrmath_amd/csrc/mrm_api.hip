// mrm_api.hip -- the C ABI of libmrmath_gpu.so: library state, error reporting,
// host-pointer entries with the reference's stub-switch signatures, and the
// device-pointer entries.
//
// Host-pointer entries stage operands through grow-only device buffers owned by
// the library.  One mutex serialises them, which makes the entries reentrant for
// the reference's threaded callers (ThrMatrixMultEx2 calls the blocked pointer
// from up to 64 pool threads, src/ThreadedMatrixOperations.pas:706-717).

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "mrm_common.h"
#include "mrm_host.h"
#include "mrm_ozaki.h"

namespace {

thread_local int t_err = MRM_OK;
thread_local char t_msg[512] = "";

struct Context {
    std::mutex init_mu;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
};
Context g_ctx;

}  // namespace

void mrm_set_error(int code, const char* fmt, ...) {
    t_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
}

static void clear_error() {
    t_err = MRM_OK;
    t_msg[0] = 0;
}

// ---------------------------------------------------------------------------
// Staging buffers for host-pointer entries.
// ---------------------------------------------------------------------------
std::mutex& mrm_host_mutex() {
    static std::mutex mu;
    return mu;
}

int mrm_staging(int slot, size_t bytes, void** out) {
    static void* bufs[MRM_NUM_SLOTS] = {};
    static size_t caps[MRM_NUM_SLOTS] = {};
    if (slot < 0 || slot >= MRM_NUM_SLOTS) return MRM_BAD_ARG;
    if (bytes == 0) bytes = 256;
    if (caps[slot] < bytes) {
        if (bufs[slot]) {
            (void)hipStreamSynchronize(g_ctx.stream);
            (void)hipFree(bufs[slot]);
            bufs[slot] = nullptr;
            caps[slot] = 0;
        }
        size_t cap = (bytes + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
        hipError_t e = hipMalloc(&bufs[slot], cap);
        if (e != hipSuccess) {
            mrm_set_error(MRM_NO_MEMORY, "hipMalloc(%zu) failed: %s", cap, hipGetErrorString(e));
            return MRM_NO_MEMORY;
        }
        caps[slot] = cap;
    }
    *out = bufs[slot];
    return MRM_OK;
}

hipStream_t mrm_stream() { return g_ctx.stream; }

int mrm_ensure_init() {
    if (g_ctx.ready) return MRM_OK;
    std::lock_guard<std::mutex> lk(g_ctx.init_mu);
    if (g_ctx.ready) return MRM_OK;
    int dev = 0;
    MRM_HIP_TRY(hipGetDevice(&dev));
    MRM_HIP_TRY(hipSetDevice(dev));
    hipStream_t s;
    MRM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_ctx.device = dev;
    g_ctx.stream = s;
    g_ctx.ready = true;
    return MRM_OK;
}

// Device entries run on the caller's stream; NULL is HIP's null stream, so work
// the caller queued there (e.g. by torch on its default stream) is ordered first.
hipStream_t mrm_resolve_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Host-side GEMM staging.
// ---------------------------------------------------------------------------
static int copy_in(void* dst, const void* src, int64_t lw, int64_t w, int64_t h, hipStream_t s) {
    MRM_HIP_TRY(hipMemcpy2DAsync(dst, size_t(w) * 8, src, size_t(lw), size_t(w) * 8, size_t(h),
                                 hipMemcpyHostToDevice, s));
    return MRM_OK;
}

// dims: A is (rowsA x colsA) with LW1, B is (rowsB x colsB) with LW2, C is M x N.
// slices == 0: native FP64 GEMM; 1..8: INT8-slice emulation (workspace in slot 3)
static int host_gemm(int trans, double* dest, int64_t destLW, const double* mt1, const double* mt2, int64_t w1,
                     int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2, int op, int slices = 0) {
    clear_error();
    if (w1 == 0 || h1 == 0 || w2 == 0 || h2 == 0) return MRM_OK;  // src/BlockedMult.pas:374-375
    if (w1 < 0 || h1 < 0 || w2 < 0 || h2 < 0 || op < 0 || op > 2) {
        mrm_set_error(MRM_BAD_ARG, "negative dimension or bad op");
        return MRM_BAD_ARG;
    }
    int64_t M, N, K;
    if (trans == MRM_NN) {
        if (w1 != h2) { mrm_set_error(MRM_BAD_DIMS, "Dimension error: width1 != height2"); return MRM_BAD_DIMS; }
        M = h1; N = w2; K = w1;
    } else if (trans == MRM_TN) {
        if (h1 != h2) { mrm_set_error(MRM_BAD_DIMS, "Dimension error: height1 != height2"); return MRM_BAD_DIMS; }
        M = w1; N = w2; K = h1;
    } else {
        if (w1 != w2) { mrm_set_error(MRM_BAD_DIMS, "Dimension error: width1 != width2"); return MRM_BAD_DIMS; }
        M = h1; N = h2; K = w1;
    }
    if (LW1 < w1 * 8 || LW2 < w2 * 8 || destLW < N * 8 || (LW1 | LW2 | destLW) & 7) {
        mrm_set_error(MRM_BAD_DIMS, "Line widths do not match");
        return MRM_BAD_DIMS;
    }
    if (!dest || !mt1 || !mt2) { mrm_set_error(MRM_BAD_ARG, "null pointer"); return MRM_BAD_ARG; }
    if (slices < 0 || slices > MRM_OZAKI_MAX_SLICES) {
        mrm_set_error(MRM_BAD_ARG, "slice count %d outside 1..%d", slices, MRM_OZAKI_MAX_SLICES);
        return MRM_BAD_ARG;
    }
    int rc = mrm_ensure_init();
    if (rc) return rc;

    std::lock_guard<std::mutex> lk(mrm_host_mutex());
    hipStream_t s = mrm_stream();
    void *dA, *dB, *dC;
    // one operand used twice (X^T X, X X^T) is staged once, which also lets
    // the device product take its symmetric path
    const bool same = mt1 == mt2 && LW1 == LW2 && w1 == w2 && h1 == h2;
    if ((rc = mrm_staging(0, size_t(w1 * h1) * 8, &dA))) return rc;
    if (!same && (rc = mrm_staging(1, size_t(w2 * h2) * 8, &dB))) return rc;
    if ((rc = mrm_staging(2, size_t(M * N) * 8, &dC))) return rc;
    if ((rc = copy_in(dA, mt1, LW1, w1, h1, s))) return rc;
    if (same) dB = dA;
    else if ((rc = copy_in(dB, mt2, LW2, w2, h2, s))) return rc;
    if (op != MRM_DO_NONE && (rc = copy_in(dC, dest, destLW, N, M, s))) return rc;
    if (slices == 0) {
        rc = mrm_launch_dgemm(trans, M, N, K, static_cast<const double*>(dA), w1 * 8, static_cast<const double*>(dB),
                              w2 * 8, static_cast<double*>(dC), N * 8, op, s);
    } else {
        const size_t wb = mrm_ozaki_workspace(M, N, K, slices);
        void* dW;
        if ((rc = mrm_staging(3, wb, &dW))) return rc;
        rc = mrm_launch_dgemm_ozaki(trans, M, N, K, static_cast<const double*>(dA), w1 * 8,
                                    static_cast<const double*>(dB), w2 * 8, static_cast<double*>(dC), N * 8, op, slices,
                                    dW, wb, s);
    }
    if (rc) return rc;
    MRM_HIP_TRY(hipMemcpy2DAsync(dest, size_t(destLW), dC, size_t(N) * 8, size_t(N) * 8, size_t(M),
                                 hipMemcpyDeviceToHost, s));
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

extern "C" {

int32_t mrm_init(void) {
    clear_error();
    return mrm_ensure_init();
}

void mrm_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_ctx.init_mu);
    if (!g_ctx.ready) return;
    (void)hipStreamSynchronize(g_ctx.stream);
    // Staging buffers and the stream live until process exit: other threads may
    // still hold results in flight, and the reference never tears down its tables.
}

int32_t mrm_last_error(void) { return t_err; }
const char* mrm_last_error_message(void) { return t_msg; }
int32_t mrm_version(void) { return 0x000100; }

void mrm_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
              int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2) {
    host_gemm(MRM_NN, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2,
              MRM_DO_NONE);
}

void mrm_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                 int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2) {
    host_gemm(MRM_NT, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2,
              MRM_DO_NONE);
}

void mrm_blocked_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                      int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2,
                      int64_t blockSize, int32_t op, double* mem) {
    (void)blockSize;
    (void)mem;
    host_gemm(MRM_NN, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2, op);
}

void mrm_blocked_mult_t1(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                         int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                         int64_t LineWidth2, int64_t blockSize, int32_t op, double* mem) {
    (void)blockSize;
    (void)mem;
    host_gemm(MRM_TN, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2, op);
}

void mrm_blocked_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                         int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                         int64_t LineWidth2, int64_t blockSize, int32_t op, double* mem) {
    (void)blockSize;
    (void)mem;
    host_gemm(MRM_NT, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2, op);
}

int32_t mrm_dev_dgemm(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, void* stream) {
    clear_error();
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((M > 0 && N > 0) && (!C || (K > 0 && (!A || !B)))) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    return mrm_launch_dgemm(trans, M, N, K, A, lda, B, ldb, C, ldc, op, mrm_resolve_stream(stream));
}

int64_t mrm_ozaki_workspace_bytes(int64_t M, int64_t N, int64_t K, int32_t slices) {
    if (M < 0 || N < 0 || K < 0 || slices < 1 || slices > MRM_OZAKI_MAX_SLICES) return 0;
    return int64_t(mrm_ozaki_workspace(M, N, K, slices));
}

int32_t mrm_dev_dgemm_ozaki(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, int32_t slices,
                            void* work, int64_t work_bytes, void* stream) {
    clear_error();
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((M > 0 && N > 0) && (!C || (K > 0 && (!A || !B)))) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    if (work_bytes < 0) work_bytes = 0;
    return mrm_launch_dgemm_ozaki(trans, M, N, K, A, lda, B, ldb, C, ldc, op, slices, work, size_t(work_bytes),
                                  mrm_resolve_stream(stream));
}

int32_t mrm_mult_ozaki(int32_t trans, double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                       int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                       int64_t LineWidth2, int32_t op, int32_t slices) {
    if (trans < 0 || trans > 2) {
        clear_error();
        mrm_set_error(MRM_BAD_ARG, "bad trans");
        return MRM_BAD_ARG;
    }
    if (slices < 1) {
        clear_error();
        mrm_set_error(MRM_BAD_ARG, "slice count %d outside 1..%d", slices, MRM_OZAKI_MAX_SLICES);
        return MRM_BAD_ARG;
    }
    return host_gemm(trans, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2,
                     op, slices);
}

}  // extern "C"
