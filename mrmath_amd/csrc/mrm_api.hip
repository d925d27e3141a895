// mrm_api.hip -- the C ABI of libmrmath_gpu.so: library state, error reporting,
// host-pointer entries with the reference's stub-switch signatures, and the
// device-pointer entries.
//
// Host-pointer GEMM entries take a context (a stream and grow-only staging
// buffers) from a pool for the duration of the call, so calls from several
// threads run concurrently: the reference's threaded callers invoke the
// blocked pointer from up to 64 pool threads on disjoint tiles
// (src/ThreadedMatrixOperations.pas:706-717).  The LU / solve entries share
// one stream and staging set behind a mutex.

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "mrm_common.h"
#include "mrm_host.h"
#include "mrm_lu.h"
#include "mrm_ozaki.h"

namespace {

thread_local int t_err = MRM_OK;
thread_local char t_msg[512] = "";

struct Context {
    std::mutex init_mu;
    std::atomic<bool> ready{false};
    bool atexit_registered = false;
    int device = 0;
    hipStream_t stream = nullptr;
};
Context g_ctx;

// Staging of the mutex-serialised entries (LU / solve / inverse / determinant).
void* g_bufs[MRM_NUM_SLOTS] = {};
size_t g_caps[MRM_NUM_SLOTS] = {};

// Per-call context of a host-pointer GEMM entry.
struct GemmCtx {
    hipStream_t stream = nullptr;
    void* bufs[4] = {};
    size_t caps[4] = {};
};
std::mutex g_pool_mu;
std::vector<GemmCtx*> g_pool;   // idle contexts
int g_pool_live = 0;            // contexts handed out

std::atomic<int64_t> g_calls{0}, g_h2d{0}, g_d2h{0}, g_gemms{0};

int grow_buffer(void*& buf, size_t& cap, size_t bytes, hipStream_t s) {
    if (bytes == 0) bytes = 256;
    if (cap >= bytes) return MRM_OK;
    if (buf) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(buf);
        buf = nullptr;
        cap = 0;
    }
    const size_t c = (bytes + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
    hipError_t e = hipMalloc(&buf, c);
    if (e != hipSuccess) {
        mrm_set_error(MRM_NO_MEMORY, "hipMalloc(%zu) failed: %s", c, hipGetErrorString(e));
        return MRM_NO_MEMORY;
    }
    cap = c;
    return MRM_OK;
}

void free_ctx(GemmCtx* c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 4; ++i)
        if (c->bufs[i]) (void)hipFree(c->bufs[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int acquire_ctx(GemmCtx** out) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (!g_pool.empty()) {
            *out = g_pool.back();
            g_pool.pop_back();
            ++g_pool_live;
            return MRM_OK;
        }
    }
    GemmCtx* c = new GemmCtx();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        mrm_set_error(MRM_HIP_ERROR, "hipStreamCreate failed");
        return MRM_HIP_ERROR;
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    ++g_pool_live;
    *out = c;
    return MRM_OK;
}

void release_ctx(GemmCtx* c) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(c);
    --g_pool_live;
}

// Releases every library-owned HIP object (see mrm_shutdown).
void pinned_release_all();  // the pinned copy buffers (below)

void release_all() {
    std::lock_guard<std::mutex> lk(g_ctx.init_mu);
    if (!g_ctx.ready.load()) return;
    (void)hipStreamSynchronize(g_ctx.stream);
    {
        std::lock_guard<std::mutex> pk(g_pool_mu);
        for (GemmCtx* c : g_pool) free_ctx(c);
        g_pool.clear();
    }
    {
        std::lock_guard<std::mutex> hk(mrm_host_mutex());
        for (int i = 0; i < MRM_NUM_SLOTS; ++i) {
            if (g_bufs[i]) (void)hipFree(g_bufs[i]);
            g_bufs[i] = nullptr;
            g_caps[i] = 0;
        }
        mrm::lu_release_workspace();
        mrm::lu_lookahead_release();  // the LU entries use its streams under this mutex
        mrm::ir_release_workspace();
        mrm::gmres_release_workspace();
    }
    pinned_release_all();
    (void)hipStreamDestroy(g_ctx.stream);
    g_ctx.stream = nullptr;
    // give the stream-ordered allocations of the split products back
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, g_ctx.device) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
    g_ctx.ready.store(false);
}

void release_at_exit() { release_all(); }

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

std::mutex& mrm_host_mutex() {
    static std::mutex mu;
    return mu;
}

int mrm_staging(int slot, size_t bytes, void** out) {
    if (slot < 0 || slot >= MRM_NUM_SLOTS) return MRM_BAD_ARG;
    int rc = grow_buffer(g_bufs[slot], g_caps[slot], bytes, g_ctx.stream);
    if (rc) return rc;
    *out = g_bufs[slot];
    return MRM_OK;
}

hipStream_t mrm_stream() { return g_ctx.stream; }

int mrm_ensure_init() {
    if (g_ctx.ready.load()) return MRM_OK;
    std::lock_guard<std::mutex> lk(g_ctx.init_mu);
    if (g_ctx.ready.load()) return MRM_OK;
    int dev = 0;
    MRM_HIP_TRY(hipGetDevice(&dev));
    MRM_HIP_TRY(hipSetDevice(dev));
    hipStream_t s;
    MRM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_ctx.device = dev;
    g_ctx.stream = s;
    // the LU lookahead's CU-masked streams, before any other work of the
    // process takes hardware queues (not available: the LU runs without)
    (void)mrm::lu_lookahead_init();
    if (!g_ctx.atexit_registered) {
        // registered after the HIP runtime initialised, so it runs before the
        // runtime's own teardown (and a profiler's finalisation)
        std::atexit(release_at_exit);
        g_ctx.atexit_registered = true;
    }
    g_ctx.ready.store(true);
    return MRM_OK;
}

// Device entries run on the caller's stream; NULL is HIP's null stream, so work
// the caller queued there (e.g. by torch on its default stream) is ordered first.
hipStream_t mrm_resolve_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Host <-> device copies of `rows` rows of `width` bytes (pitches in bytes).
// A 2-D copy with pageable host memory moves one row at a time (8 MiB in rows
// of 8 KiB: ~1 GB/s measured, against 18-34 GB/s for one 1-D copy), so rows
// that are not contiguous in host memory are gathered (scattered) by the CPU
// into (out of) a pinned buffer that moves as one 1-D copy; only the rows'
// own bytes are read or written, so callers that work on disjoint tiles of
// one matrix at the same time stay correct.  Rows not contiguous on the
// device are packed or unpacked there by a device-side 2-D copy.  A pitched
// host side makes the call wait for its copy.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_pin_mu;
std::vector<std::pair<void*, size_t>> g_pin_free;  // pinned host buffers not in use

void* pinned_take(size_t bytes, size_t* size) {
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (size_t i = 0; i < g_pin_free.size(); ++i)
            if (g_pin_free[i].second >= bytes) {
                void* p = g_pin_free[i].first;
                *size = g_pin_free[i].second;
                g_pin_free.erase(g_pin_free.begin() + std::ptrdiff_t(i));
                return p;
            }
    }
    void* p = nullptr;
    const size_t want = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *size = want;
    return p;
}

// Buffers up to 256 MiB go back to the pool (kept until mrm_shutdown); larger
// ones are freed at once, so one large padded copy does not keep its pinned
// memory.
void pinned_give(void* p, size_t size) {
    if (size > (size_t(256) << 20)) {
        (void)hipHostFree(p);
        return;
    }
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.emplace_back(p, size);
}

void pinned_release_all() {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    for (auto& b : g_pin_free) (void)hipHostFree(b.first);
    g_pin_free.clear();
}

int device_pack_buffer(void** tmp, size_t bytes, hipStream_t s) {
    if (hipMallocAsync(tmp, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        mrm_set_error(MRM_NO_MEMORY, "cannot allocate %zu bytes of copy staging", bytes);
        return MRM_NO_MEMORY;
    }
    return MRM_OK;
}
}  // namespace

int mrm_copy_h2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                 hipStream_t s) {
    if (width == 0 || rows == 0) return MRM_OK;
    const size_t bytes = width * rows;
    if (spitch == width && dpitch == width) {
        MRM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return MRM_OK;
    }
    // the packed rows on the device: dst itself, or a buffer unpacked into dst
    void* packed = dst;
    if (dpitch != width) {
        int rc = device_pack_buffer(&packed, bytes, s);
        if (rc) return rc;
    }
    hipError_t e = hipSuccess;
    void* pin = nullptr;
    size_t pin_size = 0;
    if (spitch == width) {
        e = hipMemcpyAsync(packed, src, bytes, hipMemcpyHostToDevice, s);
    } else if ((pin = pinned_take(bytes, &pin_size))) {
        for (size_t r = 0; r < rows; ++r)
            memcpy(static_cast<char*>(pin) + r * width, static_cast<const char*>(src) + r * spitch, width);
        e = hipMemcpyAsync(packed, pin, bytes, hipMemcpyHostToDevice, s);
    } else {
        e = hipMemcpy2DAsync(packed, width, src, spitch, width, rows, hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess && packed != dst)
        e = hipMemcpy2DAsync(dst, dpitch, packed, width, width, rows, hipMemcpyDeviceToDevice, s);
    if (packed != dst) {
        const hipError_t f = hipFreeAsync(packed, s);
        if (e == hipSuccess) e = f;
    }
    if (pin) {
        // the pinned buffer goes back to the pool only after its copy
        const hipError_t f = hipStreamSynchronize(s);
        if (e == hipSuccess) e = f;
        pinned_give(pin, pin_size);
    }
    MRM_HIP_TRY(e);
    return MRM_OK;
}

int mrm_copy_d2h(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                 hipStream_t s) {
    if (width == 0 || rows == 0) return MRM_OK;
    const size_t bytes = width * rows;
    if (spitch == width && dpitch == width) {
        MRM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        return MRM_OK;
    }
    const void* packed = src;
    void* tmp = nullptr;
    hipError_t e = hipSuccess;
    if (spitch != width) {
        int rc = device_pack_buffer(&tmp, bytes, s);
        if (rc) return rc;
        e = hipMemcpy2DAsync(tmp, width, src, spitch, width, rows, hipMemcpyDeviceToDevice, s);
        packed = tmp;
    }
    void* pin = nullptr;
    size_t pin_size = 0;
    if (e == hipSuccess) {
        if (dpitch == width) {
            e = hipMemcpyAsync(dst, packed, bytes, hipMemcpyDeviceToHost, s);
        } else if ((pin = pinned_take(bytes, &pin_size))) {
            e = hipMemcpyAsync(pin, packed, bytes, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e == hipSuccess)
                for (size_t r = 0; r < rows; ++r)
                    memcpy(static_cast<char*>(dst) + r * dpitch, static_cast<const char*>(pin) + r * width, width);
            pinned_give(pin, pin_size);
        } else {
            e = hipMemcpy2DAsync(dst, dpitch, packed, width, width, rows, hipMemcpyDeviceToHost, s);
        }
    }
    if (tmp) {
        const hipError_t f = hipFreeAsync(tmp, s);
        if (e == hipSuccess) e = f;
    }
    MRM_HIP_TRY(e);
    return MRM_OK;
}

// ---------------------------------------------------------------------------
// Host-side GEMM staging.
// ---------------------------------------------------------------------------
static int copy_in(void* dst, const void* src, int64_t lw, int64_t w, int64_t h, hipStream_t s) {
    int rc = mrm_copy_h2d(dst, size_t(w) * 8, src, size_t(lw), size_t(w) * 8, size_t(h), s);
    if (rc) return rc;
    g_h2d.fetch_add(int64_t(w) * h * 8);
    return MRM_OK;
}

// dims: A is (rowsA x colsA) with LW1, B is (rowsB x colsB) with LW2, C is M x N.
// slices == 0: native FP64 GEMM; 1..8: INT8-slice emulation (workspace buffer 3)
static int host_gemm_on(GemmCtx& cx, int trans, double* dest, int64_t destLW, const double* mt1, const double* mt2,
                        int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2, int64_t M,
                        int64_t N, int64_t K, int op, int slices) {
    hipStream_t s = cx.stream;
    int rc;
    // one operand used twice (X^T X, X X^T) is staged once, which also lets
    // the device product take its symmetric path
    const bool same = mt1 == mt2 && LW1 == LW2 && w1 == w2 && h1 == h2;
    if ((rc = grow_buffer(cx.bufs[0], cx.caps[0], size_t(w1 * h1) * 8, s))) return rc;
    if (!same && (rc = grow_buffer(cx.bufs[1], cx.caps[1], size_t(w2 * h2) * 8, s))) return rc;
    if ((rc = grow_buffer(cx.bufs[2], cx.caps[2], size_t(M * N) * 8, s))) return rc;
    void* dA = cx.bufs[0];
    void* dB = same ? dA : cx.bufs[1];
    void* dC = cx.bufs[2];
    if ((rc = copy_in(dA, mt1, LW1, w1, h1, s))) return rc;
    if (!same && (rc = copy_in(dB, mt2, LW2, w2, h2, s))) return rc;
    if (op != MRM_DO_NONE && (rc = copy_in(dC, dest, destLW, N, M, s))) return rc;
    if (slices == 0) {
        rc = mrm_launch_dgemm(trans, M, N, K, static_cast<const double*>(dA), w1 * 8, static_cast<const double*>(dB),
                              w2 * 8, static_cast<double*>(dC), N * 8, op, s);
    } else {
        const size_t wb = mrm_ozaki_workspace(M, N, K, slices, MRM_OZAKI_INT8);
        if ((rc = grow_buffer(cx.bufs[3], cx.caps[3], wb, s))) return rc;
        rc = mrm_launch_dgemm_ozaki(trans, M, N, K, static_cast<const double*>(dA), w1 * 8,
                                    static_cast<const double*>(dB), w2 * 8, static_cast<double*>(dC), N * 8, op, slices,
                                    MRM_OZAKI_INT8, cx.bufs[3], wb, s);
    }
    if (rc) return rc;
    g_gemms.fetch_add(1);
    if ((rc = mrm_copy_d2h(dest, size_t(destLW), dC, size_t(N) * 8, size_t(N) * 8, size_t(M), s))) return rc;
    g_d2h.fetch_add(M * N * 8);
    MRM_HIP_TRY(hipStreamSynchronize(s));
    return MRM_OK;
}

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
    g_calls.fetch_add(1);
    GemmCtx* cx;
    if ((rc = acquire_ctx(&cx))) return rc;
    rc = host_gemm_on(*cx, trans, dest, destLW, mt1, mt2, w1, h1, w2, h2, LW1, LW2, M, N, K, op, slices);
    if (rc) (void)hipStreamSynchronize(cx->stream);  // nothing of this call may still run on a reused context
    release_ctx(cx);
    return rc;
}

extern "C" {

int32_t mrm_init(void) {
    clear_error();
    return mrm_ensure_init();
}

void mrm_shutdown(void) { release_all(); }

int32_t mrm_last_error(void) { return t_err; }
const char* mrm_last_error_message(void) { return t_msg; }
int32_t mrm_version(void) { return 0x000200; }

void mrm_get_stats(mrm_stats* out) {
    if (!out) return;
    out->host_calls = g_calls.load();
    out->h2d_bytes = g_h2d.load();
    out->d2h_bytes = g_d2h.load();
    out->device_gemms = g_gemms.load();
}

void mrm_reset_stats(void) {
    g_calls.store(0);
    g_h2d.store(0);
    g_d2h.store(0);
    g_gemms.store(0);
}

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

// ThrMatrixMult / ThrMatrixMultT1 / ThrMatrixMultT2
// (src/ThreadedMatrixOperations.pas:24, 34, 40): the whole product as one call.
void mrm_thr_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                  int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2, int32_t op) {
    host_gemm(MRM_NN, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2, op);
}

void mrm_thr_mult_t1(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                     int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2,
                     int32_t op) {
    host_gemm(MRM_TN, dest, destLineWidth, mt1, mt2, width1, height1, width2, height2, LineWidth1, LineWidth2, op);
}

void mrm_thr_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                     int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2,
                     int32_t op) {
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

int64_t mrm_ozaki_workspace_bytes_ex(int64_t M, int64_t N, int64_t K, int32_t slices, int32_t kind) {
    if (M < 0 || N < 0 || K < 0 || slices < 1 || slices > MRM_OZAKI_MAX_SLICES) return 0;
    if (kind != MRM_OZAKI_INT8 && kind != MRM_OZAKI_BF16) return 0;
    return int64_t(mrm_ozaki_workspace(M, N, K, slices, kind));
}

int64_t mrm_ozaki_workspace_bytes(int64_t M, int64_t N, int64_t K, int32_t slices) {
    return mrm_ozaki_workspace_bytes_ex(M, N, K, slices, MRM_OZAKI_INT8);
}

int32_t mrm_dev_dgemm_ozaki_ex(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                               const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, int32_t slices,
                               int32_t kind, void* work, int64_t work_bytes, void* stream) {
    clear_error();
    int rc = mrm_ensure_init();
    if (rc) return rc;
    if ((M > 0 && N > 0) && (!C || (K > 0 && (!A || !B)))) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    if (work_bytes < 0) work_bytes = 0;
    return mrm_launch_dgemm_ozaki(trans, M, N, K, A, lda, B, ldb, C, ldc, op, slices, kind, work,
                                  size_t(work_bytes), mrm_resolve_stream(stream));
}

int32_t mrm_dev_dgemm_ozaki(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, int32_t slices,
                            void* work, int64_t work_bytes, void* stream) {
    return mrm_dev_dgemm_ozaki_ex(trans, M, N, K, A, lda, B, ldb, C, ldc, op, slices, MRM_OZAKI_INT8, work,
                                  work_bytes, stream);
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
