// mrm_noalloc.hip -- the allocation-free form of the device GEMM entry.
//
// mrm_dev_dgemm takes the partial products of a split-K product (the
// covariance shape X^T X: few output tiles, long K) from a stream-ordered
// allocation.  mrm_dev_dgemm_noalloc runs the same kernels with the same split
// but keeps the partial products in `scratch`, a file-scope device array of
// MRM_NOALLOC_SCRATCH_BYTES, so it calls no allocator at all: for callers
// that cannot (code under stream capture, serving loops, harnesses that
// forbid hipMalloc).
//
// Limit: the split of any shape needs at most 128 MiB (at most one wave of
// 512 workgroups, each chunk a full M x N slice: tm*tn <= 512/splits tiles of
// 128 x 128 doubles, twice that for the symmetric upper-tile grid), so every
// product fits; the entry still checks and rejects a larger need.  Every
// partial product is written in full before it is read, so nothing needs
// clearing between calls.  All calls on one GPU share its array: calls that
// overlap in time on one GPU (two streams, two threads) race.

#include <mutex>

#include "mrm_common.h"

namespace {

__device__ __align__(256) unsigned char scratch[MRM_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// The device address of this GPU's scratch (the array's host-side name is a
// stand-in, not a device address).  Looked up once per device, with that
// device current.
int scratch_for_current_device(void** out) {
    int dev = 0;
    MRM_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        mrm_set_error(MRM_BAD_ARG, "dgemm_noalloc: device %d beyond the %d the library tracks", dev, MAX_DEVICES);
        return MRM_BAD_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mrm_set_error(MRM_HIP_ERROR, "dgemm_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                          hipGetErrorString(e));
            return MRM_HIP_ERROR;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MRM_OK;
}

}  // namespace

extern "C" {

int32_t mrm_dev_dgemm_noalloc(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                              const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, void* stream) {
    if ((M > 0 && N > 0) && (!C || (K > 0 && (!A || !B)))) {
        mrm_set_error(MRM_BAD_ARG, "null pointer");
        return MRM_BAD_ARG;
    }
    if (trans < 0 || trans > 2 || M < 0 || N < 0 || K < 0) {
        mrm_set_error(MRM_BAD_ARG, "dgemm_noalloc: bad argument");
        return MRM_BAD_ARG;
    }
    const size_t need = mrm_dgemm_work_bytes(trans, M, N, K, A == B && lda == ldb);
    if (need > size_t(MRM_NOALLOC_SCRATCH_BYTES)) {
        mrm_set_error(MRM_BAD_DIMS, "dgemm_noalloc: the split product needs %zu bytes of scratch, the limit is %lld",
                      need, (long long)MRM_NOALLOC_SCRATCH_BYTES);
        return MRM_BAD_DIMS;
    }
    void* ptr = nullptr;
    int rc = scratch_for_current_device(&ptr);
    if (rc) return rc;
    return mrm_launch_dgemm_ws(trans, M, N, K, A, lda, B, ldb, C, ldc, op, ptr, size_t(MRM_NOALLOC_SCRATCH_BYTES),
                               reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
