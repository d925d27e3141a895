// mrm_common.h -- shared definitions for the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double mrm_d4 __attribute__((ext_vector_type(4)));
typedef float mrm_f4 __attribute__((ext_vector_type(4)));

// Status codes, MRM_NN/TN/NT and MRM_DO_* come from the public header.
#include "mrmath_gpu.h"

// Internal launchers (all device pointers, byte strides, enqueue on `stream`).
int mrm_launch_dgemm(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double* C, int64_t ldc, int op, hipStream_t stream);
// The same with the K dimension cut into `splits` chunks (partial products
// reduced in chunk order): more workgroups for short, skinny products.  work
// (optional, splits * M * N doubles) holds the partial products; without it
// they come from a stream-ordered allocation.
int mrm_launch_dgemm_splitk(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, int op, int splits,
                            double* work, hipStream_t stream);

// The same with the shape's own K split and the partial products in the
// caller's work: never allocates; a split that needs more than work_bytes
// (or any split when work is NULL) is refused with MRM_BAD_DIMS.
// `batch` products C_b (op)= op(A_b) op(B_b), X_b = X + b * sx (element strides
// between the products; line widths in bytes), one launch, no K split.
int mrm_launch_dgemm_batched(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, int64_t sa,
                             const double* B, int64_t ldb, int64_t sb, double* C, int64_t ldc, int64_t sc, int batch,
                             int op, hipStream_t stream);

int mrm_launch_dgemm_ws(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double* C, int64_t ldc, int op, void* work, size_t work_bytes,
                        hipStream_t stream);
// Bytes of partial products mrm_launch_dgemm_ws needs for this shape (0: no
// split); same_operand: A and B are one matrix with one stride (A^T A, A A^T).
size_t mrm_dgemm_work_bytes(int trans, int64_t M, int64_t N, int64_t K, bool same_operand);

void mrm_set_error(int code, const char* fmt, ...);

#define MRM_HIP_TRY(expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            mrm_set_error(MRM_HIP_ERROR, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                             \
            return MRM_HIP_ERROR;                                                          \
        }                                                                                  \
    } while (0)
