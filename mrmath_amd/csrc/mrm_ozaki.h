// mrm_ozaki.h -- FP64 GEMM emulated with INT8 or BF16 slices (internal launcher).
#pragma once

#include "mrm_common.h"

// At most this many 7-bit slices per operand (56 significant bits).
#define MRM_OZAKI_MAX_SLICES 8
// K is processed in chunks of at most this many; within a chunk the sum of the
// (d+1) slice products of one diagonal stays below 2^31:
// 8 * 16384 * 127^2 = 2113929216 < 2147483648.
#define MRM_OZAKI_KCHUNK 16384

int mrm_launch_dgemm_ozaki(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double* C, int64_t ldc, int op, int slices, int kind, void* work,
                           size_t work_bytes, hipStream_t stream);
size_t mrm_ozaki_workspace(int64_t M, int64_t N, int64_t K, int slices, int kind);
