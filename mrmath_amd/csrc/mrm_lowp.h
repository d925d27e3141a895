// mrm_lowp.h -- low-precision (FP32 / BF16 / split BF16) trailing update of the mixed-precision LU.
#pragma once

#include "mrm_common.h"

// Product precisions; equal to the public MRM_SOLVE_* modes.
#define MRM_LOWP_FP32 MRM_SOLVE_FP32_IR
#define MRM_LOWP_BF16 MRM_SOLVE_BF16_IR
#define MRM_LOWP_BF16X3 MRM_SOLVE_BF16X3_IR

// Device bytes of packed operands for an M x N x K update.
size_t mrm_lowp_workspace(int64_t M, int64_t N, int64_t K, int mode);

// C (M x N) -= A (M x K) * B (K x N), FP64 storage, element strides, the
// product rounded to `mode` and accumulated in FP32.
int mrm_launch_lowp_update(int mode, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double* C, int64_t ldc, void* work, size_t work_bytes, hipStream_t stream);
