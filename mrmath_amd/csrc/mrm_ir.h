// mrm_ir.h -- linear solve with low-precision LU and FP64 iterative refinement.
#pragma once

#include "mrm_common.h"

namespace mrm {

struct IrStats {
    int iterations = 0;              // refinement steps of the returned solution
    double backward_error = 0.0;     // its normwise backward error
    int fell_back = 0;               // 1: the low-precision factors stalled, FP64 was used
    int lowp_iterations = 0;         // steps taken with the low-precision factors before that
    double lowp_backward_error = 0.0;
};

// X (n x nrhs) = A^-1 B, all device memory, element strides.  The LU of a copy
// of A is computed in `mode` (MRM_SOLVE_*), then refined in FP64 until the
// normwise backward error is <= tol or max_iter steps; a stalled low-precision
// refinement falls back to mode FP64.  Synchronises `s`.  Caller holds
// mrm_host_mutex().
int lineq_solve_ir(int mode, const double* A, int64_t lda, int n, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int nrhs, int max_iter, double tol, IrStats* stats, hipStream_t s);

}  // namespace mrm
