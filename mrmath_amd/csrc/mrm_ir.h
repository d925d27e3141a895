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
    int gmres = 0;                   // 1: the corrections came from GMRES (the plain step stalled)
    int krylov_steps = 0;            // Krylov steps over all GMRES corrections of the returned solution
    int lowp_krylov_steps = 0;       // the same for the low-precision attempt before a fall-back
};

// GMRES corrections of the refinement: Krylov dimension and relative
// tolerance on the preconditioned residual of one correction solve.
constexpr int IR_GMRES_DIM = 40;
constexpr double IR_GMRES_TOL = 1e-8;

// X (n x nrhs) = A^-1 B, all device memory, element strides.  The LU of a copy
// of A is computed in `mode` (MRM_SOLVE_*), then refined in FP64 until the
// normwise backward error is <= tol or max_iter steps.  When a plain step
// with low-precision factors gains less than tenfold, the remaining
// corrections come from GMRES preconditioned by the same factors (GMRES-IR);
// only if a GMRES step also gains less than tenfold is the factorisation
// redone in FP64.  Synchronises `s`.  Caller holds
// mrm_host_mutex().
int lineq_solve_ir(int mode, const double* A, int64_t lda, int n, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int nrhs, int max_iter, double tol, IrStats* stats, hipStream_t s);

}  // namespace mrm
