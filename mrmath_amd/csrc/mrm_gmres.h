// mrm_gmres.h -- GMRES correction solves preconditioned by LU factors.
#pragma once

#include "mrm_common.h"

// Largest Krylov dimension of one correction solve.
#define MRM_GMRES_MAX_DIM 64

namespace mrm {

// D (n x nrhs, row stride nrhs) ~= A^-1 R by left-preconditioned GMRES on
// (LU)^-1 P A d = (LU)^-1 P r, one Krylov space per column, at most m steps;
// a column stops once its preconditioned residual is <= tol times its initial
// one.  LU / perm / dinv: the factors as lu_factor and lu_solve_inverses
// leave them (lda = n).  *steps: Krylov steps taken.  Synchronises s.  Caller
// holds mrm_host_mutex().
int gmres_correction(const double* A, int64_t lda, int n, const double* LU, const int* perm, const double* dinv,
                     const double* R, int64_t ldr, double* D, int nrhs, int m, double tol, int* steps, hipStream_t s);

void gmres_release_workspace();

}  // namespace mrm
