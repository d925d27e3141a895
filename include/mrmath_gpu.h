/*
 * mrmath_gpu.h -- C ABI of libmrmath_gpu.so, the gfx950 (MI355X) backend for
 * mrMath's dense FP64 multiply and LU/solve path.
 *
 * Conventions are the reference's (src/MatrixConst.pas, src/Matrix.pas):
 * matrices are row-major, "width" is the number of columns, "height" the
 * number of rows, every LineWidth is the row stride in BYTES (a multiple of 8),
 * and sizes are 64-bit (NativeInt).  Enumerations are passed as int32.
 *
 * Two families of entry points:
 *   - host-pointer entries with exactly the reference's function-pointer
 *     signatures, for installation into the stub switch (they stage the
 *     operands through device memory and return when the result is back in
 *     host memory);
 *   - device-pointer entries (mrm_dev_*) that enqueue on a caller-supplied HIP
 *     stream (a hipStream_t passed as void*; NULL = HIP's null stream) and
 *     return at once.
 *
 * Errors: the reference's GEMM procedures return nothing, so the GEMM entries
 * record failures in a thread-local status read with mrm_last_error() /
 * mrm_last_error_message().  LU / solve entries return the status directly.
 * Every entry is safe to call from several host threads at once; the
 * host-pointer GEMM entries of different threads run concurrently (each
 * call takes a stream and staging buffers of its own from a pool), the LU /
 * solve entries one at a time.
 */
#ifndef MRMATH_GPU_H
#define MRMATH_GPU_H

#include <stdint.h>

#ifndef MRM_API
#define MRM_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. MRM_SINGULAR equals the reference's leSingular. */
#define MRM_OK 0
#define MRM_SINGULAR 1
#define MRM_BAD_DIMS 2
#define MRM_BAD_ARG 3
#define MRM_HIP_ERROR 4
#define MRM_NO_MEMORY 5
#define MRM_NOT_INITIALISED 6
#define MRM_NOT_CONVERGED 7
/* The LU panel's workgroups wait on each other's exchange words; a wait that
 * passed its deadline (they were not resident together) also on the
 * cooperative relaunch the library makes after a first timeout. */
#define MRM_TIMED_OUT 8

/* TMatrixMultDestOperation (reference src/MatrixConst.pas:147). */
#define MRM_DO_NONE 0
#define MRM_DO_ADD 1
#define MRM_DO_SUB 2

/* Operand layouts for the device entries. */
#define MRM_NN 0 /* C = A * B   */
#define MRM_TN 1 /* C = A^T * B (MultT1) */
#define MRM_NT 2 /* C = A * B^T (MultT2) */

/* Factorisation precisions of the linear solvers: plain FP64, or the trailing
 * updates on the FP32 / BF16 matrix cores with FP64 iterative refinement.
 * MRM_SOLVE_BF16X3_IR splits each operand into two BF16 pieces, x = xh + xl,
 * and forms Ah Bh + Ah Bl + Al Bh on the BF16 matrix cores (three products,
 * error about 2^-16 instead of 2^-8 of one BF16 product). */
#define MRM_SOLVE_FP64 0
#define MRM_SOLVE_FP32_IR 1
#define MRM_SOLVE_BF16_IR 2
#define MRM_SOLVE_BF16X3_IR 3

/* ---- library state ------------------------------------------------------ */

/* Selects the current HIP device, creates the library stream.  Called
 * implicitly by the first entry; replaces nothing in the reference (its
 * kernels are bound at unit initialisation, src/MatrixASMStubSwitch.pas:2080-2084). */
MRM_API int32_t mrm_init(void);
/* Waits for the library's work and releases its streams and device buffers
 * (also run at process exit).  No other thread may be inside the library;
 * the next entry initialises it again.  Batched plans stay valid until
 * mrm_batched_destroy. */
MRM_API void mrm_shutdown(void);
MRM_API int32_t mrm_last_error(void);
MRM_API const char* mrm_last_error_message(void);
/* Library version, e.g. 0x000200 = 0.2.0. */
MRM_API int32_t mrm_version(void);

/* Counters of the host-pointer GEMM entries since the last reset: calls,
 * bytes copied host -> device and device -> host, device products run.  No
 * reference counterpart (diagnostics). */
typedef struct mrm_stats {
    int64_t host_calls;
    int64_t h2d_bytes;
    int64_t d2h_bytes;
    int64_t device_gemms;
} mrm_stats;
MRM_API void mrm_get_stats(mrm_stats* out);
MRM_API void mrm_reset_stats(void);

/* Test hook, no reference counterpart: sets the deadline of the LU panel's
 * waits on a plain (non-cooperative) launch in 100 MHz wall-clock ticks;
 * 0 makes every such wait give up at once, so the factorisation reports a
 * timeout and is relaunched cooperatively from the original matrix (whose
 * waits keep the default 2 s); ticks < 0 restores the default.  Returns the
 * number of such relaunches so far in the process. */
MRM_API int64_t mrm_lu_panel_test_deadline(int64_t ticks);

/* Query, no reference counterpart: 1 when the LU factorisation overlaps the
 * factorisation of each outer block with the previous block's trailing
 * update on two CU-masked streams of the library (orders 6144 to 32768 on a
 * device of 256 CUs in 8 XCDs), 0 when it runs one kernel after the other.
 * Initialises the library. */
MRM_API int32_t mrm_lu_lookahead_available(void);

/* ---- host-pointer GEMM entries (stub-switch surface) --------------------- */

/* TMatrixMultFunc (src/MatrixASMStubSwitch.pas:351), the type of multFunc:
 * dest(height1 x width2) = mt1(height1 x width1) * mt2(height2 x width2),
 * width1 = height2.  Replaces FMAMatrixMult (src/FMAMatrixOperations.pas:94-143). */
MRM_API void mrm_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
              int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2);

/* TMatrixMultFunc, the type of multT2Func: dest(height1 x height2) = mt1 * mt2^T,
 * width1 = width2.  Replaces FMAMatrixMultTransposed (src/FMAMatrixOperations.pas:55-92). */
MRM_API void mrm_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                 int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2);

/* TMatrixBlockedMultfunc (src/MatrixASMStubSwitch.pas:352), blockedMultFunc:
 * dest (op)= mt1 * mt2.  Replaces BlockMatrixMultiplication (src/BlockedMult.pas:350-474).
 * blockSize and mem are accepted for ABI compatibility and ignored. */
MRM_API void mrm_blocked_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                      int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2,
                      int64_t blockSize, int32_t op, double* mem);

/* blockedMultT1Func: dest(width1 x width2) (op)= mt1^T * mt2, height1 = height2.
 * Replaces BlockMatrixMultiplicationT1 (src/BlockedMult.pas:477-593). */
MRM_API void mrm_blocked_mult_t1(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                         int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                         int64_t LineWidth2, int64_t blockSize, int32_t op, double* mem);

/* blockedMultT2Func: dest(height1 x height2) (op)= mt1 * mt2^T, width1 = width2.
 * Replaces BlockMatrixMultiplicationT2 (src/BlockedMult.pas:596-724). */
MRM_API void mrm_blocked_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                         int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                         int64_t LineWidth2, int64_t blockSize, int32_t op, double* mem);

/* ---- device-pointer GEMM -------------------------------------------------- */

/* C(M x N) (op)= opA(A) * opB(B) with trans = MRM_NN / MRM_TN / MRM_NT, all
 * pointers in device memory, strides in bytes.  The kernel behind every
 * GEMM entry above; no reference counterpart (device-resident callers). */
MRM_API int32_t mrm_dev_dgemm(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, void* stream);

/* Bytes of the device array mrm_dev_dgemm_noalloc keeps its partial products in. */
#define MRM_NOALLOC_SCRATCH_BYTES (128ll << 20)

/* mrm_dev_dgemm without any allocation, for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  Same arguments, kernels and K
 * split as mrm_dev_dgemm, so the result is the same bit for bit; a split
 * product (few output tiles, long K: X^T X) keeps its partial products in a
 * library-owned device array of MRM_NOALLOC_SCRATCH_BYTES per GPU instead of
 * a stream-ordered allocation.  Every split fits (at most 128 MiB); a larger
 * need would be refused with MRM_BAD_DIMS.  All calls on one GPU share that
 * array: calls that overlap in time on one GPU race. */
MRM_API int32_t mrm_dev_dgemm_noalloc(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A,
                                      int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op,
                                      void* stream);

/* ---- threaded callers: one device GEMM instead of up to 64 tile calls ---- */

/* ThrMatrixMult (reference src/ThreadedMatrixOperations.pas:729-735, the body
 * of TThreadedMatrix.Mult, src/ThreadedMatrix.pas:347-357): dest (op)=
 * mt1 * mt2 with width1 = height2.  The reference cuts the product into up to
 * 64 tiles for its thread pool (ThrMatrixMultEx2, :612-718); here the whole
 * product is one staging of each operand and one GEMM launch. */
MRM_API void mrm_thr_mult(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2, int64_t width1,
                          int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1, int64_t LineWidth2,
                          int32_t op);
/* ThrMatrixMultT1 (src/ThreadedMatrixOperations.pas:34): dest (op)= mt1^T * mt2. */
MRM_API void mrm_thr_mult_t1(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                             int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                             int64_t LineWidth2, int32_t op);
/* ThrMatrixMultT2 (src/ThreadedMatrixOperations.pas:40): dest (op)= mt1 * mt2^T. */
MRM_API void mrm_thr_mult_t2(double* dest, int64_t destLineWidth, const double* mt1, const double* mt2,
                             int64_t width1, int64_t height1, int64_t width2, int64_t height2, int64_t LineWidth1,
                             int64_t LineWidth2, int32_t op);

/* ---- FP64-accurate GEMM from INT8 slices (Ozaki scheme) ------------------ */

/* Device bytes of workspace mrm_dev_dgemm_ozaki needs for these sizes. */
MRM_API int64_t mrm_ozaki_workspace_bytes(int64_t M, int64_t N, int64_t K, int32_t slices);

/* C(M x N) (op)= opA(A) * opB(B) in FP64, computed on the INT8 matrix cores:
 * every row of opA(A) and column of opB(B) is scaled by a power of two and cut
 * into `slices` (1..8) signed 7-bit slices; slice products are summed exactly
 * in int32 and in FP64 across slices.  Per element the error is at most
 * 4 (slices+1) K 2^(-7 slices) max|row i of opA(A)| max|column j of opB(B)|
 * plus FP64 rounding of the slice sums (6-7 slices reach FP64 accuracy).
 * Inputs must be finite.  `work` is device memory
 * of at least mrm_ozaki_workspace_bytes(); it must not be shared by calls in
 * flight at the same time.  Asynchronous on `stream`.  The GEMM the reference
 * computes with MatrixMult / MatrixMultT1 / MatrixMultT2 (src/MatrixASMStubSwitch.pas
 * :769-863); the method has no reference counterpart. */
MRM_API int32_t mrm_dev_dgemm_ozaki(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                                    const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op, int32_t slices,
                                    void* work, int64_t work_bytes, void* stream);

/* Slice kinds of mrm_dev_dgemm_ozaki_ex: signed 7-bit digits held as INT8
 * (v_mfma_i32_32x32x32_i8, int32 sums) or as BF16 (v_mfma_f32_32x32x16_bf16,
 * FP32 sums over 128 k, exact, added into int32).  Both give the same digits
 * and the same exact diagonal sums, so their results are equal bit for bit. */
#define MRM_OZAKI_INT8 0
#define MRM_OZAKI_BF16 1

/* Device bytes of workspace mrm_dev_dgemm_ozaki_ex needs (BF16 slices take
 * two bytes a digit); 0 for a bad size, slice count or kind. */
MRM_API int64_t mrm_ozaki_workspace_bytes_ex(int64_t M, int64_t N, int64_t K, int32_t slices, int32_t kind);

/* mrm_dev_dgemm_ozaki with the slice kind chosen (MRM_OZAKI_INT8 or
 * MRM_OZAKI_BF16); the same GEMM, arguments and error bound. */
MRM_API int32_t mrm_dev_dgemm_ozaki_ex(int32_t trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                                       const double* B, int64_t ldb, double* C, int64_t ldc, int32_t op,
                                       int32_t slices, int32_t kind, void* work, int64_t work_bytes, void* stream);

/* Host-pointer form of mrm_dev_dgemm_ozaki with the arguments of
 * TMatrixBlockedMultfunc (src/MatrixASMStubSwitch.pas:352) plus trans and
 * slices; the library holds the workspace.  Returns the status. */
MRM_API int32_t mrm_mult_ozaki(int32_t trans, double* dest, int64_t destLineWidth, const double* mt1,
                               const double* mt2, int64_t width1, int64_t height1, int64_t width2, int64_t height2,
                               int64_t LineWidth1, int64_t LineWidth2, int32_t op, int32_t slices);

/* ---- LU factorisation and linear solve ------------------------------------ */

/* MatrixLUDecompInPlace (reference src/LinAlgLU.pas:283-295, recursion :181-281):
 * in-place LU with partial pivoting of the width x width matrix A; L unit lower
 * below the diagonal, U on and above it; indx[i] = row exchanged with row i at
 * step i.  Returns MRM_OK (leOk) or MRM_SINGULAR (leSingular). */
MRM_API int32_t mrm_lu_decomp_inplace(double* A, int64_t LineWidthA, int64_t width, int32_t* indx);

/* MatrixLUBackSubst (reference src/LinAlgLU.pas:315-385): solves LU x = P b for
 * the one right-hand side B (element stride LineWidthB bytes) in place. */
MRM_API int32_t mrm_lu_backsubst(const double* LUDecomp, int64_t LineWidthLU, int64_t width, const int32_t* indx,
                                 double* B, int64_t LineWidthB);

/* MatrixLinEQSolve (reference src/LinAlgLU.pas:481-592): X (width x width2) =
 * A^-1 B, followed by NumRefinments FP64 refinement steps on every column. */
MRM_API int32_t mrm_lineq_solve(const double* A, int64_t LineWidthA, int64_t width, const double* B,
                                int64_t LineWidthB, double* X, int64_t LineWidthX, int64_t width2,
                                int64_t NumRefinments);

/* MatrixInverseInPlace (reference src/LinAlgLU.pas:386-427): A (width x
 * width) is replaced by its inverse, column j being (LU)^-1 P e_j.  Returns
 * MRM_OK (leOk) or MRM_SINGULAR (leSingular; A is then left unchanged). */
MRM_API int32_t mrm_inverse_inplace(double* A, int64_t LineWidthA, int64_t width);

/* MatrixDeterminant (reference src/LinAlgLU.pas:429-478): *det = the product
 * of U's diagonal, negated once per row interchange, 0 for a singular matrix.
 * Returns MRM_OK or MRM_SINGULAR (both set *det) or an error. */
MRM_API int32_t mrm_determinant(const double* A, int64_t LineWidthA, int64_t width, double* det);

/* Device LU of the n x n matrix A (stride lda bytes): indx as above; perm
 * (optional, n entries) receives the row permutation, row i of P*A being row
 * perm[i] of A.  Synchronises `stream` to return the status.  No reference
 * counterpart (device-resident callers). */
MRM_API int32_t mrm_dev_lu_factor(double* A, int64_t lda, int64_t n, int32_t* indx, int32_t* perm, void* stream);

/* Device solve X = (LU)^-1 P B for nrhs columns using the perm of
 * mrm_dev_lu_factor.  B and X must not overlap.  Asynchronous. */
MRM_API int32_t mrm_dev_lu_solve(const double* LU, int64_t lda, int64_t n, const int32_t* perm, const double* B,
                                 int64_t ldb, double* X, int64_t ldx, int64_t nrhs, void* stream);

/* ---- mixed precision: low-precision update and LU-IR solve ---------------- */

/* Device bytes of packed operands mrm_dev_gemm_lowp needs. */
MRM_API int64_t mrm_lowp_workspace_bytes(int64_t M, int64_t N, int64_t K, int32_t mode);

/* C (M x N, FP64) -= A (M x K) * B (K x N), FP64 operands in device memory
 * (byte strides), the product rounded to FP32 (mode MRM_SOLVE_FP32_IR) or BF16
 * (MRM_SOLVE_BF16_IR; MRM_SOLVE_BF16X3_IR: two BF16 pieces per operand, three
 * products) and accumulated in FP32 on the matrix cores.  The
 * Schur-complement update of the mixed-precision LU (the reference computes it
 * in FP64 through the blocked multiply, src/LinAlgLU.pas:211-227).
 * Asynchronous on `stream`; `work` must not be shared by calls in flight. */
MRM_API int32_t mrm_dev_gemm_lowp(int32_t mode, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                                  const double* B, int64_t ldb, double* C, int64_t ldc, void* work,
                                  int64_t work_bytes, void* stream);

/* MatrixLinEQSolve on device memory (reference src/LinAlgLU.pas:481-592) with
 * a choice of factorisation precision: X (n x nrhs) = A^-1 B.  The LU of a copy
 * of A runs in `mode` (MRM_SOLVE_FP64 / _FP32_IR / _BF16_IR: trailing updates
 * on the FP32 / BF16 matrix cores, FP64 panels and pivoting), then FP64
 * refinement x += (LU)^-1 (b - A x) runs until the normwise backward error
 * max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= tol or
 * max_iter steps.  A low-precision run whose plain refinement stalls continues
 * with GMRES corrections (see mrm_dev_lineq_solve_ir_stats); one whose GMRES
 * refinement stalls too is redone in FP64 (*fell_back = 1).  Returns MRM_OK, MRM_SINGULAR or MRM_NOT_CONVERGED
 * (X then holds the last iterate).  iterations / backward_error / fell_back
 * may be NULL.  Synchronises `stream`. */
MRM_API int32_t mrm_dev_lineq_solve_ir(int32_t mode, const double* A, int64_t lda, int64_t n, const double* B,
                                       int64_t ldb, double* X, int64_t ldx, int64_t nrhs, int32_t max_iter,
                                       double tol, int32_t* iterations, double* backward_error, int32_t* fell_back,
                                       void* stream);

/* What a refined solve did (mrm_dev_lineq_solve_ir_stats). */
typedef struct mrm_solve_stats {
    int32_t iterations;         /* refinement steps of the returned solution */
    int32_t fell_back;          /* 1: the low-precision run stalled, the LU was redone in FP64 */
    int32_t gmres;              /* 1: the plain step stalled and GMRES preconditioned by the
                                   factors supplied the corrections (GMRES-IR) */
    int32_t krylov_steps;       /* Krylov steps over all GMRES corrections of the solution */
    double backward_error;      /* normwise backward error of the returned solution */
    int32_t lowp_iterations;    /* after a fall-back: steps of the low-precision attempt */
    int32_t lowp_krylov_steps;  /* and its Krylov steps */
    double lowp_backward_error; /* and its last backward error */
} mrm_solve_stats;

/* mrm_dev_lineq_solve_ir with every statistic of the run (stats may be NULL).
 * With FP32 / BF16 factors, when a plain refinement step gains less than
 * tenfold the solve continues with GMRES corrections preconditioned by the same
 * factors (one Krylov space per right-hand side, FP64 products by A, classical
 * Gram-Schmidt twice; Carson and Higham's GMRES-IR); a GMRES step that also
 * gains less than tenfold ends the attempt, and an FP64 factorisation answers.  The reference refines with the plain step only
 * (src/LinAlgLU.pas:542-586). */
MRM_API int32_t mrm_dev_lineq_solve_ir_stats(int32_t mode, const double* A, int64_t lda, int64_t n, const double* B,
                                             int64_t ldb, double* X, int64_t ldx, int64_t nrhs, int32_t max_iter,
                                             double tol, mrm_solve_stats* stats, void* stream);

/* MatrixLinEQSolve (src/LinAlgLU.pas:481-592) with a factorisation precision.
 * mode MRM_SOLVE_FP64 is mrm_lineq_solve; the IR modes refine to a backward
 * error of 1e-14 in at most NumRefinments steps (50 when 0). */
MRM_API int32_t mrm_lineq_solve_ex(const double* A, int64_t LineWidthA, int64_t width, const double* B,
                                   int64_t LineWidthB, double* X, int64_t LineWidthX, int64_t width2,
                                   int64_t NumRefinments, int32_t mode);

/* ---- covariance (the PCA caller of the multiply) -------------------------- */

/* Covariance (reference src/Statistics.pas:599-639): every row of A (height x
 * width) is an observation, every column a variable; dest (width x width) =
 * Ac^T Ac / m with Ac = A minus its column means and m = height - 1 when
 * unbiased (height otherwise).  Computed without a transposed copy: the
 * centred copy stays in device memory and Ac^T Ac takes the symmetric
 * split-K product.  A single row or column gives its variance in dest[0]
 * (MatrixVar, as the reference does).  Returns the status. */
MRM_API int32_t mrm_covariance(double* dest, int64_t destLineWidth, const double* A, int64_t LineWidthA,
                               int64_t width, int64_t height, int32_t unbiased);

/* Device bytes of workspace mrm_dev_covariance needs. */
MRM_API int64_t mrm_covariance_workspace_bytes(int64_t width, int64_t height);

/* Covariance of the device matrix X (height x width, stride ldx bytes) into
 * C (width x width, stride ldc); `work` holds the centred copy and the means
 * (mrm_covariance_workspace_bytes).  X is left unchanged.  Both dimensions
 * must be at least 2.  Asynchronous on `stream`. */
MRM_API int32_t mrm_dev_covariance(const double* X, int64_t ldx, int64_t width, int64_t height, double* C,
                                   int64_t ldc, int32_t unbiased, void* work, int64_t work_bytes, void* stream);

/* ---- batched small problems (persistent launches replayed from a hipGraph) */

/* One GEMM C = A * B (row-major, device pointers, strides in bytes). */
typedef struct mrm_gemm_desc {
    int64_t M, N, K;
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    double* C;
    int64_t ldc;
} mrm_gemm_desc;

/* One system A x = b, n <= 128 (device pointers; A and b are left unchanged). */
typedef struct mrm_lu_desc {
    int64_t n;
    const double* A;
    int64_t lda;
    const double* b;
    double* x;
} mrm_lu_desc;

/* Plans `count` FP64 GEMMs of any sizes (each M, N <= 2^20) as one persistent
 * grouped launch over T x T output tiles, T in {32, 64, 128}: tile = 0 lets the
 * library time the batch with each T and keep the fastest (this runs the
 * batch); the launch is captured into a hipGraph.  The descriptors are copied;
 * the matrices must stay allocated while the plan is used.  The callers
 * replaced are loops of MatrixMult over small matrices (the reference's
 * small-size path FMAMatrixMult, src/FMAMatrixOperations.pas:94-143). */
MRM_API int32_t mrm_batched_dgemm_plan(const mrm_gemm_desc* problems, int64_t count, int32_t tile, void** plan);

/* mrm_batched_dgemm_plan with a precision: MRM_SOLVE_FP64, or
 * MRM_SOLVE_FP32_IR -- every operand split into FP32 pieces x = xh + xl, the
 * product Ah Bh + (Ah Bl + Al Bh) formed on the FP32 matrix cores (error
 * compensation after Ootomo and Yokota) with the sums of each K chunk of 16
 * added in FP64: an error of FP32 sums over 16 products, not an FP64 result
 * (tiles 64 / 128 or 0).  The reference multiplies in FP64 only. */
MRM_API int32_t mrm_batched_dgemm_plan_ex(const mrm_gemm_desc* problems, int64_t count, int32_t tile, int32_t mode,
                                          void** plan);

/* Plans `count` linear solves A_i x_i = b_i (n_i <= 128, one right-hand side)
 * with partial pivoting, grouped by size class into persistent launches
 * captured in a hipGraph; mode MRM_SOLVE_FP64, or MRM_SOLVE_FP32_IR (FP32
 * factors, two FP64 refinement steps against A_i).  status (device, count
 * int32, may be NULL) receives MRM_OK or MRM_SINGULAR per system at each run.
 * The per-system MatrixLinEQSolve of the reference (src/LinAlgLU.pas:481-592). */
MRM_API int32_t mrm_batched_lu_plan(const mrm_lu_desc* systems, int64_t count, int32_t mode, int32_t* status,
                                    void** plan);

/* Replays a plan on `stream` (one hipGraphLaunch); asynchronous. */
MRM_API int32_t mrm_batched_run(void* plan, void* stream);
/* The GEMM tile size a plan uses (64 threads per system for solve plans). */
MRM_API int32_t mrm_batched_plan_tile(void* plan);
MRM_API void mrm_batched_destroy(void* plan);

#ifdef __cplusplus
}
#endif
#endif
