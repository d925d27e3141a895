/*
 * mrm_oracle.h -- CPU restatement of mrMath's FP64 GEMM / LU / solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * GPU library (mrmath_amd/libmrmath_gpu.so) and the CPU timing baseline of
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never calls into it.
 *
 * Conventions follow the reference (src/MatrixConst.pas, src/Matrix.pas):
 * matrices are row-major, "width" = number of columns, "height" = number of
 * rows, and every LineWidth is the row stride in BYTES.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_mt_*                 marsenneTwistOrig.txt golden file (reference
 *                            Test/TestRandom.pas:249-311), byte for byte.
 *   orc_generic_mtx_mult*    exact KATs of Test/TestSimpleMatrixOperations.pas:
 *                            1712-1729, closed forms of Test/TestMatrixClass.pas:
 *                            473-509, integer-valued inputs (exact in FP64) and
 *                            a brute-force numpy product.
 *   orc_blocked_mult*        same pins through the op (none/add/sub) algebra.
 *   orc_lu_*, orc_lineq_*,   exact KATs of Test/TestLineEquations.pas:635-831,
 *   orc_inverse, orc_det     P*A = L*U reconstruction, numpy.linalg.solve /
 *                            det / inv on random inputs, singular detection.
 *   orc_randgauss            algorithm only (Marsaglia polar); pinned through
 *                            the MT stream plus the closed-form transform.
 *   orc_ozaki_gemm           exact products of integer inputs the slices hold
 *                            completely, the a-priori error bound and its
 *                            2^-7 decay per slice, a numpy digit split.
 *   orc_lineq_solve with NumRefinments > 0: "parity unpinned" -- the reference
 *                            has no test that runs refinement.
 */
#ifndef MRM_ORACLE_H
#define MRM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MT19937 (reference src/RandomEng.pas:626-764) ---- */
typedef struct orc_mt {
    uint32_t mt[624];
    int32_t mti;
} orc_mt;

void     orc_mt_init(orc_mt* s, int32_t seed);                        /* InitMersenne */
void     orc_mt_init_by_array(orc_mt* s, const uint32_t* key, int32_t n); /* InitMersenneByArr */
uint32_t orc_mt_int32(orc_mt* s);                                     /* RandMersenneDblWord */
double   orc_mt_res53(orc_mt* s);                                     /* RandMersenneDbl */
int32_t  orc_mt_randint(orc_mt* s, int32_t range);                    /* RandMersenneInt */
double   orc_randgauss(orc_mt* s, int32_t* flag, double* g2);         /* RandGauss */

/* TDoubleMatrix.InitRandom(raMersenneTwister, seed) (src/DblMatrix.pas:2751-2759):
 * fill row-major with res53, then seed := RandInt (chaining). */
void orc_fill_rand(double* dest, int64_t lw, int64_t w, int64_t h, int32_t* seed);
/* 2*res53 - 1 fill with init_genrand(seed): the synthetic-input generator. */
void orc_fill_pm1(double* dest, int64_t lw, int64_t w, int64_t h, int32_t seed);
/* Wide-range fill (2*res53-1)*2^e, e = RandInt(41) - 20, per entry. */
void orc_fill_pm1_wide(double* dest, int64_t lw, int64_t w, int64_t h, int32_t seed, int32_t emax);

/* ---- Generic GEMMs (reference src/SimpleMatrixOperations.pas:1561-1701) ---- */
void orc_generic_mtx_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                          int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2);
void orc_generic_mtx_mult_transp(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                                 int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2);
void orc_generic_transp_mtx_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                                 int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2);

/* Blocked drivers (reference src/BlockedMult.pas:350-724): dest (op) A*B,
 * op 0 = doNone, 1 = doAdd, 2 = doSub.  trans: 0 = NN, 1 = T1 (A^T B), 2 = T2 (A B^T).
 * The tile product is the generic loop; the tile sum is accumulated per K-block
 * in a tile buffer and then applied with op, as the reference does. */
void orc_blocked_mult(int32_t trans, double* dest, int64_t destLW, const double* mt1, const double* mt2,
                      int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2,
                      int64_t blockSize, int32_t op);

/* ---- LU / solve (reference src/LinAlgLU.pas:120-592) ---- */
enum { ORC_LE_OK = 0, ORC_LE_SINGULAR = 1 };
int32_t orc_lu_decomp_inplace(double* A, int64_t LWA, int64_t width, int32_t* indx);
void    orc_lu_backsubst(const double* LU, int64_t LWLU, int64_t width, const int32_t* indx,
                         double* B, int64_t LWB);
int32_t orc_lineq_solve(const double* A, int64_t LWA, int64_t width, const double* B, int64_t LWB,
                        double* X, int64_t LWX, int64_t width2, int64_t numRefinements);
int32_t orc_inverse_inplace(double* A, int64_t LWA, int64_t width);
double  orc_determinant(const double* A, int64_t LWA, int64_t width);

/* ---- Error metrics ---- */
/* ||C - Cref||_F / (||A||_F * ||B||_F) over the logical (w x h) regions. */
double orc_gemm_normwise_err(const double* C, int64_t LWC, const double* Cref, int64_t LWR,
                             int64_t wC, int64_t hC, const double* A, int64_t LWA, int64_t wA, int64_t hA,
                             const double* B, int64_t LWB, int64_t wB, int64_t hB);
/* FP64 GEMM C (M x N) (op)= A (M x K) * B (K x N) through `slices` 7-bit
 * integer slices per operand, accumulated exactly per diagonal and K chunk
 * and in FP64 across them (mrm_ozaki_oracle.c; no reference counterpart). */
int32_t orc_ozaki_gemm(double* C, int64_t LWC, const double* A, int64_t LWA, const double* B, int64_t LWB,
                       int64_t M, int64_t N, int64_t K, int32_t op, int32_t slices, int64_t kchunk);
/* max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf), residual in long double. */
/* Covariance (src/Statistics.pas:599-639) */
void orc_covariance(double* dest, int64_t LWD, const double* A, int64_t LWA, int64_t width, int64_t height,
                    int32_t unbiased);
double orc_backward_error(const double* A, int64_t LWA, int64_t n, const double* X, int64_t LWX,
                          const double* B, int64_t LWB, int64_t nrhs);

#ifdef __cplusplus
}
#endif
#endif
