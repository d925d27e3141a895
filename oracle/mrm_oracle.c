/*
 * mrm_oracle.c -- CPU restatement of mrMath's FP64 GEMM / LU / solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline, never the
 * product).  Built with -O2 -ffp-contract=off so that no FMA contraction
 * changes the order of operations the reference's Pascal loops use.
 * Each function cites the reference file:line it restates.
 */
#include "mrm_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

#define ELEM(p, lw, x, y) ((double*)((char*)(p) + (int64_t)(y) * (lw)) + (x))
#define CELEM(p, lw, x, y) ((const double*)((const char*)(p) + (int64_t)(y) * (lw)) + (x))

/* ------------------------------------------------------------------------ */
/* MT19937: src/RandomEng.pas:626-764                                        */
/* ------------------------------------------------------------------------ */
enum { MT_N = 624, MT_M = 397 };
#define MT_MATRIX_A 0x9908b0dfu
#define MT_UPPER 0x80000000u
#define MT_LOWER 0x7fffffffu

/* InitMersenne (src/RandomEng.pas:626-644); seed 0 (time seed) is not used here. */
void orc_mt_init(orc_mt* s, int32_t seed)
{
    s->mt[0] = (uint32_t)seed;
    for (uint32_t i = 1; i < MT_N; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + i;
    s->mti = MT_N;
}

/* InitMersenneByArr (src/RandomEng.pas:646-692) */
void orc_mt_init_by_array(orc_mt* s, const uint32_t* key, int32_t n)
{
    orc_mt_init(s, 19650218);
    uint32_t i = 1, j = 0;
    int32_t k = (MT_N > n ? MT_N : n);
    for (; k > 0; k--) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1664525u)) + key[j] + j;
        i++; j++;
        if (i >= MT_N) { s->mt[0] = s->mt[MT_N - 1]; i = 1; }
        if (j >= (uint32_t)n) j = 0;
    }
    for (k = MT_N - 1; k > 0; k--) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1566083941u)) - i;
        i++;
        if (i >= MT_N) { s->mt[0] = s->mt[MT_N - 1]; i = 1; }
    }
    s->mt[0] = 0x80000000u;
}

/* RandMersenneDblWord (src/RandomEng.pas:694-732) */
uint32_t orc_mt_int32(orc_mt* s)
{
    static const uint32_t mag01[2] = { 0u, MT_MATRIX_A };
    uint32_t y;
    if (s->mti > MT_N) orc_mt_init(s, 5489);
    if (s->mti >= MT_N) {
        int kk;
        for (kk = 0; kk < MT_N - MT_M; kk++) {
            y = (s->mt[kk] & MT_UPPER) | (s->mt[kk + 1] & MT_LOWER);
            s->mt[kk] = s->mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < MT_N - 1; kk++) {
            y = (s->mt[kk] & MT_UPPER) | (s->mt[kk + 1] & MT_LOWER);
            s->mt[kk] = s->mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (s->mt[MT_N - 1] & MT_UPPER) | (s->mt[0] & MT_LOWER);
        s->mt[MT_N - 1] = s->mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1u];
        s->mti = 0;
    }
    y = s->mt[s->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* RandMersenneDbl (src/RandomEng.pas:749-764): genrand_res53 */
double orc_mt_res53(orc_mt* s)
{
    const double multA = 67108864.0;
    const double multB = 1.0 / 9007199254740992.0;
    double x = (double)(orc_mt_int32(s) >> 5);
    double y = (double)(orc_mt_int32(s) >> 6);
    return (x * multA + y) * multB;
}

/* RandMersenneInt (src/RandomEng.pas:734-740) */
int32_t orc_mt_randint(orc_mt* s, int32_t range)
{
    uint32_t y = orc_mt_int32(s);
    return (int32_t)(y & 0x7fffffffu) % range;
}

/* RandGauss, Marsaglia polar (src/RandomEng.pas:283-307) */
double orc_randgauss(orc_mt* s, int32_t* flag, double* g2)
{
    double r;
    if (*flag) {
        r = *g2;
    } else {
        double v1, v2, rsq;
        do {
            v1 = 2.0 * orc_mt_res53(s) - 1.0;
            v2 = 2.0 * orc_mt_res53(s) - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (!(rsq < 1.0 && rsq != 0.0));
        r = sqrt(-2.0 * log(rsq) / rsq);
        *g2 = v1 * r;
        r = v2 * r;
    }
    *flag = !*flag;
    return r;
}

/* TDoubleMatrix.InitRandom (src/DblMatrix.pas:2751-2759) with the row-major
 * fill order of GenericMtxFunc (src/SimpleMatrixOperations.pas:2325-2339). */
void orc_fill_rand(double* dest, int64_t lw, int64_t w, int64_t h, int32_t* seed)
{
    orc_mt s;
    orc_mt_init(&s, *seed);
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++)
            *ELEM(dest, lw, x, y) = orc_mt_res53(&s);
    *seed = orc_mt_randint(&s, 2147483647); /* RandInt = fRandInt(MaxInt) */
}

void orc_fill_pm1(double* dest, int64_t lw, int64_t w, int64_t h, int32_t seed)
{
    orc_mt s;
    orc_mt_init(&s, seed);
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++)
            *ELEM(dest, lw, x, y) = 2.0 * orc_mt_res53(&s) - 1.0;
}

void orc_fill_pm1_wide(double* dest, int64_t lw, int64_t w, int64_t h, int32_t seed, int32_t emax)
{
    orc_mt s;
    orc_mt_init(&s, seed);
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) {
            double v = 2.0 * orc_mt_res53(&s) - 1.0;
            int e = orc_mt_randint(&s, 2 * emax + 1) - emax;
            *ELEM(dest, lw, x, y) = ldexp(v, e);
        }
}

/* ------------------------------------------------------------------------ */
/* Generic GEMMs                                                             */
/* ------------------------------------------------------------------------ */

/* GenericMtxMult (src/SimpleMatrixOperations.pas:1561-1591): dest = mt1*mt2,
 * dest is h1 x w2, sum over idx = 0..w1-1 in order, separate mul and add. */
void orc_generic_mtx_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                          int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2)
{
    (void)h2;
    for (int64_t y = 0; y < h1; y++) {
        const double* a = CELEM(mt1, LW1, 0, y);
        for (int64_t x = 0; x < w2; x++) {
            double acc = 0.0;
            for (int64_t k = 0; k < w1; k++) {
                double p = a[k] * *CELEM(mt2, LW2, x, k);
                acc = acc + p;
            }
            *ELEM(dest, destLW, x, y) = acc;
        }
    }
}

/* GenericMtxMultTransp (src/SimpleMatrixOperations.pas:1593-1617):
 * dest = mt1 * mt2^T, dest is h1 x h2, K = w1 = w2. */
void orc_generic_mtx_mult_transp(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                                 int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2)
{
    (void)w2;
    for (int64_t y = 0; y < h1; y++) {
        const double* a = CELEM(mt1, LW1, 0, y);
        for (int64_t x = 0; x < h2; x++) {
            const double* b = CELEM(mt2, LW2, 0, x);
            double acc = 0.0;
            for (int64_t k = 0; k < w1; k++) {
                double p = a[k] * b[k];
                acc = acc + p;
            }
            *ELEM(dest, destLW, x, y) = acc;
        }
    }
}

/* GenericTranspMtxMult (src/SimpleMatrixOperations.pas:1669-1701):
 * dest = mt1^T * mt2, dest is w1 x w2, K = h1 = h2. */
void orc_generic_transp_mtx_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                                 int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2)
{
    (void)h1;
    for (int64_t y = 0; y < w1; y++) {
        for (int64_t x = 0; x < w2; x++) {
            double acc = 0.0;
            for (int64_t k = 0; k < h2; k++) {
                double p = *CELEM(mt1, LW1, y, k) * *CELEM(mt2, LW2, x, k);
                acc = acc + p;
            }
            *ELEM(dest, destLW, x, y) = acc;
        }
    }
}

/* Next2Pwr (src/MathUtilFunc.pas:555-560) */
static int64_t next2pwr(int64_t num, int64_t maxSize)
{
    int64_t r = 1;
    while (r < maxSize && r < num) r <<= 1;
    return r;
}

/* Blocked drivers: BlockMatrixMultiplication / T1 / T2 (src/BlockedMult.pas:350-724).
 * Per output tile: actBlk = 0; for each K-block: actBlk += tile product
 * (src/BlockedMult.pas:423-456); then dest = / += / -= actBlk (:459-463). */
void orc_blocked_mult(int32_t trans, double* dest, int64_t destLW, const double* mt1, const double* mt2,
                      int64_t w1, int64_t h1, int64_t w2, int64_t h2, int64_t LW1, int64_t LW2,
                      int64_t bs, int32_t op)
{
    int64_t M, N, K;
    if (w1 == 0 || w2 == 0 || h1 == 0 || h2 == 0) return;   /* :374-375 */
    if (trans == 0)      { M = h1; N = w2; K = w1; }
    else if (trans == 1) { M = w1; N = w2; K = h1; }
    else                 { M = h1; N = h2; K = w1; }
    if (bs < 2) bs = 2;
    if (bs > K) bs = next2pwr(K, bs);  /* the "mem = nil" adjustment (:381-382) */

    double* act = (double*)malloc(sizeof(double) * bs * bs);
    double* mul = (double*)malloc(sizeof(double) * bs * bs);
    for (int64_t by = 0; by < M; by += bs) {
        int64_t bh = (M - by < bs) ? M - by : bs;
        for (int64_t bx = 0; bx < N; bx += bs) {
            int64_t bw = (N - bx < bs) ? N - bx : bs;
            memset(act, 0, sizeof(double) * bs * bs);
            for (int64_t bk = 0; bk < K; bk += bs) {
                int64_t gw = (K - bk < bs) ? K - bk : bs;
                for (int64_t y = 0; y < bh; y++)
                    for (int64_t x = 0; x < bw; x++) {
                        double acc = 0.0;
                        for (int64_t k = 0; k < gw; k++) {
                            double a, b;
                            if (trans == 1) a = *CELEM(mt1, LW1, by + y, bk + k);
                            else            a = *CELEM(mt1, LW1, bk + k, by + y);
                            if (trans == 2) b = *CELEM(mt2, LW2, bk + k, bx + x);
                            else            b = *CELEM(mt2, LW2, bx + x, bk + k);
                            double p = a * b;
                            acc = acc + p;
                        }
                        mul[y * bs + x] = acc;
                    }
                for (int64_t i = 0; i < bh * bs; i++) act[i] = act[i] + mul[i];
            }
            for (int64_t y = 0; y < bh; y++)
                for (int64_t x = 0; x < bw; x++) {
                    double* d = ELEM(dest, destLW, bx + x, by + y);
                    if (op == 0)      *d = act[y * bs + x];
                    else if (op == 1) *d = *d + act[y * bs + x];
                    else              *d = *d - act[y * bs + x];
                }
        }
    }
    free(act);
    free(mul);
}

/* ------------------------------------------------------------------------ */
/* LU decomposition and solve                                                */
/* ------------------------------------------------------------------------ */

/* LUSwap (src/LinAlgLU.pas:120-136): row i <-> row indx[i] for i = k1..k2 */
static void lu_swap(double* A, int64_t LW, int64_t width, int64_t k1, int64_t k2, const int32_t* indx)
{
    for (int64_t i = k1; i <= k2; i++) {
        if (indx[i] != i) {
            double* r1 = ELEM(A, LW, 0, i);
            double* r2 = ELEM(A, LW, 0, indx[i]);
            for (int64_t x = 0; x < width; x++) { double t = r1[x]; r1[x] = r2[x]; r2[x] = t; }
        }
    }
}

/* LUBacksup (src/LinAlgLU.pas:138-169): B <- L^-1 B, L unit lower (height x height). */
static void lu_backsup(const double* A, int64_t width, int64_t height, double* B, int64_t LW)
{
    for (int64_t k = 0; k < height; k++) {
        const double* bk = CELEM(B, LW, 0, k);
        for (int64_t i = k + 1; i < height; i++) {
            double l = *CELEM(A, LW, k, i);
            double* bi = ELEM(B, LW, 0, i);
            for (int64_t j = 0; j < width; j++) {
                double p = bk[j] * l;
                bi[j] = bi[j] - p;
            }
        }
    }
}

/* InternalRecursiveMatrixLUDecompInPlace (src/LinAlgLU.pas:181-281). */
static int32_t lu_rec(double* A, int64_t width, int64_t height, int32_t* indx, int64_t LW)
{
    int64_t mn = width < height ? width : height;
    if (mn > 1) {
        int64_t nleft = mn / 2;
        int64_t nright = width - nleft;
        int32_t r = lu_rec(A, nleft, height, indx, LW);
        if (r != ORC_LE_OK) return r;

        double* a12 = ELEM(A, LW, nleft, 0);
        lu_swap(a12, LW, nright, 0, nleft - 1, indx);
        if (nright > 1) lu_backsup(A, nright, nleft, a12, LW);

        /* A22 = A22 - A21*A12 (product first, then subtract: the doSub
         * semantics of MatrixMultEx at :218-226) */
        double* a22 = ELEM(A, LW, nleft, nleft);
        const double* a21 = CELEM(A, LW, 0, nleft);
        int64_t mh = height - nleft;
        if (mh > 0 && nright > 0) {
            double* tmp = (double*)malloc(sizeof(double) * mh * nright);
            orc_generic_mtx_mult(tmp, nright * 8, a21, a12, nleft, mh, nright, nleft, LW, LW);
            for (int64_t y = 0; y < mh; y++) {
                double* c = ELEM(a22, LW, 0, y);
                for (int64_t x = 0; x < nright; x++) c[x] = c[x] - tmp[y * nright + x];
            }
            free(tmp);
        }
        r = lu_rec(a22, nright, height - nleft, indx + nleft, LW);
        if (r != ORC_LE_OK) return r;
        for (int64_t i = nleft; i < width; i++) indx[i] += (int32_t)nleft;
        lu_swap(A, LW, nleft, nleft, mn - 1, indx);
        return ORC_LE_OK;
    }
    /* leaf: one column -- first max-abs pivot, reciprocal scale, swap (:239-280) */
    double maxVal = 0.0, absMax = 0.0;
    int64_t idx = -1;
    for (int64_t i = 0; i < height; i++) {
        double v = *CELEM(A, LW, 0, i);
        if (fabs(v) > absMax) { idx = i; maxVal = v; absMax = fabs(v); }
    }
    indx[0] = (int32_t)idx;
    if (fabs(maxVal) > 10.0 / DBL_MAX) {
        double inv = 1.0 / maxVal;
        for (int64_t i = 0; i < height; i++) {
            double* p = ELEM(A, LW, 0, i);
            *p = (*p * inv) + 0.0;
        }
        *ELEM(A, LW, 0, idx) = *A;
        *A = maxVal;
        return ORC_LE_OK;
    }
    return ORC_LE_SINGULAR;
}

/* MatrixLUDecompInPlace (src/LinAlgLU.pas:283-295) */
int32_t orc_lu_decomp_inplace(double* A, int64_t LWA, int64_t width, int32_t* indx)
{
    memset(indx, 0, sizeof(int32_t) * width);
    return lu_rec(A, width, width, indx, LWA);
}

/* MatrixLUBackSubst (src/LinAlgLU.pas:315-385) for one strided RHS column. */
void orc_lu_backsubst(const double* LU, int64_t LWLU, int64_t width, const int32_t* indx,
                      double* B, int64_t LWB)
{
    int64_t ii = -1;
    for (int64_t i = 0; i < width; i++) {
        int64_t ip = indx[i];
        double* pB = ELEM(B, LWB, 0, ip);
        double sum = *pB;
        *pB = *ELEM(B, LWB, 0, i);
        if (ii >= 0) {
            const double* row = CELEM(LU, LWLU, 0, i);
            for (int64_t j = ii; j < i; j++) {
                double p = row[j] * *ELEM(B, LWB, 0, j);
                sum = sum - p;
            }
        } else if (sum != 0.0) {
            ii = i;
        }
        *ELEM(B, LWB, 0, i) = sum;
    }
    for (int64_t i = width - 1; i >= 0; i--) {
        double sum = *ELEM(B, LWB, 0, i);
        const double* row = CELEM(LU, LWLU, 0, i);
        for (int64_t j = i + 1; j < width; j++) {
            double p = row[j] * *ELEM(B, LWB, 0, j);
            sum = sum - p;
        }
        *ELEM(B, LWB, 0, i) = sum / row[i];
    }
}

/* MatrixLinEQSolve (src/LinAlgLU.pas:481-592).  The residual walks X with
 * the column stride LWX; the reference walks it with inc(pX), which equals
 * this only for one right-hand side (LWX = 8) -- refinement parity unpinned. */
int32_t orc_lineq_solve(const double* A, int64_t LWA, int64_t width, const double* B, int64_t LWB,
                        double* X, int64_t LWX, int64_t width2, int64_t numRefinements)
{
    double* lu = (double*)malloc(sizeof(double) * width * width);
    int32_t* indx = (int32_t*)malloc(sizeof(int32_t) * width);
    for (int64_t y = 0; y < width; y++)
        memcpy(lu + y * width, CELEM(A, LWA, 0, y), sizeof(double) * width);
    int32_t r = orc_lu_decomp_inplace(lu, width * 8, width, indx);
    if (r != ORC_LE_OK) { free(lu); free(indx); return r; }
    for (int64_t c = 0; c < width2; c++) {
        for (int64_t j = 0; j < width; j++) *ELEM(X, LWX, c, j) = *CELEM(B, LWB, c, j);
        orc_lu_backsubst(lu, width * 8, width, indx, X + c, LWX);
    }
    if (numRefinements > 0) {
        double* row = (double*)malloc(sizeof(double) * width);
        for (int64_t c = 0; c < width2; c++) {
            for (int64_t it = 0; it < numRefinements; it++) {
                for (int64_t i = 0; i < width; i++) {
                    const double* a = CELEM(A, LWA, 0, i);
                    double sdp = -*CELEM(B, LWB, c, i);
                    for (int64_t j = 0; j < width; j++) {
                        double p = *ELEM(X, LWX, c, j) * a[j];
                        sdp = sdp + p;
                    }
                    row[i] = sdp;
                }
                orc_lu_backsubst(lu, width * 8, width, indx, row, 8);
                for (int64_t i = 0; i < width; i++) *ELEM(X, LWX, c, i) -= row[i];
            }
        }
        free(row);
    }
    free(lu);
    free(indx);
    return ORC_LE_OK;
}

/* MatrixInverseInPlace (src/LinAlgLU.pas:388-431) */
int32_t orc_inverse_inplace(double* A, int64_t LWA, int64_t width)
{
    int64_t w = width + (width & 1);
    double* Y = (double*)calloc((size_t)(w * w), sizeof(double));
    int32_t* indx = (int32_t*)malloc(sizeof(int32_t) * width);
    double* col = (double*)malloc(sizeof(double) * width);
    for (int64_t y = 0; y < width; y++) memcpy(Y + y * w, CELEM(A, LWA, 0, y), sizeof(double) * width);
    int32_t r = orc_lu_decomp_inplace(Y, w * 8, width, indx);
    if (r == ORC_LE_OK) {
        for (int64_t j = 0; j < width; j++) {
            for (int64_t i = 0; i < width; i++) col[i] = 0.0;
            col[j] = 1.0;
            orc_lu_backsubst(Y, w * 8, width, indx, col, 8);
            for (int64_t i = 0; i < width; i++) *ELEM(A, LWA, j, i) = col[i];
        }
    }
    free(Y); free(indx); free(col);
    return r;
}

/* MatrixDeterminant (src/LinAlgLU.pas:433-479) */
double orc_determinant(const double* A, int64_t LWA, int64_t width)
{
    int64_t w = width + (width & 1);
    double* lu = (double*)calloc((size_t)(w * w), sizeof(double));
    int32_t* indx = (int32_t*)calloc((size_t)width, sizeof(int32_t));
    for (int64_t y = 0; y < width; y++) memcpy(lu + y * w, CELEM(A, LWA, 0, y), sizeof(double) * width);
    double res;
    if (lu_rec(lu, width, width, indx, w * 8) != ORC_LE_OK) {
        res = 0.0;
    } else {
        res = 1.0;
        for (int64_t i = 0; i < width; i++) if (indx[i] != i) res = -res;
        for (int64_t i = 0; i < width; i++) res = res * lu[i * w + i];
    }
    free(lu); free(indx);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Error metrics                                                             */
/* ------------------------------------------------------------------------ */
static long double frob2(const double* M, int64_t LW, int64_t w, int64_t h)
{
    long double s = 0.0L;
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) { long double v = *CELEM(M, LW, x, y); s += v * v; }
    return s;
}

double orc_gemm_normwise_err(const double* C, int64_t LWC, const double* Cref, int64_t LWR,
                             int64_t wC, int64_t hC, const double* A, int64_t LWA, int64_t wA, int64_t hA,
                             const double* B, int64_t LWB, int64_t wB, int64_t hB)
{
    long double d = 0.0L;
    for (int64_t y = 0; y < hC; y++)
        for (int64_t x = 0; x < wC; x++) {
            long double v = (long double)*CELEM(C, LWC, x, y) - (long double)*CELEM(Cref, LWR, x, y);
            d += v * v;
        }
    long double na = sqrtl(frob2(A, LWA, wA, hA)), nb = sqrtl(frob2(B, LWB, wB, hB));
    if (na == 0.0L || nb == 0.0L) return (double)sqrtl(d);
    return (double)(sqrtl(d) / (na * nb));
}

double orc_backward_error(const double* A, int64_t LWA, int64_t n, const double* X, int64_t LWX,
                          const double* B, int64_t LWB, int64_t nrhs)
{
    long double anorm = 0.0L;
    for (int64_t i = 0; i < n; i++) {
        long double s = 0.0L;
        for (int64_t j = 0; j < n; j++) s += fabsl((long double)*CELEM(A, LWA, j, i));
        if (s > anorm) anorm = s;
    }
    double worst = 0.0;
    long double* r = (long double*)malloc(sizeof(long double) * n);
    for (int64_t c = 0; c < nrhs; c++) {
        long double xn = 0.0L, bn = 0.0L, rn = 0.0L;
        for (int64_t i = 0; i < n; i++) r[i] = (long double)*CELEM(B, LWB, c, i);
        for (int64_t i = 0; i < n; i++) {
            const double* a = CELEM(A, LWA, 0, i);
            long double s = r[i];
            for (int64_t j = 0; j < n; j++) s -= (long double)a[j] * (long double)*CELEM(X, LWX, c, j);
            r[i] = s;
        }
        for (int64_t i = 0; i < n; i++) {
            long double xv = fabsl((long double)*CELEM(X, LWX, c, i));
            long double bv = fabsl((long double)*CELEM(B, LWB, c, i));
            if (xv > xn) xn = xv;
            if (bv > bn) bn = bv;
            if (fabsl(r[i]) > rn) rn = fabsl(r[i]);
        }
        long double den = anorm * xn + bn;
        double e = den > 0.0L ? (double)(rn / den) : (double)rn;
        if (e > worst) worst = e;
    }
    free(r);
    return worst;
}

/* Covariance (src/Statistics.pas:599-639): column means by a sequential sum
 * over the rows divided by the height (GenericMtxMean, column-wise,
 * src/SimpleMatrixOperations.pas:432-445); ac = A - mean (MatrixSubVec);
 * dest = ac^T * ac as the generic product of the transposed copy
 * (MatrixTranspose + MatrixMult, :626-632); then dest = dest * (1/m) + 0
 * (MatrixScaleAndAdd, :633; GenericMtxScaleAndAdd,
 * src/SimpleMatrixOperations.pas:1220-1232), m = height - 1 when unbiased.
 * A single row or column gives MatrixVar's value in dest[0]
 * (GenericMtxVar, src/SimpleMatrixOperations.pas:450-520). */
void orc_covariance(double* dest, int64_t LWD, const double* A, int64_t LWA, int64_t width, int64_t height,
                    int32_t unbiased)
{
    if (height == 1 || width == 1) {
        const int64_t n = width * height;
        double val = 0.0;
        for (int64_t i = 0; i < n; i++) val = val + (height == 1 ? *CELEM(A, LWA, i, 0) : *CELEM(A, LWA, 0, i));
        const double mean = val / (double)n;
        double v = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const double d = (height == 1 ? *CELEM(A, LWA, i, 0) : *CELEM(A, LWA, 0, i)) - mean;
            v = v + d * d;
        }
        dest[0] = unbiased ? v / (double)(n - 1 > 1 ? n - 1 : 1) : v / (double)n;
        return;
    }
    double* mean = (double*)malloc(sizeof(double) * width);
    double* ac = (double*)malloc(sizeof(double) * width * height);
    double* tmp = (double*)malloc(sizeof(double) * width * height);
    for (int64_t x = 0; x < width; x++) {
        double val = 0.0;
        for (int64_t y = 0; y < height; y++) val = val + *CELEM(A, LWA, x, y);
        mean[x] = val / (double)height;
    }
    for (int64_t y = 0; y < height; y++)
        for (int64_t x = 0; x < width; x++) ac[y * width + x] = *CELEM(A, LWA, x, y) - mean[x];
    /* tmp (width x height) = ac^T */
    for (int64_t y = 0; y < height; y++)
        for (int64_t x = 0; x < width; x++) tmp[x * height + y] = ac[y * width + x];
    orc_generic_mtx_mult(dest, LWD, tmp, ac, height, width, width, height, height * 8, width * 8);
    const int64_t m = unbiased ? height - 1 : height;
    const double scale = 1.0 / (double)m;
    for (int64_t y = 0; y < width; y++)
        for (int64_t x = 0; x < width; x++) *ELEM(dest, LWD, x, y) = (*ELEM(dest, LWD, x, y) * scale) + 0.0;
    free(mean); free(ac); free(tmp);
}
