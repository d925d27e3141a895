/*
 * mrm_cpu_lu.c -- CPU restatement of mrMath's THREADED LU factorisation and
 * linear solve, the CPU timing baseline of the LU half of the metric.
 *
 * TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and tests).  Restates
 * (src/LinAlgLU.pas):
 *   - InternalThrMatrixLUDecomp (:674-786): the recursive LU of
 *     InternalRecursiveMatrixLUDecompInPlace with two threaded parts: the
 *     unit-lower solve A12 <- L11^-1 A12 split by columns over the threads
 *     (ThrMatrixUSubst, :627-672, for nRight > 64), and the Schur update
 *     A22 -= A21 A12 through the threaded blocked multiply with block size 48
 *     (ThrMatrixMultEx, when nRight or the height below exceeds 48; here the
 *     ThrMatrixMultEx2 tiling of mrm_cpu_fast.c with its AVX2/FMA kernel,
 *     also for the 512..1024 sizes the reference gives to its inner-block
 *     variant); smaller updates form the product and subtract it;
 *   - ThrMatrixLUDecomp (:788-803) and ThrMatrixLinEQSolve (:986-1128): copy,
 *     factor, per-right-hand-side MatrixLUBackSubst with the columns split
 *     over the threads, then NumRefinments sequential FP64 steps per column.
 * The leaf (pivot search, reciprocal scaling, swap) and the substitutions are
 * scalar loops, as the reference's Pascal is: this file is built with
 * -fno-tree-vectorize; the multiply comes from mrm_cpu_fast.c.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ELEM(p, lw, x, y) ((double*)((char*)(p) + (int64_t)(y) * (lw)) + (x))
#define CELEM(p, lw, x, y) ((const double*)((const char*)(p) + (int64_t)(y) * (lw)) + (x))

int orc_fast_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2, int64_t w1, int64_t h1,
                  int64_t w2, int64_t LW1, int64_t LW2, int64_t bs, int op, int numUseCPUCores);
void orc_lu_backsubst(const double* LU, int64_t LWLU, int64_t width, const int32_t* indx, double* B, int64_t LWB);

enum { LE_OK = 0, LE_SINGULAR = 1 };
#define BLK_MULT_SIZE 48   /* cBlkMultSize, src/LinAlgLU.pas:171 */
#define MIN_THR_SIZE 64    /* cMinThrMultSize, :677 */
#define MAX_CORES 64       /* cMaxNumCores */

/* LUSwap (:120-136) */
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

/* LUBacksup (:138-169) on `width` columns of B */
static void lu_backsup(const double* A, int64_t width, int64_t height, double* B, int64_t LW)
{
    for (int64_t k = 0; k < height; k++) {
        const double* bk = CELEM(B, LW, 0, k);
        for (int64_t i = k + 1; i < height; i++) {
            const double l = *CELEM(A, LW, k, i);
            double* bi = ELEM(B, LW, 0, i);
            for (int64_t j = 0; j < width; j++) bi[j] = bi[j] - bk[j] * l;
        }
    }
}

typedef struct {
    const double* A; int64_t width, height; double* B; int64_t LW;
} usubst_job;

static void* usubst_run(void* p)
{
    usubst_job* j = (usubst_job*)p;
    lu_backsup(j->A, j->width, j->height, j->B, j->LW);
    return NULL;
}

/* ThrMatrixUSubst (:627-672): columns in chunks of max(32, ceil(width/cores)) */
static void thr_usubst(const double* A, int64_t width, int64_t height, double* B, int64_t LW, int cores)
{
    int64_t thrSize = width / cores + ((width % cores) != 0);
    if (thrSize < 32) thrSize = 32;
    usubst_job jobs[MAX_CORES];
    pthread_t th[MAX_CORES];
    int n = 0;
    for (int i = 0; i < cores; i++) {
        int64_t w = thrSize;
        if (w * (i + 1) >= width) w = width - i * thrSize;
        jobs[n].A = A; jobs[n].width = w; jobs[n].height = height; jobs[n].B = B + i * thrSize; jobs[n].LW = LW;
        n++;
        if (thrSize * (i + 1) >= width) break;
    }
    for (int i = 0; i < n - 1; i++) pthread_create(&th[i], NULL, usubst_run, &jobs[i]);
    usubst_run(&jobs[n - 1]);
    for (int i = 0; i < n - 1; i++) pthread_join(th[i], NULL);
}

static int32_t thr_lu_rec(double* A, int64_t width, int64_t height, int32_t* indx, int64_t LW, double* mem,
                          int cores)
{
    const int64_t mn = width < height ? width : height;
    if (mn > 1) {
        const int64_t nleft = mn / 2, nright = width - nleft;
        int32_t r = thr_lu_rec(A, nleft, height, indx, LW, mem, cores);
        if (r != LE_OK) return r;
        double* a12 = ELEM(A, LW, nleft, 0);
        lu_swap(a12, LW, nright, 0, nleft - 1, indx);
        if (nright > 1) {
            if (nright > MIN_THR_SIZE && cores > 1) thr_usubst(A, nright, nleft, a12, LW, cores);
            else lu_backsup(A, nright, nleft, a12, LW);
        }
        double* a22 = ELEM(A, LW, nleft, nleft);
        const double* a21 = CELEM(A, LW, 0, nleft);
        const int64_t mh = height - nleft;
        if (nright > BLK_MULT_SIZE || mh > BLK_MULT_SIZE) {
            orc_fast_mult(a22, LW, a21, a12, nleft, mh, nright, LW, LW, BLK_MULT_SIZE, 2, cores);
        } else if (mh > 0 && nright > 0) {
            /* MatrixMult into the scratch, then MatrixSub (:221-225) */
            const int64_t mlw = (nright + (nright & 1)) * 8;
            orc_fast_mult(mem, mlw, a21, a12, nleft, mh, nright, LW, LW, BLK_MULT_SIZE, 0, 1);
            for (int64_t y = 0; y < mh; y++) {
                double* c = ELEM(a22, LW, 0, y);
                const double* m = CELEM(mem, mlw, 0, y);
                for (int64_t x = 0; x < nright; x++) c[x] = c[x] - m[x];
            }
        }
        r = thr_lu_rec(a22, nright, height - nleft, indx + nleft, LW, mem, cores);
        if (r != LE_OK) return r;
        for (int64_t i = nleft; i < width; i++) indx[i] += (int32_t)nleft;
        lu_swap(A, LW, nleft, nleft, mn - 1, indx);
        return LE_OK;
    }
    /* leaf (:751-785): first max-abs pivot, reciprocal scale, swap */
    double maxVal = 0.0, absMax = 0.0;
    int64_t idx = -1;
    for (int64_t i = 0; i < height; i++) {
        const double v = *CELEM(A, LW, 0, i);
        if (fabs(v) > absMax) { idx = i; maxVal = v; absMax = fabs(v); }
    }
    indx[0] = (int32_t)idx;
    if (fabs(maxVal) > 10.0 / DBL_MAX) {
        const double inv = 1.0 / maxVal;
        for (int64_t i = 0; i < height; i++) {
            double* p = ELEM(A, LW, 0, i);
            *p = (*p * inv) + 0.0;
        }
        *ELEM(A, LW, 0, idx) = *A;
        *A = maxVal;
        return LE_OK;
    }
    return LE_SINGULAR;
}

/* ThrMatrixLUDecomp (:788-803) */
int32_t orc_thr_lu_decomp(double* A, int64_t LWA, int64_t width, int32_t* indx, int cores)
{
    if (cores < 1) cores = 1;
    if (cores > MAX_CORES) cores = MAX_CORES;
    double* mem = (double*)aligned_alloc(64, sizeof(double) * (BLK_MULT_SIZE + 2) * (BLK_MULT_SIZE + 2));
    memset(indx, 0, sizeof(int32_t) * width);
    const int32_t r = thr_lu_rec(A, width, width, indx, LWA, mem, cores);
    free(mem);
    return r;
}

typedef struct {
    const double* LU; int64_t LWLU, n; const int32_t* indx; double* X; int64_t LWX; int64_t c0, c1;
} backsub_job;

static void* backsub_run(void* p)
{
    backsub_job* j = (backsub_job*)p;
    for (int64_t c = j->c0; c < j->c1; c++) orc_lu_backsubst(j->LU, j->LWLU, j->n, j->indx, j->X + c, j->LWX);
    return NULL;
}

/* ThrMatrixLinEQSolve (:986-1128).  Returns leOk / leSingular. */
int32_t orc_thr_lineq_solve(const double* A, int64_t LWA, int64_t width, const double* B, int64_t LWB, double* X,
                            int64_t LWX, int64_t width2, int64_t numRefinements, int cores)
{
    if (cores < 1) cores = 1;
    if (cores > MAX_CORES) cores = MAX_CORES;
    const int64_t w = (width & 3) ? width + 4 - (width & 3) : width;
    const int64_t LULW = w * 8;
    double* LU = (double*)aligned_alloc(64, sizeof(double) * w * width);
    int32_t* indx = (int32_t*)malloc(sizeof(int32_t) * width);
    for (int64_t y = 0; y < width; y++) memcpy(ELEM(LU, LULW, 0, y), CELEM(A, LWA, 0, y), sizeof(double) * width);
    int32_t r = orc_thr_lu_decomp(LU, LULW, width, indx, cores);
    if (r != LE_OK) { free(LU); free(indx); return r; }
    for (int64_t y = 0; y < width; y++) memcpy(ELEM(X, LWX, 0, y), CELEM(B, LWB, 0, y), sizeof(double) * width2);
    /* columns over the threads: ceil(width2 / cores) each */
    int64_t thrSize = width2 / cores + ((width2 % cores) != 0);
    if (thrSize < 1) thrSize = 1;
    const int nthr = (int)(thrSize < cores ? thrSize : cores);
    backsub_job jobs[MAX_CORES];
    pthread_t th[MAX_CORES];
    int n = 0;
    for (int i = 0; i < nthr; i++) {
        const int64_t c0 = i * thrSize;
        int64_t c1 = (i == nthr - 1) ? width2 : c0 + thrSize;
        if (c1 > width2) c1 = width2;
        if (c0 >= c1) break;
        jobs[n] = (backsub_job){LU, LULW, width, indx, X, LWX, c0, c1};
        n++;
    }
    for (int i = 0; i < n - 1; i++) pthread_create(&th[i], NULL, backsub_run, &jobs[i]);
    if (n > 0) backsub_run(&jobs[n - 1]);
    for (int i = 0; i < n - 1; i++) pthread_join(th[i], NULL);
    if (numRefinements > 0) {
        double* row = (double*)malloc(sizeof(double) * width);
        for (int64_t c = 0; c < width2; c++) {
            for (int64_t it = 0; it < numRefinements; it++) {
                for (int64_t i = 0; i < width; i++) {
                    const double* a = CELEM(A, LWA, 0, i);
                    double sdp = -*CELEM(B, LWB, c, i);
                    for (int64_t j = 0; j < width; j++) sdp = sdp + *CELEM(X, LWX, c, j) * a[j];
                    row[i] = sdp;
                }
                orc_lu_backsubst(LU, LULW, width, indx, row, 8);
                for (int64_t i = 0; i < width; i++) *ELEM(X, LWX, c, i) -= row[i];
            }
        }
        free(row);
    }
    free(LU);
    free(indx);
    return LE_OK;
}
