/*
 * mrm_cpu_fast.c -- CPU restatement of mrMath's FAST FP64 GEMM path, used as
 * the CPU timing baseline.
 *
 * TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and tests).  Built
 * with -O3 -mavx2 -mfma.  Restates, from the algorithm and not from any
 * instruction bytes:
 *   - the blocked driver BlockMatrixMultiplication (src/BlockedMult.pas:350-474):
 *     per (row-block, col-block) tile, per K-block transpose the B tile, run the
 *     transposed product, add into the tile accumulator, then apply op;
 *   - the FMA transposed micro-kernel FMAMatrixMultTransposed
 *     (src/FMAMatrixOperations.pas:55-92; algorithm of
 *     src/FMAMatrixMultTransposedOperationsx64.pas:40-188): two outputs at a time,
 *     four 4-lane FMA accumulators per output stepping 16 doubles of K, lanes
 *     reduced at the end;
 *   - the threaded tiling ThrMatrixMultEx2 and its core heuristic
 *     NumCoresToUseForMult (src/ThreadedMatrixOperations.pas:486-503, 612-718).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ELEM(p, lw, x, y) ((double*)((char*)(p) + (int64_t)(y) * (lw)) + (x))
#define CELEM(p, lw, x, y) ((const double*)((const char*)(p) + (int64_t)(y) * (lw)) + (x))

static double hsum4(__m256d v)
{
    __m128d lo = _mm256_castpd256_pd128(v), hi = _mm256_extractf128_pd(v, 1);
    __m128d s = _mm_add_pd(lo, hi);
    return _mm_cvtsd_f64(_mm_add_sd(s, _mm_unpackhi_pd(s, s)));
}

/* dest(h x w) = A(h x k) * Bt(w x k)^T; both operands K-contiguous. */
static void mult_transposed(double* dest, int64_t destLW, const double* A, int64_t LWA,
                            const double* Bt, int64_t LWB, int64_t k, int64_t h, int64_t w)
{
    int64_t k16 = k & ~(int64_t)15;
    for (int64_t y = 0; y < h; y++) {
        const double* a = CELEM(A, LWA, 0, y);
        int64_t x = 0;
        for (; x + 1 < w; x += 2) {
            const double* b0 = CELEM(Bt, LWB, 0, x);
            const double* b1 = CELEM(Bt, LWB, 0, x + 1);
            __m256d c00 = _mm256_setzero_pd(), c01 = _mm256_setzero_pd(), c02 = _mm256_setzero_pd(), c03 = _mm256_setzero_pd();
            __m256d c10 = _mm256_setzero_pd(), c11 = _mm256_setzero_pd(), c12 = _mm256_setzero_pd(), c13 = _mm256_setzero_pd();
            for (int64_t i = 0; i < k16; i += 16) {
                __m256d a0 = _mm256_loadu_pd(a + i), a1 = _mm256_loadu_pd(a + i + 4);
                __m256d a2 = _mm256_loadu_pd(a + i + 8), a3 = _mm256_loadu_pd(a + i + 12);
                c00 = _mm256_fmadd_pd(a0, _mm256_loadu_pd(b0 + i), c00);
                c01 = _mm256_fmadd_pd(a1, _mm256_loadu_pd(b0 + i + 4), c01);
                c02 = _mm256_fmadd_pd(a2, _mm256_loadu_pd(b0 + i + 8), c02);
                c03 = _mm256_fmadd_pd(a3, _mm256_loadu_pd(b0 + i + 12), c03);
                c10 = _mm256_fmadd_pd(a0, _mm256_loadu_pd(b1 + i), c10);
                c11 = _mm256_fmadd_pd(a1, _mm256_loadu_pd(b1 + i + 4), c11);
                c12 = _mm256_fmadd_pd(a2, _mm256_loadu_pd(b1 + i + 8), c12);
                c13 = _mm256_fmadd_pd(a3, _mm256_loadu_pd(b1 + i + 12), c13);
            }
            double s0 = hsum4(_mm256_add_pd(_mm256_add_pd(c00, c01), _mm256_add_pd(c02, c03)));
            double s1 = hsum4(_mm256_add_pd(_mm256_add_pd(c10, c11), _mm256_add_pd(c12, c13)));
            for (int64_t i = k16; i < k; i++) { s0 += a[i] * b0[i]; s1 += a[i] * b1[i]; }
            *ELEM(dest, destLW, x, y) = s0;
            *ELEM(dest, destLW, x + 1, y) = s1;
        }
        for (; x < w; x++) {
            const double* b0 = CELEM(Bt, LWB, 0, x);
            double s = 0.0;
            for (int64_t i = 0; i < k; i++) s += a[i] * b0[i];
            *ELEM(dest, destLW, x, y) = s;
        }
    }
}

/* BlockMatrixMultiplication, op 0/1/2 = none/add/sub. */
static void blocked_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                         int64_t w1, int64_t h1, int64_t w2, int64_t LW1, int64_t LW2, int64_t bs, int op)
{
    double* act = (double*)aligned_alloc(64, sizeof(double) * bs * bs * 3);
    double* mul = act + bs * bs;
    double* trn = mul + bs * bs;
    for (int64_t by = 0; by < h1; by += bs) {
        int64_t bh = h1 - by < bs ? h1 - by : bs;
        for (int64_t bx = 0; bx < w2; bx += bs) {
            int64_t bw = w2 - bx < bs ? w2 - bx : bs;
            memset(act, 0, sizeof(double) * bs * bs);
            for (int64_t bk = 0; bk < w1; bk += bs) {
                int64_t gw = w1 - bk < bs ? w1 - bk : bs;
                for (int64_t k = 0; k < gw; k++) {            /* MatrixTranspose of the B tile */
                    const double* br = CELEM(mt2, LW2, bx, bk + k);
                    for (int64_t x = 0; x < bw; x++) trn[x * bs + k] = br[x];
                }
                mult_transposed(mul, bs * 8, CELEM(mt1, LW1, bk, by), LW1, trn, bs * 8, gw, bh, bw);
                for (int64_t y = 0; y < bh; y++)
                    for (int64_t x = 0; x < bw; x++) act[y * bs + x] += mul[y * bs + x];
            }
            for (int64_t y = 0; y < bh; y++) {
                double* d = ELEM(dest, destLW, bx, by + y);
                for (int64_t x = 0; x < bw; x++) {
                    if (op == 0) d[x] = act[y * bs + x];
                    else if (op == 1) d[x] += act[y * bs + x];
                    else d[x] -= act[y * bs + x];
                }
            }
        }
    }
    free(act);
}

typedef struct {
    double* dest; int64_t destLW; const double* mt1; const double* mt2;
    int64_t w1, h1, w2, LW1, LW2, bs; int op;
} tile_job;

static void* run_job(void* p)
{
    tile_job* j = (tile_job*)p;
    blocked_mult(j->dest, j->destLW, j->mt1, j->mt2, j->w1, j->h1, j->w2, j->LW1, j->LW2, j->bs, j->op);
    return NULL;
}

/* NumCoresToUseForMult (src/ThreadedMatrixOperations.pas:486-503) */
int orc_num_cores_for_mult(int64_t w1, int64_t h1, int64_t w2, int64_t bs, int numUseCPUCores)
{
    if (w1 <= bs && w2 <= bs) {
        int r = 1;
        if (h1 > bs) { int64_t c = 1 + h1 / bs; r = (int)(c < numUseCPUCores ? c : numUseCPUCores); }
        return r;
    }
    int64_t nb = h1 / bs + w2 / bs;
    int64_t r = nb < numUseCPUCores ? nb : numUseCPUCores;
    return (int)(r < 1 ? 1 : r);
}

/* ThrMatrixMultEx2 (src/ThreadedMatrixOperations.pas:612-718): 2-D grid of
 * output tiles, one blocked multiply per thread.  Returns the threads used. */
int orc_fast_mult(double* dest, int64_t destLW, const double* mt1, const double* mt2,
                  int64_t w1, int64_t h1, int64_t w2, int64_t LW1, int64_t LW2,
                  int64_t bs, int op, int numUseCPUCores)
{
    if (numUseCPUCores > 64) numUseCPUCores = 64;      /* cMaxNumCores */
    int used = orc_num_cores_for_mult(w1, h1, w2, bs, numUseCPUCores);
    if (used <= 1) {
        blocked_mult(dest, destLW, mt1, mt2, w1, h1, w2, LW1, LW2, bs, op);
        return 1;
    }
    int64_t cpud2 = used / 2, thrWidth = w2, thrHeight, heightCores;
    if (cpud2 > 1) {
        heightCores = cpud2;
        cpud2 = used / heightCores;
        thrWidth = w2 / cpud2 + ((w2 % cpud2) != 0);
        thrWidth += thrWidth & 1;
        thrHeight = h1 / heightCores + ((h1 % heightCores) != 0);
    } else {
        cpud2 = 1;
        thrHeight = h1 / used + ((h1 % used) != 0);
        thrHeight += thrHeight & 1;
        heightCores = used;
    }
    tile_job jobs[64];
    pthread_t th[64];
    int n = 0;
    for (int64_t j = 0; j < cpud2; j++)
        for (int64_t i = 0; i < heightCores; i++) {
            int64_t hh = h1 > (i + 1) * thrHeight ? thrHeight : h1 - i * thrHeight;
            int64_t ww = w2 > (j + 1) * thrWidth ? thrWidth : w2 - j * thrWidth;
            if (hh <= 0 || ww <= 0) break;
            tile_job* t = &jobs[n++];
            t->dest = ELEM(dest, destLW, j * thrWidth, i * thrHeight);
            t->destLW = destLW;
            t->mt1 = CELEM(mt1, LW1, 0, i * thrHeight);
            t->mt2 = mt2 + j * thrWidth;
            t->w1 = w1; t->h1 = hh; t->w2 = ww; t->LW1 = LW1; t->LW2 = LW2; t->bs = bs; t->op = op;
        }
    for (int i = 0; i < n - 1; i++) pthread_create(&th[i], NULL, run_job, &jobs[i]);
    run_job(&jobs[n - 1]);
    for (int i = 0; i < n - 1; i++) pthread_join(th[i], NULL);
    return n;
}
