/*
 * mrm_ozaki_oracle.c -- CPU restatement of the INT8-slice FP64 GEMM.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of mrm_dev_dgemm_ozaki; see
 * mrm_oracle.h for who may load it).
 *
 * The reference multiplies in FP64 only (src/SimpleMatrixOperations.pas:
 * 1561-1591 is its oracle GEMM); the split-precision method follows the
 * published Ozaki scheme with integer slices (Ozaki, Ogita, Oishi, Rump,
 * "Error-free transformations of matrix multiplication", Numer. Algorithms
 * 2012; Ootomo, Ozaki, Yokota, "DGEMM on integer matrix multiplication unit",
 * IJHPCA 2024):
 *   row scale 2^ea_i with max_k |A_ik| < 2^ea_i (0 for a zero row), column
 *   scale 2^eb_j likewise for B; digits d_t = trunc(128 y_t), y_{t+1} =
 *   128 y_t - d_t from y_0 = A_ik 2^-ea_i; for each K chunk and each diagonal
 *   d = 0..s-1, S_ij = sum_{t=0..d} sum_k DA_t[i][k] DB_(d-t)[j][k] (exact),
 *   then C_ij (op)= S_ij 2^(ea_i + eb_j - 7(d+2)) in FP64, the very first term
 *   applying op and every later one adding (subtracting for op = sub).
 * Every step up to the FP64 accumulation is exact, so the GPU result must equal
 * this one bit for bit.
 *
 * Pinning (tests/test_oracle.py): integer inputs whose digits the slices hold
 * completely give the exact product (equal to orc_generic_mtx_mult); the error
 * against the FP64 GEMM falls by about 2^-7 per added slice and stays under
 * the bound 2 K 2^(-7 s) max|A_i.| max|B_.j|; a single-slice product of inputs
 * that are multiples of 2^(e-7) equals the exact product; and a brute-force
 * numpy restatement of the digit split agrees on small random inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mrm_oracle.h"

#define ELEM(P, LW, x, y) ((const double*)((const char*)(P) + (int64_t)(y) * (LW)) + (x))
#define MELEM(P, LW, x, y) ((double*)((char*)(P) + (int64_t)(y) * (LW)) + (x))

static int scale_exp(double amax) {
    int e = 0;
    if (amax > 0.0) (void)frexp(amax, &e);
    return e;
}

static void digits(double x, int e, int s, int8_t* out, int64_t pitch) {
    double y = ldexp(x, -e);
    for (int t = 0; t < s; t++) {
        y *= 128.0;
        const double d = trunc(y);
        y -= d;
        out[t * pitch] = (int8_t)d;
    }
}

/* C (M x N) (op)= A (M x K) * B (K x N), row-major, strides in bytes.
 * Returns 0, or 1 if a chunk sum left the int32 range (the GPU's accumulator
 * would have wrapped), or 2 on bad arguments / no memory. */
int32_t orc_ozaki_gemm(double* C, int64_t LWC, const double* A, int64_t LWA, const double* B, int64_t LWB,
                       int64_t M, int64_t N, int64_t K, int32_t op, int32_t slices, int64_t kchunk) {
    if (slices < 1 || slices > 8 || kchunk < 1 || M < 0 || N < 0 || K < 0) return 2;
    if (M == 0 || N == 0) return 0;
    const int s = slices;
    int* ea = (int*)malloc(sizeof(int) * (size_t)M);
    int* eb = (int*)malloc(sizeof(int) * (size_t)N);
    int8_t* DA = (int8_t*)malloc((size_t)s * (size_t)M * (size_t)(K ? K : 1));
    int8_t* DB = (int8_t*)malloc((size_t)s * (size_t)N * (size_t)(K ? K : 1));
    if (!ea || !eb || !DA || !DB) {
        free(ea); free(eb); free(DA); free(DB);
        return 2;
    }
    for (int64_t i = 0; i < M; i++) {
        double m = 0.0;
        for (int64_t k = 0; k < K; k++) m = fmax(m, fabs(*ELEM(A, LWA, k, i)));
        ea[i] = scale_exp(m);
        for (int64_t k = 0; k < K; k++) digits(*ELEM(A, LWA, k, i), ea[i], s, DA + i * K + k, M * K);
    }
    for (int64_t j = 0; j < N; j++) {
        double m = 0.0;
        for (int64_t k = 0; k < K; k++) m = fmax(m, fabs(*ELEM(B, LWB, j, k)));
        eb[j] = scale_exp(m);
        for (int64_t k = 0; k < K; k++) digits(*ELEM(B, LWB, j, k), eb[j], s, DB + j * K + k, N * K);
    }
    int rc = 0;
    int first = 1;
    if (K == 0) {
        for (int64_t i = 0; i < M; i++)
            for (int64_t j = 0; j < N; j++)
                if (op == 0) *MELEM(C, LWC, j, i) = 0.0;
    }
    for (int64_t k0 = 0; k0 < K; k0 += kchunk) {
        const int64_t k1 = k0 + kchunk < K ? k0 + kchunk : K;
        for (int d = 0; d < s; d++) {
            const int o = first ? op : (op == 2 ? 2 : 1);
            first = 0;
            for (int64_t i = 0; i < M; i++) {
                for (int64_t j = 0; j < N; j++) {
                    int64_t acc = 0;
                    for (int t = 0; t <= d; t++) {
                        const int8_t* a = DA + (int64_t)t * M * K + i * K;
                        const int8_t* b = DB + (int64_t)(d - t) * N * K + j * K;
                        int32_t part = 0;
                        for (int64_t k = k0; k < k1; k++) part += (int32_t)a[k] * (int32_t)b[k];
                        acc += part;
                    }
                    if (acc > INT32_MAX || acc < INT32_MIN) rc = 1;
                    const double x = ldexp((double)acc, ea[i] + eb[j] - 7 * (d + 2));
                    double* c = MELEM(C, LWC, j, i);
                    if (o == 0) *c = x;
                    else if (o == 1) *c = *c + x;
                    else *c = *c - x;
                }
            }
        }
    }
    free(ea); free(eb); free(DA); free(DB);
    return rc;
}
