// mrm_lu.h -- internal interface of the LU / triangular-solve kernels.
#pragma once

#include "mrm_common.h"

// Status a panel kernel reports when a wait on its peer workgroups' exchange
// words passed the deadline (the workgroups were not all resident at once).
#define MRM_LU_TIMEOUT 100

namespace mrm {

constexpr int LU_NB = 256;             // outer block: K of the trailing-update GEMM (FP64)
constexpr int LU_NB_LOWP = 256;        // outer block of the low-precision trailing update
inline int lu_outer_block(int mode) { return mode == MRM_SOLVE_FP64 ? LU_NB : LU_NB_LOWP; }
constexpr int LU_W = 32;               // panel width of the panel kernel
constexpr int LU_TRSM_R = 64;          // rows per diagonal block of the blocked triangular solves
constexpr int LU_MAX_PANEL_WG = 128;   // panel workgroups (all resident; candidates fit in LDS)
constexpr int LU_PANEL_WG_TARGET = 32; // preferred panel workgroups before packing more rows per thread
// Lookahead: from this order on, the panel stream factors block k + 1 on its
// own CUs while the update stream applies block k to the columns beyond it
// (after applying it to block k + 1 first).
constexpr int LU_LA_MIN_N = 6144;
constexpr int LU_LA_MAX_N = 32768;      // up to here a panel has at most 32 workgroups, one per panel CU
constexpr int LU_LA_CUS_PER_XCD = 4;   // panel stream: 4 of the 32 CUs of each of the 8 XCDs

// Device workspace of the factorisation (one per library instance; users are
// serialised by the host mutex).
struct LuWork {
    int* status = nullptr;        // MRM_OK / MRM_SINGULAR / MRM_LU_TIMEOUT
    void* xch = nullptr;          // pivot exchange area
    void* lowp = nullptr;         // packed operands of the low-precision trailing update
    size_t lowp_bytes = 0;
    double* backup = nullptr;     // copy of an in-place caller's matrix for a relaunch
    size_t backup_elems = 0;
    // panel launch: 0 = plain launch when the device holds the whole grid,
    // 1 = always cooperative (the relaunch after a timeout)
    int coop = 0;
    // deadline of the panel waits on a plain launch in 100 MHz wall-clock
    // ticks; 0 makes every wait give up at once (a host-forced timeout)
    uint64_t plain_timeout = 0;
    // CUs the panel kernel may use (0: the whole device); bounds the grid of
    // a plain launch
    int panel_cus = 0;
};

constexpr uint64_t LU_PANEL_TIMEOUT_TICKS = 2ull * 100000000ull;  // 2 s

// Factors A in place like lu_factor, then synchronises s and reads the
// status.  A panel timeout (its workgroups were not co-resident) is retried
// once with cooperative panel launches after restore() has put the original
// matrix back into A.  Returns MRM_OK, MRM_SINGULAR, MRM_TIMED_OUT or an error.
template <typename Restore>
int lu_factor_sync(double* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s, int mode,
                   Restore restore);
// Host-selected deadline of plain panel launches (MRM_LU_TIMEOUT test hook)
// and the count of cooperative relaunches so far.
uint64_t lu_plain_panel_timeout();
void lu_count_relaunch();

size_t lu_xch_bytes();
// The lookahead's two streams: the panel stream (LU_LA_CUS_PER_XCD CUs of
// every XCD) and the update stream (the other CUs).  Created at library
// initialisation, before other work of the process can have taken the
// hardware queues (with queues created after a hipGraph with parallel
// branches had run, the two streams did not overlap: n = 16384 took 186 ms
// against 136 ms).  Returns 1 when they are ready, -1 when the device has no
// 8 x 32 CUs.  lu_lookahead_release destroys them (library shutdown).
int lu_lookahead_init();
void lu_lookahead_release();
// The streams exist and belong to the current device.
bool lu_lookahead_usable();

// Element-stride GEMM in the arithmetic type T (double: FP64 MFMA).
template <typename T>
int gemm_t(int trans, int64_t M, int64_t N, int64_t K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
           int64_t ldc, int op, hipStream_t s);

template <>
inline int gemm_t<double>(int trans, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B,
                          int64_t ldb, double* C, int64_t ldc, int op, hipStream_t s) {
    return mrm_launch_dgemm(trans, M, N, K, A, lda * 8, B, ldb * 8, C, ldc * 8, op, s);
}

// In-place LU with partial pivoting of the n x n matrix A (element stride lda).
// indx: reference pivot format; perm (optional): row i of P*A is row perm[i] of A.
// mode MRM_SOLVE_FP32_IR / MRM_SOLVE_BF16_IR computes the outer trailing
// updates in that precision (ws.lowp must hold lu_lowp_bytes(n, mode)).
template <typename T>
int lu_factor(T* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s, int mode = MRM_SOLVE_FP64);

// Factors the m x w panel at P (w <= LU_W) with partial pivoting in one
// launch whose workgroups exchange candidates through flagged words (plain
// launch, or cooperative when ws.coop is set or the device cannot hold the
// grid); pivots go to indx[piv_base + j] as absolute rows.
template <typename T>
int lu_panel(T* P, int64_t lda, int m, int w, int piv_base, int* indx, int* perm, LuWork& ws, hipStream_t s);

// Packed-operand bytes of the low-precision trailing updates of an n x n LU.
size_t lu_lowp_bytes(int n, int mode);
// Grows ws.lowp to lu_lowp_bytes(n, mode); caller holds mrm_host_mutex().
int lu_lowp_reserve(LuWork& ws, int n, int mode);

// X = (LU)^-1 P B for nrhs columns.  dinv (optional): the diagonal-block
// inverses from lu_solve_inverses for these factors, reused across solves.
template <typename T>
int lu_solve(const T* LU, int64_t lda, int n, const int* perm, const T* B, int64_t ldb, T* X, int64_t ldx, int nrhs,
             hipStream_t s, const double* dinv = nullptr);
// Doubles lu_solve_inverses writes for an n x n factorisation (0 when the
// solve needs none), and the writer.
size_t lu_solve_inverse_elems(int n);
int lu_solve_inverses(const double* LU, int64_t lda, int n, double* dinv, hipStream_t s);

template <typename T>
void add_inplace(T* X, int64_t ldx, const T* D, int64_t ldd, int rows, int cols, hipStream_t s);
void zero_i32(int* p, int n, hipStream_t s);

// Library-owned LU workspace (allocated on first use); caller holds mrm_host_mutex().
int lu_workspace(LuWork** out);

template <typename Restore>
int lu_factor_sync(double* A, int64_t lda, int n, int* indx, int* perm, LuWork& ws, hipStream_t s, int mode,
                   Restore restore) {
    for (int attempt = 0;; ++attempt) {
        ws.coop = attempt;
        zero_i32(ws.status, 1, s);
        int rc = lu_factor<double>(A, lda, n, indx, perm, ws, s, mode);
        ws.coop = 0;
        if (rc) return rc;
        int st = 0;
        MRM_HIP_TRY(hipMemcpyAsync(&st, ws.status, sizeof(int), hipMemcpyDeviceToHost, s));
        MRM_HIP_TRY(hipStreamSynchronize(s));
        if (st == MRM_OK) return MRM_OK;
        if (st == MRM_SINGULAR) return MRM_SINGULAR;
        if (st != MRM_LU_TIMEOUT || attempt > 0) {
            mrm_set_error(MRM_TIMED_OUT,
                          "lu: the panel workgroups' candidate exchange passed its deadline%s",
                          attempt ? " also with a cooperative launch" : "");
            return MRM_TIMED_OUT;
        }
        // the panel's workgroups were not resident together: start again from
        // the original matrix with cooperative panel launches
        lu_count_relaunch();
        if ((rc = restore())) return rc;
    }
}

}  // namespace mrm
