#!/usr/bin/env python3
"""Benchmark of mrMath's FP64 multiply / LU-solve hot path on one MI355X.

One step = one C = A*B on device-resident 8192 x 8192 float64 operands
(entries uniform [-1, 1)) through the library's device entry (the square case
of the GPU GEMM family).  Prints one JSON line; besides the headline GEMM it
carries the other measured rows of the path under "rows":

  dgemm_tn_4096 / dgemm_nt_4096   MultT1 / MultT2 at 4096^3
  syrk_200000x512                 X^T X with X 200000 x 512 (symmetric split-K GEMM),
                                  also counted against the SYRK minimum n (n + 1) k
  covariance_200000x512           Covariance of X 200000 x 512: means, centred copy,
                                  symmetric split-K product, 1/(m-1) scale
  ozaki_8192                      FP64 GEMM emulated with INT8 slices, s = 3..7,
                                  wide-range inputs (2 res53 - 1) 2^e, e in [-20, 20]
  batched                         8192 uneven FP64 GEMMs (m, n, k in [16, 256]) and
                                  20000 small solves (n in [8, 128]), hipGraph plans
  lu_ir_16384                     A X = B, n = 16384, 64 right-hand sides,
                                  A = U diag(s) V^T with kappa 1e2 / 1e6 / 1e10:
                                  plain FP64 LU vs FP32- and BF16-update LU with
                                  FP64 refinement to backward error 1e-14

  python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-rows]

--workload runs one workload alone (dgemm_nn_8192 default, dgemm_tn_4096,
dgemm_nt_4096, syrk_200000x512, covariance_200000x512, ozaki_8192, batched,
lu_ir_16384, and copy_1gib: a plain 1 GiB device copy, the calibration kernel
for the HBM byte counters).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak (AMD published)
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "dgemm_nn_8192": (0, 8192, 8192, 8192),
    "dgemm_tn_4096": (1, 4096, 4096, 4096),
    "dgemm_nt_4096": (2, 4096, 4096, 4096),
    "syrk_200000x512": (1, 512, 512, 200000),
}
INT8_PEAK_TOPS = 5000.0


def _traffic_from_profile(workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), recorded in profiles/."""
    f = ROOT / "profiles" / f"dgemm_mfma_{workload}_pmc.txt"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def host_cores() -> int:
    """CPUs this process may run on (the box's share), at most cMaxNumCores = 64."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(64, n))


def host_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return "unknown"


def cpu_baseline(n: int = 8192, budget_s: float = 20.0, lu_n: int = 4096, lu_rhs: int = 64):
    """The restated reference fast path on the host cores: the threaded
    blocked multiply (block 128, 4-lane FMA kernel, ThrMatrixMultEx2 tiling
    with NumCoresToUseForMult) on the headline n^3 product, and the restated
    ThrMatrixLinEQSolve on an lu_n system with lu_rhs right-hand sides.
    Threads: min(CPUs available, 64) as the reference's pool takes, limited
    by NumCoresToUseForMult for the multiply."""
    import numpy as np
    import oracle as orc

    threads = host_cores()
    A = orc.pm1_matrix(n, n, 8101)
    B = orc.pm1_matrix(n, n, 8102)
    C = np.zeros((n, n))
    used = orc.fast_mult(C, A, B, threads)  # warm-up
    times = []
    t_start = time.monotonic()
    while len(times) < 3 and (time.monotonic() - t_start) < budget_s:
        t0 = time.monotonic()
        used = orc.fast_mult(C, A, B, threads)
        times.append(time.monotonic() - t0)
    t = statistics.median(times)
    del A, B, C
    # one thread on a smaller product (the 8192^3 one would take about a minute)
    n1 = 2048
    A1 = orc.pm1_matrix(n1, n1, 8101)
    B1 = orc.pm1_matrix(n1, n1, 8102)
    C1 = np.zeros((n1, n1))
    orc.fast_mult(C1, A1, B1, 1)
    t1 = []
    for _ in range(3):
        t0 = time.monotonic()
        orc.fast_mult(C1, A1, B1, 1)
        t1.append(time.monotonic() - t0)
    one = {"n": n1, "threads": 1, "gflops": round(2.0 * n1 ** 3 / statistics.median(t1) / 1e9, 2),
           "sample": f"dgemm NN {n1}^3, seeds 8101/8102, median of 3 runs after 1 warm-up"}
    del A1, B1, C1
    # LU solve: one warm-up-free run (the factorisation dominates; its run
    # time is seconds), as the reference's ThrMatrixLinEQSolve would be timed
    Al = orc.pm1_matrix(lu_n, lu_n, 4401)
    Bl = orc.pm1_matrix(lu_rhs, lu_n, 4403)
    t0 = time.monotonic()
    rc, Xl = orc.thr_lineq_solve(Al, Bl, 0, threads)
    lu_s = time.monotonic() - t0
    lu_flops = 2.0 / 3.0 * lu_n ** 3 + 2.0 * lu_n ** 2 * lu_rhs
    return {
        "value": round(2.0 * n ** 3 / t / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": int(used),
        "kind": "port",
        "sample": f"dgemm NN {n}^3 (the headline product), U[-1,1) MT19937 seeds 8101/8102, median of "
                  f"{len(times)} runs after 1 warm-up, {used} threads of {threads} available (ThrMatrixMultEx2 "
                  f"tiling, block 128, AVX2/FMA kernel)",
        "host": {"cpu": host_model(), "cpus_available": threads, "nproc": os.cpu_count()},
        "one_thread": one,
        "lu_solve": {"n": lu_n, "rhs": lu_rhs, "ms": round(lu_s * 1e3, 1), "threads": threads, "rc": rc,
                     "gflops": round(lu_flops / lu_s / 1e9, 2),
                     "backward_error": orc.backward_error(Al, Xl, Bl),
                     "what": "restated ThrMatrixLinEQSolve (recursive LU, column-split LUBacksup, threaded "
                             "blocked Schur update with block 48, threaded MatrixLUBackSubst), one run"},
    }


def copy_calibration(args):
    """1 GiB device-to-device copy: reads and writes 2^30 bytes per dispatch."""
    import torch
    n = (1 << 30) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / args.steps
    print(json.dumps({"workload": "copy_1gib", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                      "ms_per_copy": round(ms, 4), "GBps": round(2 * (1 << 30) / ms / 1e6, 1)}), flush=True)


def gemm_run(lib, workload: str, steps: int, warmup: int):
    """Times the device GEMM of `workload` with HIP events around every launch
    on the stream it runs on; checks C x == A (B x) outside the timed region."""
    import torch
    from mrmath_amd import _lib

    trans, M, N, K = WORKLOADS[workload]
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(8101)
    rowsA, colsA = (K, M) if trans == 1 else (M, K)
    rowsB, colsB = (N, K) if trans == 2 else (K, N)
    A = torch.rand(rowsA, colsA, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    if workload.startswith("syrk"):
        B, colsB = A, colsA  # X^T X
    else:
        B = torch.rand(rowsB, colsB, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    C = torch.empty(M, N, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        rc = lib.mrm_dev_dgemm(trans, M, N, K, A.data_ptr(), colsA * 8, B.data_ptr(), colsB * 8,
                               C.data_ptr(), N * 8, 0, sp)
        if rc:
            _lib.check(rc)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s_, e_ in ev:
        s_.record(stream)
        step()
        e_.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s_.elapsed_time(e_) for s_, e_ in ev]

    x = torch.randn(N, dtype=torch.float64, device=dev, generator=g)
    Am = A.t() if trans == 1 else A
    Bm = B.t() if trans == 2 else B
    rel = (torch.linalg.vector_norm(C @ x - Am @ (Bm @ x)) /
           (torch.linalg.vector_norm(A) * torch.linalg.vector_norm(B) * torch.linalg.vector_norm(x))).item()
    if not rel <= 4 * K * 2.0 ** -53:
        print(json.dumps({"error": f"{workload}: GEMM result failed the linearity check", "rel": rel}),
              file=sys.stderr)
        sys.exit(1)
    flops = 2.0 * M * N * K
    avg_ms = statistics.mean(kernel_ms)
    return {
        "M": M, "N": N, "K": K, "op": ["NN", "TN (MultT1)", "NT (MultT2)"][trans],
        "ms_per_step": wall * 1e3 / steps,
        "tflops": flops / (wall / steps) / 1e12,
        "avg_launch_ms": avg_ms,
        "achieved_tflops": flops / (avg_ms * 1e-3) / 1e12,
        "linearity_rel_err": rel,
    }


def covariance_row(lib, h: int = 200000, w: int = 512, reps: int = 10):
    """Covariance (src/Statistics.pas:599-639) of X 200000 x 512 on the device
    entry: column means, centred copy, symmetric split product, scale; HIP
    events around each call, median."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(2001)
    X = torch.rand(h, w, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    Cv = torch.empty(w, w, dtype=torch.float64, device=dev)
    wb = lib.mrm_covariance_workspace_bytes(w, h)
    W = torch.empty(wb, dtype=torch.uint8, device=dev)
    sp = torch.cuda.current_stream().cuda_stream

    def call():
        rc = lib.mrm_dev_covariance(X.data_ptr(), w * 8, w, h, Cv.data_ptr(), w * 8, 1, W.data_ptr(), wb, sp)
        assert rc == 0, rc

    for _ in range(3):
        call()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        call()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ms = statistics.median(ts)
    # algorithmic traffic: X read twice (means, centring), the centred copy
    # written once and read once by the product, the w x w result
    traffic = 8.0 * (4 * h * w + w * w)
    syrk = w * (w + 1) * h
    Xc = X - X.mean(dim=0)
    ref = (Xc[:, :4].t() @ Xc) / (h - 1)
    err = float(torch.linalg.norm(Cv[:4] - ref) / (torch.linalg.norm(Xc[:, :4]) * torch.linalg.norm(Xc) / (h - 1)))
    return {"ms": round(ms, 4), "syrk_tflops": round(syrk / (ms * 1e-3) / 1e12, 3),
            "hbm_gbps_algorithmic": round(traffic / (ms * 1e-3) / 1e9, 1),
            "sampled_rows_normwise_err": err,
            "path": "mrm_dev_covariance: colsum/mean/centre kernels + symmetric split-K TN GEMM + scale"}


def ozaki_row(lib, reps: int = 3):
    """FP64 GEMM emulated with s = 3..7 INT8 slices (and s = 6, 7 BF16
    slices) at 8192^3 on wide-range inputs, against the native FP64 MFMA GEMM
    of the same operands (itself checked against the oracle in the tests)."""
    import torch

    n = 8192
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(3101)

    def wide():
        u = torch.rand(n, n, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        return u * torch.exp2(torch.randint(-20, 21, (n, n), device=dev, generator=g).double())

    A, B = wide(), wide()
    sp = torch.cuda.current_stream().cuda_stream
    Cd = torch.empty(n, n, dtype=torch.float64, device=dev)
    lib.mrm_dev_dgemm(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, Cd.data_ptr(), n * 8, 0, sp)
    na = float(torch.linalg.norm(A)) * float(torch.linalg.norm(B))
    wb = lib.mrm_ozaki_workspace_bytes(n, n, n, 7)
    W = torch.empty(wb, dtype=torch.uint8, device=dev)
    C = torch.empty_like(Cd)
    bar = 4 * n * 2.0 ** -53
    per_s = {}
    for s in range(3, 8):
        def call():
            rc = lib.mrm_dev_dgemm_ozaki(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, C.data_ptr(), n * 8,
                                         0, s, W.data_ptr(), wb, sp)
            assert rc == 0, rc
        call()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            call()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = statistics.median(ts)
        err = float(torch.linalg.norm(C - Cd)) / na
        per_s[str(s)] = {"ms": round(ms, 3), "fp64_equiv_tflops": round(2.0 * n ** 3 / ms / 1e9, 2),
                         "int8_tops": round(s * (s + 1) // 2 * 2.0 * n ** 3 / ms / 1e9, 1),
                         "normwise_err_vs_fp64_gemm": err, "meets_fp64_bar": err <= bar}
    ok = [s for s in per_s if per_s[s]["meets_fp64_bar"]]
    best = max(ok, key=lambda s: per_s[s]["fp64_equiv_tflops"]) if ok else None
    # the same GEMMs from BF16 digits (v_mfma_f32_32x32x16_bf16, exact FP32
    # sums over 128 k): bit-equal to the INT8 result, at half the INT8 rate
    wbb = lib.mrm_ozaki_workspace_bytes_ex(n, n, n, 7, 1)
    del W
    W = torch.empty(wbb, dtype=torch.uint8, device=dev)
    C8 = torch.empty_like(Cd)
    per_bf = {}
    for s in (6, 7):
        assert lib.mrm_dev_dgemm_ozaki(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, C8.data_ptr(), n * 8,
                                       0, s, W.data_ptr(), wbb, sp) == 0

        def call_bf():
            rc = lib.mrm_dev_dgemm_ozaki_ex(0, n, n, n, A.data_ptr(), n * 8, B.data_ptr(), n * 8, C.data_ptr(),
                                            n * 8, 0, s, 1, W.data_ptr(), wbb, sp)
            assert rc == 0, rc
        call_bf()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            call_bf()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = statistics.median(ts)
        per_bf[str(s)] = {"ms": round(ms, 3), "fp64_equiv_tflops": round(2.0 * n ** 3 / ms / 1e9, 2),
                          "bf16_tflops": round(s * (s + 1) // 2 * 2.0 * n ** 3 / ms / 1e9, 1),
                          "normwise_err_vs_fp64_gemm": float(torch.linalg.norm(C - Cd)) / na,
                          "bit_equal_to_int8_slices": bool(torch.equal(C, C8))}
    return {
        "config": "8192^3, entries (2 res53 - 1) 2^e, e uniform in [-20, 20]",
        "bar": f"||C - C_fp64||_F / (||A||_F ||B||_F) <= 4 k 2^-53 = {bar:.3e}",
        "slices": per_s,
        "fastest_meeting_bar": {"slices": int(best), **per_s[best],
                                "pct_of_fp64_peak": round(100 * per_s[best]["fp64_equiv_tflops"] /
                                                          FP64_PEAK_TFLOPS, 1)} if best else None,
        "kernel": "oz_fused_kernel (v_mfma_i32_32x32x32_i8), one launch per group of up to four slice diagonals",
        "bf16_slices": per_bf,
    }


def lu_ir_row(lib, n: int = 16384, nrhs: int = 64):
    """A X = B with A = U diag(s) V^T (U, V Householder-QR factors of Gaussian
    matrices, s log-spaced to kappa), B Gaussian: plain FP64 LU + solve against
    LU with FP32 / BF16 trailing updates and FP64 refinement (plain steps, then
    GMRES corrections preconditioned by the factors if those stall), all to a
    normwise backward error <= 1e-14.  Times are wall clock around the synchronous
    device entry (factorisation, solves, residuals, norms): one untimed call
    per mode first (it grows the library's work buffers), then the median of
    3 calls.  The backward error is also recomputed from an independent
    residual (torch's FP64 matmul).  At kappa 1e2 also the FP64 factorisation
    alone (mrm_dev_lu_factor on a copy of A, median of 3 after one untimed
    call), with whether it ran the lookahead on CU-masked streams."""
    import ctypes as C
    import torch

    from mrmath_amd import _lib

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(4401)
    U, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    B = torch.randn(n, nrhs, dtype=torch.float64, device=dev, generator=g)
    sp = torch.cuda.current_stream().cuda_stream
    out = {"config": f"n={n}, {nrhs} right-hand sides, tol 1e-14", "kappa": {}}
    for kappa in (1e2, 1e6, 1e10):
        s = torch.logspace(0, -float(np.log10(kappa)), n, dtype=torch.float64, device=dev)
        A = (U * s) @ V.t()
        torch.cuda.synchronize()
        res = {}
        for mode, name in ((0, "fp64"), (1, "fp32_ir"), (2, "bf16_ir"), (3, "bf16x3_ir")):
            X = torch.zeros(n, nrhs, dtype=torch.float64, device=dev)
            st = _lib.SolveStats()
            ts = []
            for rep in range(4):
                torch.cuda.synchronize()
                t = time.perf_counter()
                rc = lib.mrm_dev_lineq_solve_ir_stats(mode, A.data_ptr(), n * 8, n, B.data_ptr(), nrhs * 8,
                                                      X.data_ptr(), nrhs * 8, nrhs, 50, 1e-14, C.byref(st), sp)
                torch.cuda.synchronize()
                if rep > 0:
                    ts.append(time.perf_counter() - t)
            R = B - A @ X
            berr = float((R.abs().amax(dim=0) / (A.abs().sum(dim=1).amax() * X.abs().amax(dim=0) +
                                                 B.abs().amax(dim=0))).max())
            del R
            res[name] = {"rc": rc, "ms": round(1e3 * statistics.median(ts), 1), "refinement_steps": st.iterations,
                         "gmres_ir": bool(st.gmres), "krylov_steps": st.krylov_steps,
                         "backward_error": st.backward_error, "backward_error_independent": berr,
                         "fell_back_to_fp64": bool(st.fell_back)}
            if st.fell_back:
                res[name]["low_precision_attempt"] = {"steps": st.lowp_iterations,
                                                      "krylov_steps": st.lowp_krylov_steps,
                                                      "backward_error": st.lowp_backward_error}
        for name in ("fp32_ir", "bf16_ir", "bf16x3_ir"):
            res[name]["speedup_vs_fp64"] = round(res["fp64"]["ms"] / res[name]["ms"], 3)
        out["kappa"][f"{kappa:g}"] = res
        if kappa == 1e2:
            # the FP64 factorisation alone (mrm_dev_lu_factor on a copy of A)
            LU = torch.empty_like(A)
            indx = torch.empty(n, dtype=torch.int32, device=dev)
            perm = torch.empty(n, dtype=torch.int32, device=dev)
            ts = []
            for rep in range(4):
                LU.copy_(A)
                torch.cuda.synchronize()
                t = time.perf_counter()
                rc = lib.mrm_dev_lu_factor(LU.data_ptr(), n * 8, n, indx.data_ptr(), perm.data_ptr(), sp)
                torch.cuda.synchronize()
                if rep > 0:
                    ts.append(time.perf_counter() - t)
            ms = 1e3 * statistics.median(ts)
            out["fp64_factor"] = {"rc": rc, "ms": round(ms, 1), "tflops": round(2 * n ** 3 / 3 / ms / 1e9, 2),
                                  "lookahead": bool(lib.mrm_lu_lookahead_available())}
            del LU
        del A
    return out


def batched_row(lib, reps: int = 5):
    """8192 GEMMs with m, n, k uniform in [16, 256] (seed 5101; FP64, and FP32
    pieces with error compensation) and 20000 solves with n uniform in
    [8, 128], A Gaussian + n I, one right-hand side (seed 5201; FP64, and FP32
    factors with FP64 refinement), each as one hipGraph replay of its plan
    (median of `reps`)."""
    import torch
    from mrmath_amd import batched as bt

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5101)
    sizes = [tuple(int(v) for v in rng.integers(16, 257, 3)) for _ in range(8192)]
    g = torch.Generator(device=dev).manual_seed(5101)
    probs = []
    for m, n, k in sizes:
        probs.append((torch.rand(m, k, dtype=torch.float64, device=dev, generator=g) * 2 - 1,
                      torch.rand(k, n, dtype=torch.float64, device=dev, generator=g) * 2 - 1,
                      torch.empty(m, n, dtype=torch.float64, device=dev)))
    flops = sum(2.0 * m * n * k for m, n, k in sizes)
    nbytes = sum(8.0 * (m * k + k * n + m * n) for m, n, k in sizes)

    def timed(plan):
        plan.run()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.run()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    gp = bt.dgemm_plan(probs, 0)
    ms = timed(gp)
    floor_ms = max(flops / (FP64_PEAK_TFLOPS * 1e12), nbytes / (HBM_PEAK_GBS * 1e9)) * 1e3
    out = {"gemm_8192": {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 2), "tile": gp.tile,
                         "gbps": round(nbytes / ms / 1e6, 1), "roofline_floor_ms": round(floor_ms, 4),
                         "frac_of_floor": round(floor_ms / ms, 4)}}
    gp.close()
    # FP32 pieces with error compensation, FP64 chunk sums (the FP32 mode):
    # time, and the normwise error of a sample against the FP64 plan's result
    ref = [p[2].clone() for p in probs[:64]]
    gp = bt.dgemm_plan(probs, 0, 1)
    ms = timed(gp)
    err = max(float(torch.linalg.norm(p[2] - r) / (torch.linalg.norm(p[0]) * torch.linalg.norm(p[1])))
              for p, r in zip(probs[:64], ref))
    out["gemm_8192_fp32_compensated"] = {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 2), "tile": gp.tile,
                                         "normwise_err_vs_fp64_sample": err}
    gp.close()
    del probs, ref
    rng = np.random.default_rng(5201)
    ns = [int(v) for v in rng.integers(8, 129, 20000)]
    g = torch.Generator(device=dev).manual_seed(5201)
    systems = []
    for n in ns:
        A = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        A.diagonal().add_(float(n))
        systems.append((A, torch.randn(n, dtype=torch.float64, device=dev, generator=g),
                        torch.empty(n, dtype=torch.float64, device=dev)))
    lflops = sum(2.0 / 3.0 * n ** 3 for n in ns)
    lbytes = sum(8.0 * (n * n + 2 * n) for n in ns)
    status = torch.empty(len(ns), dtype=torch.int32, device=dev)
    for mode, name in ((0, "lu_20000_fp64"), (1, "lu_20000_fp32_ir")):
        lp = bt.lu_plan(systems, mode, status)
        ms = timed(lp)
        ok = bool((status == 0).all())
        out[name] = {"ms": round(ms, 4), "systems_per_s": round(len(ns) / ms * 1e3, 0),
                     "gbps_matrices": round(lbytes / ms / 1e6, 1), "gflops": round(lflops / ms / 1e6, 1),
                     "all_nonsingular": ok}
        lp.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=None,
                    choices=sorted(WORKLOADS) + ["ozaki_8192", "lu_ir_16384", "batched", "copy_1gib",
                                                 "covariance_200000x512"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-rows", action="store_true", help="headline GEMM only")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(each GPU's share of a sharded GEMM is the single-GPU workload).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mrmath_amd import _lib

    lib = _lib.load()
    torch.cuda.set_device(0)
    if args.workload == "copy_1gib":
        return copy_calibration(args)
    if args.workload == "ozaki_8192":
        print(json.dumps({"workload": "ozaki_8192", **ozaki_row(lib)}), flush=True)
        return
    if args.workload == "batched":
        print(json.dumps({"workload": "batched", **batched_row(lib)}), flush=True)
        return
    if args.workload == "covariance_200000x512":
        print(json.dumps({"workload": "covariance_200000x512", **covariance_row(lib)}), flush=True)
        return
    if args.workload == "lu_ir_16384":
        print(json.dumps({"workload": "lu_ir_16384", **lu_ir_row(lib)}), flush=True)
        return

    workload = args.workload or "dgemm_nn_8192"
    r = gemm_run(lib, workload, args.steps, args.warmup)
    M, N, K = r["M"], r["N"], r["K"]
    out = {
        "metric": "GEMM TFLOP/s (% peak) at FP64 accuracy",
        "value": round(r["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform [-1,1) float64, device-resident)",
        "config": {"workload": workload, "M": M, "N": N, "K": K, "op": r["op"],
                   "path": "mrm_dev_dgemm -> dgemm_mfma_kernel (v_mfma_f64_16x16x4_f64)"},
        "pct_of_peak": round(100.0 * r["tflops"] / FP64_PEAK_TFLOPS, 2),
        "roofline": {
            "bound": "mfma",
            "achieved": round(r["achieved_tflops"], 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(r["achieved_tflops"] / FP64_PEAK_TFLOPS, 4),
            "traffic": _traffic_from_profile(workload),
            "kernel": "dgemm_mfma_kernel",
            "avg_kernel_ms": round(r["avg_launch_ms"], 4),
        },
        "check": {"linearity_rel_err": r["linearity_rel_err"], "bound": 4 * K * 2.0 ** -53},
    }
    if args.workload is None and not args.no_rows:
        rows = {}
        for w in ("dgemm_tn_4096", "dgemm_nt_4096", "syrk_200000x512"):
            # ten untimed launches first: the first tens of milliseconds of
            # GEMM work after a pause run at a lower clock
            q = gemm_run(lib, w, 10, 10)
            rows[w] = {"tflops": round(q["tflops"], 3), "ms": round(q["ms_per_step"], 4),
                       "pct_of_fp64_peak": round(100 * q["tflops"] / FP64_PEAK_TFLOPS, 2),
                       "linearity_rel_err": q["linearity_rel_err"]}
            if w.startswith("syrk"):
                rows[w]["hbm_gbps_algorithmic"] = round((q["M"] * q["K"] + q["M"] * q["N"]) * 8 /
                                                        (q["ms_per_step"] * 1e-3) / 1e9, 1)
                # X^T X is symmetric: the library computes the 128 x 128 tiles on
                # and above the diagonal and mirrors them; tflops counts the full
                # product 2 M N K, executed_tflops what the matrix cores ran
                t = -(-q["M"] // 128)
                frac = t * (t + 1) / 2 / (t * t)
                rows[w]["symmetric_tiles"] = f"{t * (t + 1) // 2}/{t * t}"
                rows[w]["executed_tflops"] = round(q["tflops"] * frac, 3)
                # the SYRK minimum n (n + 1) k of a symmetric product
                syrk = q["M"] * (q["M"] + 1) * q["K"]
                rows[w]["syrk_tflops"] = round(syrk / (q["ms_per_step"] * 1e-3) / 1e12, 3)
                rows[w]["syrk_pct_of_fp64_peak"] = round(100 * rows[w]["syrk_tflops"] / FP64_PEAK_TFLOPS, 2)
                # HBM bytes per launch from the committed counter summary
                rows[w]["traffic"] = _traffic_from_profile(w)
        rows["covariance_200000x512"] = covariance_row(lib)
        rows["batched"] = batched_row(lib)
        rows["ozaki_8192"] = ozaki_row(lib)
        rows["lu_ir_16384"] = lu_ir_row(lib)
        out["rows"] = rows
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
