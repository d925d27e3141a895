#!/usr/bin/env python3
"""Benchmark of the FP64 GEMM hot path (mrMath's TDoubleMatrix.Mult / MatrixMult)
on one MI355X.

One step = one C = A*B on device-resident 8192 x 8192 float64 operands
(entries uniform [-1, 1)) through the library's device entry, i.e. the
square case of the GPU GEMM family.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

Workloads: dgemm_nn_8192 (default), dgemm_tn_4096, dgemm_nt_4096, and
copy_1gib (a plain 1 GiB device copy: the calibration kernel for the HBM byte
counters FETCH_SIZE / WRITE_SIZE).
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

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak (AMD published)
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "dgemm_nn_8192": (0, 8192, 8192, 8192),
    "dgemm_tn_4096": (1, 4096, 4096, 4096),
    "dgemm_nt_4096": (2, 4096, 4096, 4096),
}


def _traffic_from_profile(workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), recorded in profiles/."""
    f = ROOT / "profiles" / f"dgemm_mfma_{workload}_pmc.txt"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(threads: int, n: int = 2048, budget_s: float = 15.0):
    """The restated reference fast path (blocked 128 + 4-lane FMA kernel +
    ThrMatrixMultEx2 tiling) on the host cores, on an n^3 sample."""
    import numpy as np
    import oracle as orc

    A = orc.pm1_matrix(n, n, 8101)
    B = orc.pm1_matrix(n, n, 8102)
    C = np.zeros((n, n))
    used = orc.fast_mult(C, A, B, threads)  # warm-up
    times = []
    t_start = time.monotonic()
    while len(times) < 5 and (time.monotonic() - t_start) < budget_s:
        t0 = time.monotonic()
        used = orc.fast_mult(C, A, B, threads)
        times.append(time.monotonic() - t0)
    t = statistics.median(times)
    return {
        "value": round(2.0 * n ** 3 / t / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": int(used),
        "kind": "port",
        "sample": f"dgemm NN {n}^3, U[-1,1) MT19937 seeds 8101/8102, median of {len(times)} runs "
                  f"after 1 warm-up, {used} threads (ThrMatrixMultEx2 tiling, block 128, AVX2/FMA)",
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dgemm_nn_8192", choices=sorted(WORKLOADS) + ["copy_1gib"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(each GPU's share of a sharded GEMM is the single-GPU workload).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mrmath_amd import _lib

    lib = _lib.load()
    if args.workload == "copy_1gib":
        return copy_calibration(args)
    trans, M, N, K = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = torch.Generator(device=dev).manual_seed(8101)
    rowsA, colsA = (K, M) if trans == 1 else (M, K)
    rowsB, colsB = (N, K) if trans == 2 else (K, N)
    A = torch.rand(rowsA, colsA, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    B = torch.rand(rowsB, colsB, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    C = torch.empty(M, N, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        rc = lib.mrm_dev_dgemm(trans, M, N, K, A.data_ptr(), colsA * 8, B.data_ptr(), colsB * 8,
                               C.data_ptr(), N * 8, 0, sp)
        if rc:
            _lib.check(rc)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # per-launch HIP events on the stream the kernel runs on
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in ev]

    # sanity: C*x == A*(B*x) to the normwise bound (outside the timed region)
    x = torch.randn(N, dtype=torch.float64, device=dev, generator=g)
    Am = A.t() if trans == 1 else A
    Bm = B.t() if trans == 2 else B
    lhs = C @ x
    rhs = Am @ (Bm @ x)
    rel = (torch.linalg.vector_norm(lhs - rhs) /
           (torch.linalg.vector_norm(A) * torch.linalg.vector_norm(B) * torch.linalg.vector_norm(x))).item()
    if not rel <= 4 * K * 2.0 ** -53:
        print(json.dumps({"error": "GEMM result failed the linearity check", "rel": rel}), file=sys.stderr)
        sys.exit(1)

    flops = 2.0 * M * N * K
    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_ms = statistics.mean(kernel_ms)
    value = flops / (wall / args.steps) / 1e12
    achieved = flops / (avg_kernel_ms * 1e-3) / 1e12

    out = {
        "metric": "GEMM TFLOP/s (% peak) at FP64 accuracy",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform [-1,1) float64, device-resident)",
        "config": {"workload": args.workload, "M": M, "N": N, "K": K,
                   "op": ["NN", "TN (MultT1)", "NT (MultT2)"][trans],
                   "path": "mrm_dev_dgemm -> dgemm_mfma_kernel (v_mfma_f64_4x4x4_4b_f64)"},
        "pct_of_peak": round(100.0 * value / FP64_PEAK_TFLOPS, 2),
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": _traffic_from_profile(args.workload),
            "kernel": "dgemm_mfma_kernel",
            "avg_kernel_ms": round(avg_kernel_ms, 4),
        },
        "check": {"linearity_rel_err": rel, "bound": 4 * K * 2.0 ** -53},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1, 64)
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
