"""Benchmark of the batched matrix-function evaluation on one MI355X.

Workload (default): symmetric 3x3 matrices, matrix layout, order 2 (F 3x3, J 3^4, H 3^6), FP64, n = 2^22 per function
for exp, log, sqrt and pow(0.3); inputs A = Q diag(lam) Q^T, Q Haar-random, lam log-uniform in [1e-2, 1e2], with
15% exactly double, 5% triple and 10% 1e-8-gap eigenvalues (seed 0).  One step = the four functions over the n
matrices (one smf_hip_eval_matrix call each), inputs resident in HBM, outputs overwritten each step.

    python bench.py [--steps K] [--warmup W] [--n N] [--gpus 1] [--no-cpu] [--gmf]

Prints one JSON line: value = matrices/s of the whole step (4 n matrices per step); `roofline` for the dominant
kernel (smf order-2 matrix-layout kernel): algorithmic bytes per launch (A in + F, J, H out: 6624 B per matrix) over
its average launch time measured with HIP events on its stream; `cpu_baseline` the reference C library (built from
the reference sources into oracle/_ref, SIMD + OpenMP) or, without it, this repository's scalar C oracle, on a bounded
sample of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "matrices/s at order 2 (FP64 vs mixed, ≤1e-5 rel err), % HBM roofline"
HBM_PEAK_GBS = 8000.0
BYTES_O2_MATRIX = 72 + 72 + 648 + 5832          # A in, F, J, H out (FP64, matrix layout)
FUNCS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(A, seconds=10.0, chunk=1 << 17):
    """the reference C library (or the oracle) over chunks of the same inputs, all four functions, until `seconds`"""
    from oracle.pyoracle import Checker, available
    kind = "reference" if available("ref") else "port"
    chk = Checker("ref" if kind == "reference" else "oracle")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    chk.set_threads(threads)
    done, t0, lo = 0, time.perf_counter(), 0
    chk.smf_matrix("log", 0.0, A[:256], 2)                         # warm-up (libmvec binding, thread pool)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and lo < len(A):
        X = A[lo:lo + chunk]
        for k, al in FUNCS:
            rc, *_ = chk.smf_matrix(k, al, X, 2)
            assert rc == 0
            done += len(X)
        lo += chunk
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "matrices/s", "cores": threads, "kind": kind,
            "sample": f"{done} matrices ({done // 4} per function, first rows of the same inputs), {chk.name()}, "
                      f"{dt:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (the batch shards trivially "
                                   "by matrix ranges, see DESIGN.md)"}))
        sys.exit(2)

    import torch
    import matfun_amd
    from matfun_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lib = matfun_amd.library()
    n = args.n
    log(f"bench: generating {n} symmetric inputs (seed 0)")
    A = W.spd(n, seed=0)
    dA = torch.from_numpy(A).to(dev)
    F = torch.empty((n, 9), dtype=torch.float64, device=dev)
    J = torch.empty((n, 81), dtype=torch.float64, device=dev)
    H = torch.empty((n, 729), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in FUNCS]

    def step(record):
        for i, (k, al) in enumerate(FUNCS):
            if record:
                ev[i][0].record(stream)
            rc = lib.smf_hip_eval_matrix(matfun_amd.KINDS[k], al, n, dA.data_ptr(), F.data_ptr(), J.data_ptr(),
                                         H.data_ptr(), sh)
            if record:
                ev[i][1].record(stream)
            assert rc == 0, rc

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    kern_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        kern_ms += [a.elapsed_time(b) for a, b in ev]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = 4 * n / (wall / args.steps)
    avg_kernel_ms = float(np.mean(kern_ms))
    achieved = n * BYTES_O2_MATRIX / (avg_kernel_ms * 1e-3) / 1e9
    per_func = {k if k != "pow" else "pow0.3": n / (float(np.mean(kern_ms[i::4])) * 1e-3) for i, (k, _) in enumerate(FUNCS)}
    log(f"bench: {ms_per_step:.2f} ms/step, kernel avg {avg_kernel_ms:.3f} ms, {achieved:.0f} GB/s")

    # device copy peak of the same byte count order (read + write), for context
    x = torch.empty(1 << 28, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(5):
        y.copy_(x)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * x.numel() * 8 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del x, y

    out = {"metric": METRIC, "value": value, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "smf exp/log/sqrt/pow(0.3), matrix layout, order 2 (F 3x3, J 3^4, H 3^6), FP64, "
                                  "n=2^22 per function, SPD Q diag(lam) Q^T with exact/near repeated eigenvalues",
                      "n_per_function": n, "functions": [f"{k}{al if k == 'pow' else ''}" for k, al in FUNCS],
                      "order": 2, "layout": "matrix", "bytes_per_matrix": BYTES_O2_MATRIX,
                      "per_function_matrices_per_s": per_func},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "mf::smf_kernel<KIND,true,2>",
                        "avg_kernel_ms": avg_kernel_ms, "measured_copy_GBs": copy_gbs}}
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(A, seconds=args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
