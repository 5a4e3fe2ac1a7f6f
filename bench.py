"""Benchmark of the batched matrix-function evaluation on one MI355X.

Headline workload: symmetric 3x3 matrices, matrix layout, order 2 (F 3x3, J 3^4, H 3^6), FP64, n = 2^22 per
function for exp, log, sqrt and pow(0.3); inputs A = Q diag(lam) Q^T, Q Haar-random, lam log-uniform in [1e-2, 1e2],
with 15% exactly double, 5% triple and 10% 1e-8-gap eigenvalues (seed 0).  One step = the four functions over the n
matrices (one smf_hip_eval_matrix call each), inputs resident in HBM, outputs overwritten each step.

Also measured (same JSON line, key "legs"): the mixed form of the same workload (FP32 outputs; log, sqrt, pow(0.3):
exp of eigenvalues up to 1e2 is beyond the FP32 range) with its error against the FP64 form, and general matrices
(gmf_hip_eval, n = 2^20 per function, similarity / complex-pair / shear / 4I+N mix, seed 0) in FP64 and mixed form.
Also (key "legs"): the packed symmetric outputs of the headline matrices as Voigt vectors (smf_hip_eval_packed:
F 6, J 21, H 126 values per matrix, 1272 B per matrix FP64) in both precisions.
Key "mixed_vs_fp64": the mixed form against the FP64 form on smf log and smf pow with one exponent per matrix drawn
uniform in [-1, 1] (order 2, n = 2^24, the headline distribution, seed 1) plus gmf log order 1 (n = 2^22, the general
mix, seed 1): time per step of each form, the speed-up, and the mixed form's error against the FP64 form per output.
Key "hetero_batch": a heterogeneous batch of 2^20 finite elements (U{1,4,8,10,27} points; smf log o2 / smf pow o1
with per-matrix exponents / gmf exp o1 / gmf log o2) evaluated through the plan's tuned launch strategy, 20 Newton
iterations per hipGraph replay (each perturbing A inside the evaluation's kernels), FP64 and mixed (hetero_batch's
docstring).

    python bench.py [--steps K] [--warmup W] [--n N] [--gmf-n N] [--gpus 1] [--no-cpu] [--only-headline]
                    [--only-order0] [--only-batch]
                    [--mixed-n N] [--mixed-gmf-n N] [--batch-elements E] [--batch-iters I]

`value` = matrices/s of the headline step (4 n matrices per step).  `roofline`: the dominant kernel (the smf order-2
matrix-layout FP64 kernel): algorithmic bytes per launch (A in + F, J, H out: 6624 B per matrix) over its average
launch time, measured with HIP events on the launch stream.  `cpu_baseline`: the reference C library (built from the
reference sources into oracle/_ref; SIMD + OpenMP) or, without it, this repository's scalar C oracle, on a bounded
sample of the headline inputs on the host cores.
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
FUNCS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)]


def fname(k, al):
    return k if k != "pow" else f"pow{al:g}"


def bytes_per_matrix(order, out_bytes):
    return 72 + out_bytes * (9 + (81 if order >= 1 else 0) + (729 if order >= 2 else 0))


def profiled_traffic(n):
    """HBM bytes per dispatch of the headline kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/smf_order2_matrix_fp64_traffic.md, counters scaled by a 2 GiB device copy), scaled to n matrices"""
    path = os.path.join(ROOT, "profiles", "smf_order2_matrix_fp64_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch_n4194304:"):
                return float(line.split(":")[1]) * n / 4194304
    except OSError:
        pass
    return None


def profiled_valu():
    """VALU figures of the order-0 kernels from the committed rocprofv3 summary (profiles/order0_valu.md, lines
    `valu <kernel>: <key>=<value> ...`), or None"""
    path = os.path.join(ROOT, "profiles", "order0_valu.md")
    out = {}
    try:
        for line in open(path):
            if line.startswith("valu "):
                name, rest = line[5:].split(":", 1)
                out[name.strip()] = {k: float(v) for k, v in (kv.split("=") for kv in rest.split())}
    except OSError:
        return None
    return out or None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(A, seconds=10.0, chunk=1 << 17):
    """the reference C library (or the oracle) over chunks of the same inputs, all four functions, until `seconds`"""
    from oracle.pyoracle import Checker, available
    kind = "reference" if available("ref") else "port"
    chk = Checker("ref" if kind == "reference" else "oracle")
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    chk.set_threads(threads)
    done, lo = 0, 0
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
    return {"value": done / dt, "unit": "matrices/s", "cores": threads, "kind": kind, "cpu": cpu_model(),
            "host_threads_available": len(os.sched_getaffinity(0)),
            "sample": f"{done} matrices ({done // 4} per function, first rows of the headline inputs), {chk.name()}, "
                      f"order 2 matrix layout, {dt:.1f} s wall"}


def cpu_model():
    """the host CPU as /proc/cpuinfo names it, with its family / model / stepping numbers and whether it has AVX-512
    (virtual machines often report only the vendor line as the model name)"""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "flags") and k not in info:
                    info[k] = v
            if not line.strip() and info:
                break
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    extra = [f"family {info['cpu family']}" if "cpu family" in info else "",
             f"model {info['model']}" if "model" in info else "",
             f"stepping {info['stepping']}" if "stepping" in info else "",
             "AVX-512" if "avx512f" in info.get("flags", "").split() else "no AVX-512"]
    return f"{name} ({', '.join(e for e in extra if e)})"


def c0_baseline(runs=7, n=65536):
    """C0: smf log, Voigt layout, order 1 (F n x 6, J n x 6 x 6), n = 65536 SPD matrices (Q Haar-random from seed 0,
    eigenvalues log-uniform in [1e-2, 1e2], no repeated ones) on the reference C path (c/src/matfun.cc:87-89: the
    reference library built from its sources, SIMD + OpenMP on all threads given) -- median of `runs` timed calls
    after a warm-up"""
    from oracle.pyoracle import Checker, available
    from matfun_amd import workloads as W
    kind = "reference" if available("ref") else "port"
    chk = Checker("ref" if kind == "reference" else "oracle")
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    chk.set_threads(threads)
    v = W.voigt(W.spd(n, seed=0, special=False))
    rc, *_ = chk.smf("log", 0.0, v, 1)                             # warm-up
    assert rc == 0
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        rc, *_ = chk.smf("log", 0.0, v, 1)
        ts.append(time.perf_counter() - t0)
        assert rc == 0
    med = float(np.median(ts))
    return {"value": n / med, "unit": "matrices/s", "cores": threads, "kind": kind, "cpu": cpu_model(),
            "ms_median": med * 1e3, "runs": runs, "n": n,
            "sample": f"C0: smf log, Voigt, order 1, n={n}, {chk.name()}, median of {runs} after a warm-up"}, v


def timed(torch, stream, calls, steps, warmup):
    """run the closures `calls` (one step) warmup + steps times; average kernel time of each (HIP events on the
    launch stream) and the wall time per step"""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in calls]
    for _ in range(warmup):
        for c in calls:
            assert c() == 0
    torch.cuda.synchronize()
    ks = [[] for _ in calls]
    t0 = time.perf_counter()
    for _ in range(steps):
        for i, c in enumerate(calls):
            ev[i][0].record(stream)
            assert c() == 0
            ev[i][1].record(stream)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(ev):
            ks[i].append(a.elapsed_time(b))
    return [float(np.mean(k)) for k in ks], (time.perf_counter() - t0) / steps


def mixed_vs_fp64(torch, lib, W, dev, stream, args, sample=1 << 16):
    """smf log + smf pow (per-matrix exponent ~ U[-1, 1]) order 2 and gmf log order 1 (the length-sorted entry
    point, faster at order 1), FP64 form vs mixed form"""
    import matfun_amd
    n, ng = args.mixed_n, args.mixed_gmf_n
    log(f"bench: mixed vs fp64: generating {n} symmetric and {ng} general inputs (seed 1)")
    dA = torch.from_numpy(W.spd(n, seed=1)).to(dev)
    dal = torch.from_numpy(np.random.default_rng(1).uniform(-1.0, 1.0, n)).to(dev)
    dG = torch.from_numpy(W.general(ng, seed=1)).to(dev)
    sh = stream.cuda_stream
    work = torch.empty(lib.gmf_hip_sort_workspace(ng), dtype=torch.uint8, device=dev)
    res = {"workload": f"smf log + smf pow(alpha per matrix ~ U[-1,1]), matrix layout, order 2, n={n} each; gmf log "
                       f"order 1, n={ng} (gmf_hip_eval_sorted*)", "n_smf": n, "n_gmf": ng}
    t = {}
    for form, dt, eb in (("fp64", torch.float64, 8), ("mixed", torch.float32, 4)):
        mixed = form == "mixed"
        F = torch.empty(n * 9, dtype=dt, device=dev)
        J = torch.empty(n * 81, dtype=dt, device=dev)
        H = torch.empty(n * 729, dtype=dt, device=dev)
        p = (F.data_ptr(), J.data_ptr(), H.data_ptr())
        smf = lib.smf_hip_eval_matrix_mixed if mixed else lib.smf_hip_eval_matrix
        smfp = lib.smf_hip_eval_matrix_pow_mixed if mixed else lib.smf_hip_eval_matrix_pow
        gmf = lib.gmf_hip_eval_sorted_mixed if mixed else lib.gmf_hip_eval_sorted
        calls = [lambda: smf(matfun_amd.KINDS["log"], 0.0, n, dA.data_ptr(), *p, sh),
                 lambda: smfp(dal.data_ptr(), n, dA.data_ptr(), *p, sh),
                 lambda: gmf(matfun_amd.KINDS["log"], 0.0, ng, dG.data_ptr(), p[0], p[1], None, work.data_ptr(),
                             work.numel(), sh)]
        ks, wall = timed(torch, stream, calls, args.steps, args.warmup)
        b2, b1 = bytes_per_matrix(2, eb), bytes_per_matrix(1, eb)
        res[form] = {"ms_per_step": wall * 1e3, "avg_kernel_ms": {"smf_log": ks[0], "smf_pow_alpha": ks[1],
                                                                  "gmf_log_o1": ks[2]},
                     "frac_hbm": {"smf_log": n * b2 / (ks[0] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "smf_pow_alpha": n * b2 / (ks[1] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "gmf_log_o1": ng * b1 / (ks[2] * 1e-3) / 1e9 / HBM_PEAK_GBS},
                     "matrices_per_s": (2 * n + ng) / (sum(ks) * 1e-3)}
        t[form] = sum(ks)
        log(f"bench: mixed vs fp64: {form}: kernels {ks[0]:.2f} + {ks[1]:.2f} + {ks[2]:.2f} ms")
        del F, J, H
        torch.cuda.empty_cache()
    res["speedup_mixed_vs_fp64"] = t["fp64"] / t["mixed"]
    # error of the mixed form against the FP64 form (itself checked against the CPU oracle by the test suite) on the
    # first `sample` matrices of each call
    err = {}
    m = min(sample, n)
    for name, order, run in (
            ("smf_log", 2, lambda mx, o: (lib.smf_hip_eval_matrix_mixed if mx else lib.smf_hip_eval_matrix)(
                matfun_amd.KINDS["log"], 0.0, m, dA.data_ptr(), *o, sh)),
            ("smf_pow_alpha", 2, lambda mx, o: (lib.smf_hip_eval_matrix_pow_mixed if mx else
                                                lib.smf_hip_eval_matrix_pow)(dal.data_ptr(), m, dA.data_ptr(), *o, sh)),
            ("gmf_log_o1", 1, lambda mx, o: (lib.gmf_hip_eval_mixed if mx else lib.gmf_hip_eval)(
                matfun_amd.KINDS["log"], 0.0, min(m, ng), dG.data_ptr(), o[0], o[1], None, sh))):
        mm = min(m, ng) if name.startswith("gmf") else m
        outs = []
        for mx, dt in ((False, torch.float64), (True, torch.float32)):
            o = [torch.empty((mm, s), dtype=dt, device=dev) for s in (9, 81, 729)]
            assert run(mx, [x.data_ptr() for x in o]) == 0
            outs.append(o)
        torch.cuda.synchronize()
        valid = ~torch.isnan(outs[0][0]).any(dim=1)
        e = {"sample": mm, "nan_matrices": int((~valid).sum())}
        for q, nm in enumerate("FJH"[:order + 1]):
            a, b = outs[0][q], outs[1][q]
            r = (torch.linalg.norm(b.double() - a, dim=1) / torch.linalg.norm(a, dim=1))[valid]
            e[nm] = {"max": float(r.max()), "p99.9": float(torch.quantile(r[: 1 << 24], 0.999)),
                     "median": float(r.median())}
        err[name] = e
    res["mixed_err_vs_fp64"] = err
    log(f"bench: mixed vs fp64: speed-up {res['speedup_mixed_vs_fp64']:.3f}")
    del dA, dal, dG, work
    torch.cuda.empty_cache()
    return res


def hetero_batch(torch, W, dev, stream, args):
    """Heterogeneous persistent batch: E finite elements with U{1, 4, 8, 10, 27} integration points each; per element
    40% smf log order 2, 20% smf pow order 1 (exponent per matrix ~ U[-1, 1]), 20% gmf exp order 1, 20% gmf log
    order 2; inputs as the headline (smf) and general (gmf) workloads (seed 2 / 3); one step = one replay of a hipGraph
    of `iters` Newton iterations (relative N(0, 1e-3) perturbation of A, then the batch evaluation), launch strategy
    tuned first; FP64 and mixed outputs"""
    import matfun_amd
    E, iters = args.batch_elements, args.batch_iters
    rng = np.random.default_rng(2)
    pts = rng.choice(np.array([1, 4, 8, 10, 27]), E)
    cls = rng.choice(4, E, p=[0.4, 0.2, 0.2, 0.2])      # smf log o2, smf pow o1, gmf exp o1, gmf log o2
    fam = np.where(cls < 2, 0, 1)
    kind = np.array([1, 2, 0, 1])[cls]
    order = np.array([2, 1, 1, 2])[cls]
    groups = matfun_amd.group_array(fam, kind, order, pts)
    N = int(pts.sum())
    is_smf = np.repeat(fam == 0, pts)
    log(f"bench: batch: {E} elements, {N} matrices; generating inputs (seed 2 / 3)")
    A = np.empty((N, 9))
    A[is_smf] = W.spd(int(is_smf.sum()), seed=2)
    A[~is_smf] = W.general(int((~is_smf).sum()), seed=3)
    alpha = np.random.default_rng(4).uniform(-1.0, 1.0, N)
    b = matfun_amd.Batch(groups)
    dA0 = torch.from_numpy(A).to(dev)
    dA = torch.empty_like(dA0)
    dal = torch.from_numpy(alpha).to(dev)
    nJ, nH = b.nj, b.nh
    res = {"workload": f"{E} elements x U{{1,4,8,10,27}} points = {N} matrices; 40% smf log o2, 20% smf pow o1 "
                       f"(alpha per matrix ~ U[-1,1]), 20% gmf exp o1, 20% gmf log o2; {iters} Newton iterations "
                       f"(A <- A o (1 + N(0, 1e-3)), Philox4x32-10) per hipGraph replay",
           "elements": E, "matrices": N, "j_rows": nJ, "h_rows": nH, "tiles": b.ntiles, "iterations": iters}
    for prec, eb in (("fp64", 8), ("mixed", 4)):
        out = b.outputs(prec)
        dA.copy_(dA0)
        tuned = b.tune(dA, dal, out=out, precision=prec)
        seq = b.sequence(dA0, dA, out, iters, 1e-3, 12345, alpha=dal, precision=prec)
        for _ in range(args.warmup):
            seq.launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            seq.launch()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        # algorithmic bytes per iteration: A in (72 B) + F out + J / H rows out + the perturbation's read and write of
        # A (144 B) + 8 B of exponent per pow matrix
        npow = int(pts[cls == 1].sum())
        byts = N * (72 + 144 + 9 * eb) + nJ * 81 * eb + nH * 729 * eb + 8 * npow
        t_it = ms / iters
        res[prec] = {"ms_per_graph": ms, "ms_per_iteration": t_it, "matrices_per_s": N / (t_it * 1e-3),
                     "bytes_per_matrix": byts / N, "achieved_GBs": byts / (t_it * 1e-3) / 1e9,
                     "frac_hbm": byts / (t_it * 1e-3) / 1e9 / HBM_PEAK_GBS, "tuned_strategy": tuned}
        log(f"bench: batch {prec}: {t_it:.3f} ms per iteration, {N / (t_it * 1e-3) / 1e6:.1f} M matrices/s, "
            f"{res[prec]['frac_hbm']:.1%} of HBM peak, strategy {tuned}")
        seq.close()
        del out
        torch.cuda.empty_cache()
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_graph"] / res["mixed"]["ms_per_graph"]
    b.close()
    del dA0, dA, dal
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--gmf-n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--only-order0", action="store_true", help="only the order-0 legs (for the VALU counter runs)")
    ap.add_argument("--only-hencky", action="store_true", help="the headline and the Hencky legs only")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--mixed-n", type=int, default=1 << 24)
    ap.add_argument("--mixed-gmf-n", type=int, default=1 << 22)
    ap.add_argument("--batch-elements", type=int, default=1 << 20)
    ap.add_argument("--batch-iters", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--only-gmf", action="store_true", help="only the general-matrix legs (for the counter runs)")
    ap.add_argument("--only-batch", action="store_true", help="only the heterogeneous batch leg")
    ap.add_argument("--gmf-strategy", type=int, choices=(0, 1, 2), default=None,
                    help="every function, order and precision of the sorted entry points on this strategy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 (the batch shards "
                                   "trivially by contiguous matrix ranges, one process per GPU, with no "
                                   "communication: every matrix is independent)"}))
        sys.exit(2)

    import torch
    import matfun_amd
    from matfun_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lib = matfun_amd.library()
    n, ng = args.n, args.gmf_n
    tiles = gtiles = None
    if not args.no_tune:
        import matfun_amd.tune as T
        log("bench: tuning the symmetric order-2 H store tile (matfun_amd.tune)")
        tuned = T.tune_smf_tiles(layouts=("matrix",))
        tiles = {f"{k[0]} {k[1]} {k[2]}": v[0] for k, v in tuned.items()}
        log(f"bench: tiles {tiles}")
        log("bench: tuning the general-matrix workgroup shape and evaluation strategy (matfun_amd.tune)")
        gt = T.tune_gmf(orders=(0, 1, 2), n=args.gmf_n)
        gtiles = {f"{k[0]} o{k[1]} {k[2]}": {"tiles": v["tiles"][0], "strategy": v["strategy"][0],
                                             "ms": {f"tiles{a}": round(b, 4) for a, b in v["tiles"][1].items()} |
                                             {f"strategy{a}": round(b, 4) for a, b in v["strategy"][1].items()}}
                  for k, v in gt.items()}
        log("bench: gmf " + ", ".join(f"{k}: {v['tiles']}/{v['strategy']}" for k, v in gtiles.items()))
    log(f"bench: generating {n} symmetric and {ng} general inputs (seed 0)")
    A = W.spd(n, seed=0)
    dA = torch.from_numpy(A).to(dev)
    dG = torch.from_numpy(W.general(ng, seed=0)).to(dev)
    # one set of output buffers, FP64-sized; the mixed form writes FP32 into the same memory
    Fb = torch.empty(n * 9, dtype=torch.float64, device=dev)
    Jb = torch.empty(n * 81, dtype=torch.float64, device=dev)
    Hb = torch.empty(n * 729, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def leg(name, fn, X, m, out_bytes, steps, warmup, funcs=FUNCS, bpm=None, order=2, work=None):
        """one call per function in `funcs` per step; order 0 / 1 / 2 (F; F, J; F, J, H); work: the workspace
        (tensor) of the length-sorted entry points"""
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in funcs]
        ptrs = (Fb.data_ptr(), Jb.data_ptr() if order >= 1 else None, Hb.data_ptr() if order >= 2 else None)
        ws = (work.data_ptr(), work.numel()) if work is not None else ()

        def step(record):
            for i, (k, al) in enumerate(funcs):
                if record:
                    ev[i][0].record(stream)
                rc = fn(matfun_amd.KINDS[k], al, m, X.data_ptr(), *ptrs, *ws, sh)
                if record:
                    ev[i][1].record(stream)
                assert rc == 0, (name, rc)

        for _ in range(warmup):
            step(False)
        torch.cuda.synchronize()
        kern = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(True)
            torch.cuda.synchronize()
            kern += [a.elapsed_time(b) for a, b in ev]
        wall = (time.perf_counter() - t0) / steps
        kms = float(np.mean(kern))
        bpm = bpm or bytes_per_matrix(order, out_bytes)
        gbs = m * bpm / (kms * 1e-3) / 1e9
        nf = len(funcs)
        per = {fname(k, al): m / (float(np.mean(kern[i::nf])) * 1e-3) for i, (k, al) in enumerate(funcs)}
        log(f"bench: {name}: {wall * 1e3:.2f} ms/step, kernel avg {kms:.3f} ms, {gbs:.0f} GB/s "
            f"({gbs / HBM_PEAK_GBS:.1%} of HBM peak), {nf * m / wall / 1e6:.1f} M matrices/s")
        return {"functions": [fname(k, al) for k, al in funcs], "order": order, "matrices_per_s": nf * m / wall,
                "ms_per_step": wall * 1e3, "n_per_function": m, "bytes_per_matrix": bpm, "avg_kernel_ms": kms,
                "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "per_function_matrices_per_s": per}

    def mixed_error(fn64, fnmx, X, funcs, m=16384):
        """max over m matrices of the per-matrix normwise relative error of the mixed form against the FP64 form
        (itself checked against the CPU oracle by the test suite), per function and output"""
        res = {}
        for k, al in funcs:
            outs = []
            for fn, dt in ((fn64, torch.float64), (fnmx, torch.float32)):
                o = [torch.empty((m, s), dtype=dt, device=dev) for s in (9, 81, 729)]
                assert fn(matfun_amd.KINDS[k], al, m, X.data_ptr(), *[t.data_ptr() for t in o], sh) == 0
                outs.append(o)
            torch.cuda.synchronize()
            e = {}
            valid = ~torch.isnan(outs[0][0]).any(dim=1)       # (matrices outside the domain: NaN in both forms)
            for nm, a, b in zip("FJH", outs[0], outs[1]):
                err = torch.linalg.norm(b.double() - a, dim=1) / torch.linalg.norm(a, dim=1)
                e[nm] = float(err[valid].max())
            e["nan_matrices"] = int((~valid).sum())
            e["nan_agree"] = bool((torch.isnan(outs[1][0]).any(dim=1) == ~valid).all())
            res[fname(k, al)] = e
        return res

    def hencky_leg(name, fn, dt, steps, warmup):
        """mf_hip_hencky(_mixed) over the n resident results L, J, H of smf log at order 2 (the step after the
        evaluation in a Newton iteration on psi = 1/2 |log X|^2): reads L, J, H, writes psi, g, K"""
        fill = lib.smf_hip_eval_matrix_mixed if dt == torch.float32 else lib.smf_hip_eval_matrix
        assert fill(1, 0.0, n, dA.data_ptr(), Fb.data_ptr(), Jb.data_ptr(), Hb.data_ptr(), sh) == 0
        psi = torch.empty(n, dtype=dt, device=dev)
        g = torch.empty(n * 9, dtype=dt, device=dev)
        K = torch.empty(n * 81, dtype=dt, device=dev)
        call = lambda: fn(n, Fb.data_ptr(), Jb.data_ptr(), Hb.data_ptr(), psi.data_ptr(), g.data_ptr(), K.data_ptr(),
                          sh)
        for _ in range(warmup):
            assert call() == 0
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            assert call() == 0
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        bpm = (9 + 81 + 729 + 1 + 9 + 81) * (4 if dt == torch.float32 else 8)
        gbs = n * bpm / (ms * 1e-3) / 1e9
        log(f"bench: {name}: kernel avg {ms:.3f} ms, {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS:.1%} of HBM peak), "
            f"{n / ms / 1e3:.1f} M matrices/s")
        return {"matrices_per_s": n / (ms * 1e-3), "avg_kernel_ms": ms, "n": n, "bytes_per_matrix": bpm,
                "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}

    def fused_leg(steps, warmup, gen=False):
        """smf_hip_hencky_matrix (gmf_hip_hencky: gen) -- log X, psi, g, K of the n symmetric (ng general) inputs in
        one kernel, no J / H in memory"""
        m = ng if gen else n
        outs = [torch.empty(m * k, dtype=torch.float64, device=dev) for k in (9, 1, 9, 81)]
        fn, X = (lib.gmf_hip_hencky, dG) if gen else (lib.smf_hip_hencky_matrix, dA)
        call = lambda: fn(m, X.data_ptr(), *[o.data_ptr() for o in outs], sh)
        for _ in range(warmup):
            assert call() == 0
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            assert call() == 0
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        bpm = 72 + 8 * (9 + 1 + 9 + 81)
        gbs = m * bpm / (ms * 1e-3) / 1e9
        res = {"matrices_per_s": m / (ms * 1e-3), "avg_kernel_ms": ms, "n": m, "bytes_per_matrix": bpm,
               "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
        if gen:                                    # (the evaluation alone, for reference: gmf log at order 2)
            Jg = torch.empty(m * 81, dtype=torch.float64, device=dev)
            Hg = torch.empty(m * 729, dtype=torch.float64, device=dev)
            assert lib.gmf_hip_eval(1, 0.0, m, dG.data_ptr(), outs[0].data_ptr(), Jg.data_ptr(), Hg.data_ptr(),
                                    sh) == 0
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(steps):
                assert lib.gmf_hip_eval(1, 0.0, m, dG.data_ptr(), outs[0].data_ptr(), Jg.data_ptr(), Hg.data_ptr(),
                                        sh) == 0
            e1.record(stream)
            torch.cuda.synchronize()
            composed = e0.elapsed_time(e1) / steps * 1e-3 + m / legs["hencky_fp64"]["matrices_per_s"]
            del Jg, Hg
        else:
            composed = m / head["per_function_matrices_per_s"]["log"] + m / legs["hencky_fp64"]["matrices_per_s"]
        res["speedup_vs_eval_plus_step"] = composed / (ms * 1e-3)
        log(f"bench: hencky fused {'gmf' if gen else 'smf'} fp64: kernel avg {ms:.3f} ms (n = {m}), "
            f"{m / ms / 1e3:.1f} M matrices/s, {composed / (ms * 1e-3):.2f}x log order 2 + hencky "
            f"({composed * 1e3:.3f} ms)")
        return res

    funcs_mx = FUNCS[1:]          # (exp of eigenvalues up to 1e2 is beyond the FP32 range of the mixed form)

    def order0_legs():
        """order 0 (F only: 144 B per matrix FP64, 108 B mixed): bound by the per-matrix arithmetic (FP64 square roots,
        trigonometric root, transcendentals), not by HBM; the VALU counters of these kernels are read from
        profiles/order0_valu.md"""
        o0 = {"smf_fp64": leg("smf fp64 order 0", lib.smf_hip_eval_matrix, dA, n, 8, args.steps, args.warmup, order=0),
              "smf_mixed": leg("smf mixed order 0", lib.smf_hip_eval_matrix_mixed, dA, n, 4, args.steps, args.warmup,
                               funcs_mx, order=0),
              "gmf_fp64": leg("gmf fp64 order 0", lib.gmf_hip_eval, dG, ng, 8, args.steps, args.warmup, order=0),
              "gmf_mixed": leg("gmf mixed order 0", lib.gmf_hip_eval_mixed, dG, ng, 4, args.steps, args.warmup,
                               order=0)}
        o0["valu"] = profiled_valu()
        return o0

    if args.gmf_strategy is not None:
        for k in range(5):
            for o in range(3):
                for x in range(2):
                    assert lib.gmf_hip_set_sort_strategy(k, o, x, args.gmf_strategy) == 0
    if args.only_order0:
        print(json.dumps({"order0": order0_legs()}), flush=True)
        return
    if args.only_batch:
        print(json.dumps({"hetero_batch": hetero_batch(torch, W, dev, stream, args)}), flush=True)
        return
    if args.only_gmf:
        work = torch.empty(lib.gmf_hip_sort_workspace(ng), dtype=torch.uint8, device=dev)
        g = {}
        for o in (2, 1):
            g[f"fp64_o{o}"] = leg(f"gmf fp64 order {o}", lib.gmf_hip_eval, dG, ng, 8, args.steps, args.warmup, order=o)
            g[f"mixed_o{o}"] = leg(f"gmf mixed order {o}", lib.gmf_hip_eval_mixed, dG, ng, 4, args.steps, args.warmup,
                                   order=o)
            g[f"fp64_o{o}_sorted"] = leg(f"gmf fp64 order {o} sorted", lib.gmf_hip_eval_sorted, dG, ng, 8, args.steps,
                                         args.warmup, order=o, work=work)
            g[f"mixed_o{o}_sorted"] = leg(f"gmf mixed order {o} sorted", lib.gmf_hip_eval_sorted_mixed, dG, ng, 4,
                                          args.steps, args.warmup, order=o, work=work)
        print(json.dumps({"gmf": g, "strategy": args.gmf_strategy}), flush=True)
        return
    head = leg("smf fp64", lib.smf_hip_eval_matrix, dA, n, 8, args.steps, args.warmup)
    legs = {"smf_fp64": head}
    # the ceilings of the same output shape over the same buffers: mf_hip_store_probe (pure 16-byte stores of F, J,
    # H, no loads, no arithmetic) against the headline kernel's write rate (6552 of its 6624 B are stores), and
    # mf_hip_store_probe_in (the same after reading the input rows as the kernel does: its whole traffic, no
    # arithmetic) against the kernel's time
    probe = lambda: lib.mf_hip_store_probe(n, 9, Fb.data_ptr(), Jb.data_ptr(), Hb.data_ptr(), sh)
    probe_in = lambda: lib.mf_hip_store_probe_in(n, dA.data_ptr(), 9, 9, Fb.data_ptr(), Jb.data_ptr(), Hb.data_ptr(),
                                                 sh)
    pms, _ = timed(torch, stream, [probe, probe_in], max(args.steps, 5), args.warmup)
    wbytes = n * (9 + 81 + 729) * 8
    ceiling = {"probe_ms": pms[0], "probe_write_GBs": wbytes / (pms[0] * 1e-3) / 1e9,
               "kernel_write_GBs": wbytes / (head["avg_kernel_ms"] * 1e-3) / 1e9,
               "frac_of_write_ceiling": pms[0] / head["avg_kernel_ms"],
               "probe_in_ms": pms[1], "probe_in_GBs": n * 6624 / (pms[1] * 1e-3) / 1e9,
               "probe_in_frac_hbm": n * 6624 / (pms[1] * 1e-3) / 8e12,
               "frac_of_io_ceiling": pms[1] / head["avg_kernel_ms"]}
    log(f"bench: write ceiling (store probe, same shape): {pms[0]:.3f} ms = {ceiling['probe_write_GBs']:.0f} GB/s; "
        f"headline kernel {ceiling['kernel_write_GBs']:.0f} GB/s = {ceiling['frac_of_write_ceiling']:.1%} of it; "
        f"with the input reads (the kernel's whole traffic): {pms[1]:.3f} ms = {ceiling['probe_in_GBs']:.0f} GB/s, "
        f"kernel at {ceiling['frac_of_io_ceiling']:.1%} of it")
    if not (args.only_headline or args.only_hencky):
        legs["smf_mixed"] = leg("smf mixed", lib.smf_hip_eval_matrix_mixed, dA, n, 4, args.steps, args.warmup,
                                funcs_mx)
        # the mixed form's own ceiling: the FP32 output shape after the same input reads (3348 B per matrix)
        probe_in32 = lambda: lib.mf_hip_store_probe_in_f32(n, dA.data_ptr(), 9, 9, Fb.data_ptr(), Jb.data_ptr(),
                                                           Hb.data_ptr(), sh)
        pm, _ = timed(torch, stream, [probe_in32], max(args.steps, 5), args.warmup)
        legs["smf_mixed"]["io_ceiling"] = {"probe_in_ms": pm[0], "probe_in_GBs": n * 3348 / (pm[0] * 1e-3) / 1e9,
                                           "frac_of_io_ceiling": pm[0] / legs["smf_mixed"]["avg_kernel_ms"]}
        log(f"bench: smf mixed against its ceiling (FP32 store probe with the input reads): {pm[0]:.3f} ms, kernel at "
            f"{legs['smf_mixed']['io_ceiling']['frac_of_io_ceiling']:.1%} of it")
        # packed symmetric outputs (F 6, J 21, H 126 per matrix) of the same matrices as Voigt vectors
        dV = torch.from_numpy(W.voigt(A)).to(dev)
        legs["smf_packed_fp64"] = leg("smf packed fp64", lib.smf_hip_eval_packed, dV, n, 8, args.steps, args.warmup,
                                      bpm=48 + 8 * 153)
        legs["smf_packed_mixed"] = leg("smf packed mixed", lib.smf_hip_eval_packed_mixed, dV, n, 4, args.steps,
                                       args.warmup, funcs_mx, bpm=48 + 4 * 153)
        del dV
        legs["gmf_fp64"] = leg("gmf fp64", lib.gmf_hip_eval, dG, ng, 8, args.steps, args.warmup)
        legs["gmf_mixed"] = leg("gmf mixed", lib.gmf_hip_eval_mixed, dG, ng, 4, args.steps, args.warmup)
        # the same through the length-sorted entry points (workspace allocated once, outside the timed region)
        work = torch.empty(lib.gmf_hip_sort_workspace(ng), dtype=torch.uint8, device=dev)
        legs["gmf_fp64_sorted"] = leg("gmf fp64 sorted", lib.gmf_hip_eval_sorted, dG, ng, 8, args.steps, args.warmup,
                                      work=work)
        legs["gmf_mixed_sorted"] = leg("gmf mixed sorted", lib.gmf_hip_eval_sorted_mixed, dG, ng, 4, args.steps,
                                       args.warmup, work=work)
        legs["order0"] = order0_legs()
        del work
        legs["smf_mixed"]["max_err_vs_fp64"] = mixed_error(lib.smf_hip_eval_matrix, lib.smf_hip_eval_matrix_mixed,
                                                           dA, funcs_mx)
        legs["gmf_mixed"]["max_err_vs_fp64"] = mixed_error(lib.gmf_hip_eval, lib.gmf_hip_eval_mixed, dG, FUNCS)
        # FP64 vs mixed on the same functions (smf log, sqrt, pow(0.3) -- exp of the headline eigenvalues is beyond the
        # FP32 range -- and gmf exp, log, sqrt, pow(0.3)): summed kernel time per function, each form through its
        # faster general-matrix entry point (plain or length-sorted) per function
        t64_smf = sum(n / head["per_function_matrices_per_s"][fname(k, al)] for k, al in funcs_mx)
        tmx_smf = sum(n / legs["smf_mixed"]["per_function_matrices_per_s"][fname(k, al)] for k, al in funcs_mx)

        def gmf_time(form):
            a, b = legs[f"gmf_{form}"], legs[f"gmf_{form}_sorted"]
            return sum(ng / max(a["per_function_matrices_per_s"][f], b["per_function_matrices_per_s"][f])
                       for f in a["per_function_matrices_per_s"])
        t64, tmx = t64_smf + gmf_time("fp64"), tmx_smf + gmf_time("mixed")
        legs["smf_mixed"]["speedup_vs_fp64"] = t64_smf / tmx_smf
        legs["gmf_mixed"]["speedup_vs_fp64"] = gmf_time("fp64") / gmf_time("mixed")
        legs["mixed_vs_fp64_smf_plus_gmf"] = t64 / tmx
        legs["mixed_vs_fp64_smf_plus_gmf_functions"] = ("smf " + ", ".join(fname(k, al) for k, al in funcs_mx) +
                                                        "; gmf exp, log, sqrt, pow0.3 (faster of plain and sorted)")

    if not args.only_headline:
        legs["hencky_fp64"] = hencky_leg("hencky fp64", lib.mf_hip_hencky, torch.float64, args.steps, args.warmup)
        legs["hencky_mixed"] = hencky_leg("hencky mixed", lib.mf_hip_hencky_mixed, torch.float32, args.steps,
                                          args.warmup)
        legs["hencky_fused_smf_fp64"] = fused_leg(args.steps, args.warmup)
        legs["hencky_fused_gmf_fp64"] = fused_leg(args.steps, args.warmup, gen=True)
    if not (args.only_headline or args.only_hencky):
        del Fb, Jb, Hb
        torch.cuda.empty_cache()
        legs_c = mixed_vs_fp64(torch, lib, W, dev, stream, args)
        legs_b = hetero_batch(torch, W, dev, stream, args)
    # device copy bandwidth (read + write of 2 GiB), for context next to the 8 TB/s spec
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

    out = {"metric": METRIC, "value": head["matrices_per_s"], "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "smf exp/log/sqrt/pow(0.3), matrix layout, order 2 (F 3x3, J 3^4, H 3^6), FP64, "
                                  "n=2^22 per function, SPD Q diag(lam) Q^T with exact/near repeated eigenvalues",
                      "n_per_function": n, "functions": ["exp", "log", "sqrt", "pow0.3"], "order": 2,
                      "layout": "matrix", "bytes_per_matrix": head["bytes_per_matrix"]},
           "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": head["frac"], "traffic": profiled_traffic(n), "kernel": "mf::smf_kernel<KIND,true,2,double>",
                        "avg_kernel_ms": head["avg_kernel_ms"], "measured_copy_GBs": copy_gbs,
                        "write_ceiling": ceiling},
           "legs": legs, "tuned_tiles": tiles, "tuned_gmf": gtiles}
    if not (args.only_headline or args.only_hencky):
        out["mixed_vs_fp64"] = legs_c
        out["hetero_batch"] = legs_b
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(A, seconds=args.cpu_seconds)
        c0, v0 = c0_baseline()
        # the HIP path on the same C0 inputs (smf_hip_eval, Voigt, order 1; inputs and outputs resident on the GPU)
        dv = torch.from_numpy(v0).to(dev)
        F0 = torch.empty(len(v0) * 6, dtype=torch.float64, device=dev)
        J0 = torch.empty(len(v0) * 36, dtype=torch.float64, device=dev)
        ks, _ = timed(torch, stream, [lambda: lib.smf_hip_eval(1, 0.0, len(v0), dv.data_ptr(), F0.data_ptr(),
                                                               J0.data_ptr(), None, sh)], 20, 3)
        c0["hip_same_inputs"] = {"matrices_per_s": len(v0) / (ks[0] * 1e-3), "avg_call_ms": ks[0]}
        out["cpu_baseline"]["c0"] = c0
        log(f"bench: C0 (smf log Voigt order 1, n = {len(v0)}): CPU {c0['value'] / 1e6:.2f} M matrices/s on "
            f"{c0['cores']} threads ({c0['cpu']}); HIP {len(v0) / (ks[0] * 1e-3) / 1e6:.1f} M matrices/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
