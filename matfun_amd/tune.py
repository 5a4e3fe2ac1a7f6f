"""Autotuner of the kernels' run-time choices, each timed on the device and the fastest kept:

- the symmetric order-2 kernels' H store tile: 1 (10 matrices per phase-B pass, eigenvectors and second divided
  differences in their own LDS block) or 10, 8, 5 (matrices per pass, that block sharing the first phase's LDS
  region) -- tune_smf_tiles, smf_hip_set_tile;
- the general-matrix kernels' workgroup shape: 1, 2 or 4 wavefronts (tiles of 64 matrices) per workgroup sharing out
  their group series -- and the evaluation strategy of the workspace entry points (0 unsorted, 1 length-sorted
  permutation, 2 staged) -- tune_gmf, gmf_hip_set_tiles / gmf_hip_set_sort_strategy.

    import matfun_amd.tune
    table = matfun_amd.tune.tune_smf_tiles()      # times every choice, keeps the fastest (smf_hip_set_tile)
    gtable = matfun_amd.tune.tune_gmf()

Per function (exp, log, sqrt, pow with one exponent, pow with per-matrix exponents), layout (matrix, Voigt) and
precision (FP64, mixed): a smaller tile takes less LDS per wavefront, so more wavefronts share a compute unit, and
leaves more lanes idle while H is assembled; which wins depends on the function's register use and output width.
No result depends on the tile or the workgroup shape (checked bitwise by tests/test_gpu_tune.py); the staged strategy
changes results by rounding only (tests/test_gpu_sorted.py).  The timing uses HIP events on torch's
current stream over seeded SPD inputs resident on the device: per candidate the median of `reps` launches, the
candidates interleaved.
"""
import numpy as np

import matfun_amd

TILES = (1, 10, 8, 5)
SLOTS = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3, "pow_alpha": 4}


def _time_candidates(set_tile, fn, tiles, rounds):
    """median over `rounds` of each candidate's time, the candidates taken in turn within a round (so that a drift
    of the device's clocks or temperature affects them alike)"""
    import torch
    s = torch.cuda.current_stream()
    ev = {t: [] for t in tiles}
    for t in tiles:                                       # (warm-up, one launch each)
        set_tile(t)
        assert fn() == 0
    torch.cuda.synchronize()
    for _ in range(rounds):
        for t in tiles:
            set_tile(t)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            assert fn() == 0
            e1.record(s)
            ev[t].append((e0, e1))
    torch.cuda.synchronize()
    return {t: float(np.median([a.elapsed_time(b) for a, b in ev[t]])) for t in tiles}


def tune_smf_tiles(funcs=tuple(SLOTS), layouts=("matrix", "voigt"), precisions=("fp64", "mixed"), n=1 << 21, reps=7,
                   seed=0):
    """times each tile choice, sets the fastest; returns {(func, layout, precision): (tile, {tile: ms})}"""
    import torch
    from matfun_amd import workloads as W
    lib = matfun_amd.library()
    A = W.spd(n, seed=seed, log10_range=(-2.0, np.log10(60.0)))   # (exp stays inside the FP32 range)
    st = torch.cuda.current_stream().cuda_stream
    out = {}
    dal = torch.from_numpy(np.random.default_rng(seed).uniform(-1.0, 1.0, n)).cuda()
    for layout in layouts:
        X = torch.from_numpy(A if layout == "matrix" else W.voigt(A)).cuda()
        nf = 9 if layout == "matrix" else 6
        for prec in precisions:
            dt = torch.float32 if prec == "mixed" else torch.float64
            F = torch.empty(n * nf, dtype=dt, device="cuda")
            J = torch.empty(n * nf ** 2, dtype=dt, device="cuda")
            H = torch.empty(n * nf ** 3, dtype=dt, device="cuda")
            p = (F.data_ptr(), J.data_ptr(), H.data_ptr())
            mx = prec == "mixed"
            for func in funcs:
                slot = SLOTS[func]
                if func == "pow_alpha":
                    fn0 = {("matrix", False): lib.smf_hip_eval_matrix_pow, ("matrix", True): lib.smf_hip_eval_matrix_pow_mixed,
                           ("voigt", False): lib.smf_hip_eval_pow, ("voigt", True): lib.smf_hip_eval_pow_mixed}[(layout, mx)]
                    call = lambda: fn0(dal.data_ptr(), n, X.data_ptr(), *p, st)
                else:
                    fn0 = {("matrix", False): lib.smf_hip_eval_matrix, ("matrix", True): lib.smf_hip_eval_matrix_mixed,
                           ("voigt", False): lib.smf_hip_eval, ("voigt", True): lib.smf_hip_eval_mixed}[(layout, mx)]
                    al = 0.3 if func == "pow" else 0.0
                    call = lambda: fn0(matfun_amd.KINDS[func], al, n, X.data_ptr(), *p, st)
                ms = _time_candidates(lambda t: lib.smf_hip_set_tile(slot, layout == "matrix", mx, t), call, TILES,
                                      reps)
                best = min(ms, key=ms.get)
                assert lib.smf_hip_set_tile(slot, layout == "matrix", mx, best) == 0
                out[(func, layout, prec)] = (best, ms)
            del F, J, H
            torch.cuda.empty_cache()
    return out


GMF_TILES = (1, 2, 4)
GMF_STRATEGIES = (0, 1, 2)


def tune_gmf(funcs=tuple(SLOTS), orders=(0, 1, 2), precisions=("fp64", "mixed"), n=1 << 20, reps=5, seed=0):
    """per function, order and precision: the workgroup shape of gmf_hip_eval* (tiles 1 / 2 / 4), then the strategy of
    gmf_hip_eval_sorted* (0 / 1 / 2; 0 runs the tuned shape), over the general mix (matfun_amd.workloads.general).
    Returns {(func, order, precision): {"tiles": (best, {tiles: ms}), "strategy": (best, {strategy: ms})}}"""
    import torch
    from matfun_amd import workloads as W
    lib = matfun_amd.library()
    X = torch.from_numpy(W.general(n, seed=seed)).cuda()
    dal = torch.from_numpy(np.random.default_rng(seed).uniform(-1.0, 1.0, n)).cuda()
    st = torch.cuda.current_stream().cuda_stream
    nb = lib.gmf_hip_sort_workspace(n)
    work = torch.empty(nb, dtype=torch.uint8, device="cuda")
    out = {}
    for prec in precisions:
        mx = prec == "mixed"
        dt = torch.float32 if mx else torch.float64
        for order in orders:
            F = torch.empty(n * 9, dtype=dt, device="cuda")
            J = torch.empty(n * 81, dtype=dt, device="cuda") if order >= 1 else None
            H = torch.empty(n * 729, dtype=dt, device="cuda") if order >= 2 else None
            p = (F.data_ptr(), None if J is None else J.data_ptr(), None if H is None else H.data_ptr())
            for func in funcs:
                slot = SLOTS[func]
                if func == "pow_alpha":
                    plain = lib.gmf_hip_eval_pow_mixed if mx else lib.gmf_hip_eval_pow
                    srt = lib.gmf_hip_eval_pow_sorted_mixed if mx else lib.gmf_hip_eval_pow_sorted
                    call = lambda: plain(dal.data_ptr(), n, X.data_ptr(), *p, st)
                    scall = lambda: srt(dal.data_ptr(), n, X.data_ptr(), *p, work.data_ptr(), nb, st)
                else:
                    plain = lib.gmf_hip_eval_mixed if mx else lib.gmf_hip_eval
                    srt = lib.gmf_hip_eval_sorted_mixed if mx else lib.gmf_hip_eval_sorted
                    al = 0.3 if func == "pow" else 0.0
                    kind = matfun_amd.KINDS[func]
                    call = lambda: plain(kind, al, n, X.data_ptr(), *p, st)
                    scall = lambda: srt(kind, al, n, X.data_ptr(), *p, work.data_ptr(), nb, st)
                ms = _time_candidates(lambda t: lib.gmf_hip_set_tiles(slot, order, mx, t), call, GMF_TILES, reps)
                best = min(ms, key=ms.get)
                assert lib.gmf_hip_set_tiles(slot, order, mx, best) == 0
                sms = _time_candidates(lambda t: lib.gmf_hip_set_sort_strategy(slot, order, mx, t), scall,
                                       GMF_STRATEGIES, reps)
                sbest = min(sms, key=sms.get)
                assert lib.gmf_hip_set_sort_strategy(slot, order, mx, sbest) == 0
                out[(func, order, prec)] = {"tiles": (best, ms), "strategy": (sbest, sms)}
            del F, J, H
            torch.cuda.empty_cache()
    return out
