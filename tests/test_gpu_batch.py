"""Heterogeneous batches (mf_hip_plan_* / mf_hip_batch_eval*, an extension without counterpart in the reference API)
on the GPU.  A batch evaluates each matrix with the tile code of the homogeneous entry points, so the contract is
bitwise equality with them: every group's F / J / H equal smf_hip_eval_matrix / gmf_hip_eval on the same matrices
(pow: smf_hip_eval_matrix_pow / gmf_hip_eval_pow with the same exponents), in both precisions, for every launch
strategy.  The homogeneous entry points are pinned to the oracle by the other GPU tests.  Newton driver: the Philox
perturbation against its CPU build (the generator is pinned to the published known answers on the CPU), and a
graph-captured sequence against the same steps launched eagerly, over several replays with eager work between."""
import ctypes
import os

import numpy as np
import pytest

from matfun_amd import workloads as W

pytestmark = pytest.mark.gpu
KIND = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3}


def groups_all_classes(seed):
    """every (family, function, order), counts 0 .. 70 (tile boundaries 63 / 64 / 65 among them), shuffled"""
    rng = np.random.default_rng(seed)
    gs = []
    for fam in ("smf", "gmf"):
        for kind in ("exp", "log", "sqrt", "pow"):
            for order in (0, 1, 2):
                for count in rng.choice([0, 1, 5, 27, 63, 64, 65, 70], 3, replace=False):
                    gs.append((fam, kind, order, int(count), float(rng.uniform(-1.0, 1.0))))
    rng.shuffle(gs)
    return gs


def inputs(groups, seed):
    """A (N, 9): SPD matrices for smf groups (with repeated eigenvalues), general ones for gmf groups"""
    n = sum(g[3] for g in groups)
    S = W.spd(max(n, 1), seed=seed)
    G = W.general(max(n, 1), seed=seed + 1)
    A = np.empty((n, 9))
    lo = 0
    for g in groups:
        A[lo:lo + g[3]] = (S if g[0] == "smf" else G)[lo:lo + g[3]]
        lo += g[3]
    return A


def staged_group(fam, kind, order):
    """the groups that batch strategy 3 evaluates through the staged general-matrix passes"""
    return fam == "gmf" and kind != "exp" and order >= 1


def reference(groups, A, alpha, mixed, staged=False):
    """the homogeneous calls group by group; returns F, J, H stacked in the batch layout.  staged: the groups of
    staged_group through gmf_hip_eval_sorted / gmf_hip_eval_pow_sorted (at the strategy the library's table holds)"""
    from tests.gpu_util import call, call_pow, call_sorted
    Fs, Js, Hs = [], [], []
    lo = 0
    for fam, kind, order, count, al in groups:
        if count == 0:
            continue
        X = A[lo:lo + count]
        family = "matrix" if fam == "smf" else "general"
        a = None
        if kind == "pow":
            a = alpha[lo:lo + count] if alpha is not None else np.full(count, al)
        if staged and staged_group(fam, kind, order):
            rc, F, J, H = call_sorted(kind, 0.0, X, order, mixed=mixed, alphas=a)
        elif kind == "pow":
            rc, F, J, H = call_pow(family, a, X, order, mixed=mixed)
        else:
            rc, F, J, H = call(family, kind, 0.0, X, order, mixed=mixed)
        assert rc == 0
        Fs.append(F)
        if order >= 1:
            Js.append(J)
        if order >= 2:
            Hs.append(H)
        lo += count
    cat = lambda xs, w: np.concatenate(xs) if xs else np.zeros((0, w))
    return cat(Fs, 9), cat(Js, 81), cat(Hs, 729)


def run_batch(b, A, alpha, mixed):
    import torch
    dA = torch.from_numpy(A).cuda()
    dal = None if alpha is None else torch.from_numpy(alpha).cuda()
    F, J, H = b.eval(dA, dal, precision="mixed" if mixed else "fp64")
    torch.cuda.synchronize()
    g = lambda t: t.reshape(t.shape[0], -1).cpu().double().numpy()
    return g(F), g(J), g(H)


def same(x, y):
    return x.shape == y.shape and np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("per_matrix_alpha", [False, True])
def test_batch_equals_homogeneous_calls(mixed, per_matrix_alpha):
    import matfun_amd
    gs = groups_all_classes(601)
    A = inputs(gs, 602)
    alpha = np.random.default_rng(603).uniform(-1.0, 1.0, len(A)) if per_matrix_alpha else None
    b = matfun_amd.Batch(gs)
    assert (b.n, b.nclasses) == (len(A), 24)
    print(f"batch n={b.n} nj={b.nj} nh={b.nh} tiles={b.ntiles} mixed={mixed} alpha={per_matrix_alpha}", flush=True)
    got = run_batch(b, A, alpha, mixed)
    ref = reference(gs, A, alpha, mixed)
    for name, x, y in zip("FJH", got, ref):
        assert same(x, y), (name, x.shape, y.shape, np.nanmax(np.abs(x - y)) if x.shape == y.shape else None)
    b.close()


def test_every_strategy_gives_the_same_results():
    import matfun_amd
    gs = groups_all_classes(604)
    A = inputs(gs, 605)
    b = matfun_amd.Batch(gs)
    base = None
    for mode, per_cu, order in [(0, 0, 0), (0, 0, 1), (0, 2, 1), (0, 6, 0), (1, 0, 0), (1, 2, 0), (1, 8, 0),
                                (2, 0, 0), (2, 2, 0)]:
        print(f"strategy mode={mode} per_cu={per_cu} order={order}", flush=True)
        b.set_strategy(mode, per_cu, order)
        got = run_batch(b, A, None, False)
        if base is None:
            base = got
        for x, y in zip(got, base):
            assert same(x, y), (mode, per_cu, order)
    b.close()


@pytest.fixture
def staged_sorted():
    """every entry of gmf_hip_eval_sorted's strategy table set to the staged strategy (2), restored afterwards"""
    import matfun_amd
    lib = matfun_amd.library()
    keys = [(k, o, x) for k in range(5) for o in range(3) for x in range(2)]
    saved = {key: lib.gmf_hip_get_sort_strategy(*key) for key in keys}
    for key in keys:
        assert lib.gmf_hip_set_sort_strategy(*key, 2) == 0
    yield
    for key, v in saved.items():
        lib.gmf_hip_set_sort_strategy(*key, v)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("per_matrix_alpha", [False, True])
def test_staged_strategy_equals_the_staged_sorted_calls(mixed, per_matrix_alpha, staged_sorted):
    """strategy 3: the general-matrix log / sqrt / pow classes of order >= 1 bitwise gmf_hip_eval_sorted's staged
    evaluation of the same matrices (each matrix's result does not depend on the others in its call), every other
    class bitwise the homogeneous calls; groups spread over several tiles of one class among them"""
    import matfun_amd
    gs = groups_all_classes(612) + [("gmf", k, o, 150, 0.4) for k in ("log", "pow") for o in (1, 2)]
    A = inputs(gs, 613)
    alpha = np.random.default_rng(614).uniform(-1.0, 1.0, len(A)) if per_matrix_alpha else None
    b = matfun_amd.Batch(gs)
    b.set_strategy(3, 0, 0)
    print(f"staged batch n={b.n} tiles={b.ntiles} mixed={mixed} alpha={per_matrix_alpha}", flush=True)
    got = run_batch(b, A, alpha, mixed)
    ref = reference(gs, A, alpha, mixed, staged=True)
    for name, x, y in zip("FJH", got, ref):
        assert same(x, y), (name, x.shape, y.shape, np.nanmax(np.abs(x - y)) if x.shape == y.shape else None)
    b.close()


def test_tuner_keeps_results_and_picks_a_strategy():
    import torch
    import matfun_amd
    gs = groups_all_classes(606) * 4
    A = inputs(gs, 607)
    b = matfun_amd.Batch(gs)
    before = run_batch(b, A, None, False)
    dA = torch.from_numpy(A).cuda()
    res = b.tune(dA)
    print(f"tuned: {res}", flush=True)
    assert res["mode"] in (1, 2, 3) and res["us"] > 0     # (the fused kernel, mode 0, is not a tuning candidate)
    after = run_batch(b, A, None, False)
    if res["mode"] != 3:             # (strategy 3 agrees with the others to rounding; its own test pins it)
        for x, y in zip(before, after):
            assert same(x, y)
    b.close()


def test_perturbation_matches_cpu_build_and_keeps_symmetry():
    import torch
    import matfun_amd
    lane = ctypes.CDLL(os.path.join(os.path.dirname(__file__), "lanecheck", "liblanecheck.so"))
    lane.lanecheck_perturb.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_void_p]
    gs = groups_all_classes(608)
    A = inputs(gs, 609)
    b = matfun_amd.Batch(gs)
    dA = torch.from_numpy(A).cuda()
    b.perturb(dA, 1e-3, 987654321, 7)
    got = dA.cpu().numpy()
    sym = np.concatenate([np.full(g[3], 1 if g[0] == "smf" else 0, np.int32) for g in gs])
    f = np.empty((len(A), 9))
    lane.lanecheck_perturb(987654321, 7, len(A), sym.ctypes.data, 1e-3, f.ctypes.data)
    want = A * f
    # (Box-Muller's log / cos / sin: device and host libraries may differ in the last bit)
    assert np.all(np.abs(got - want) <= 4e-16 * np.abs(want))
    S = got[sym == 1].reshape(-1, 3, 3)
    assert np.array_equal(S, S.transpose(0, 2, 1))
    b.close()


@pytest.mark.parametrize("mixed,mode", [(False, 0), (True, 0), (False, 1), (False, 2), (False, 3), (True, 3)])
def test_graph_sequence_equals_eager_steps(mixed, mode):
    """the graph's iterations perturb A inside the evaluation's kernels (the staged classes' nodes pass in mode 3),
    the eager steps call the perturbation kernel (mf_hip_perturb) and then the evaluation: A and the outputs bitwise
    equal.  mode 2 forks the class kernels onto the plan's streams: the capture follows the fork and the join; mode 3
    runs the staged classes' passes through the plan's workspace"""
    import torch
    import matfun_amd
    gs = groups_all_classes(610)
    A0 = inputs(gs, 611)
    b = matfun_amd.Batch(gs)
    b.set_strategy(mode, 0 if mode else 6, 1)
    prec = "mixed" if mixed else "fp64"
    iters, sigma, seed = 20, 1e-3, 42
    dA0 = torch.from_numpy(A0).cuda()
    # eager: A <- A0, then iters x (perturb, eval)
    dA = dA0.clone()
    out_e = b.outputs(prec)
    for it in range(1, iters + 1):
        b.perturb(dA, sigma, seed, it)
        b.eval(dA, out=out_e, precision=prec)
    torch.cuda.synchronize()
    eager = [t.cpu().numpy() for t in out_e]
    eagerA = dA.cpu().numpy()
    # graph: captured once, replayed three times with eager work and synchronisation between the replays
    dA2 = torch.empty_like(dA0)
    out_g = b.outputs(prec)
    seq = b.sequence(dA0, dA2, out_g, iters, sigma, seed, precision=prec)
    for rep in range(3):
        print(f"graph replay {rep} mixed={mixed}", flush=True)
        for t in out_g:
            t.fill_(7.0)
        torch.cuda.synchronize()
        seq.launch()
        torch.cuda.synchronize()
        assert np.array_equal(dA2.cpu().numpy(), eagerA)
        for x, y in zip(out_g, eager):
            assert np.array_equal(x.cpu().numpy(), y, equal_nan=True), rep
        b.eval(dA0, out=out_e, precision=prec)           # (eager work between replays)
        torch.cuda.synchronize()
    seq.close()
    b.close()


def test_invalid_plans_and_calls():
    import torch
    import matfun_amd
    for bad in ([("smf", "log", 3, 4)], [("xmf", "log", 1, 4)], [("smf", "tan", 1, 4)], [("smf", "log", 1, -1)],
                [("gmf", "pow", 1, 4, float("nan"))]):
        with pytest.raises((ValueError, KeyError, TypeError)):
            matfun_amd.Batch(bad)
    b = matfun_amd.Batch([("smf", "log", 2, 10)])
    A = torch.from_numpy(W.spd(10, seed=612)).cuda()
    F, J, H = b.outputs()
    lib = matfun_amd.library()
    assert lib.mf_hip_batch_eval(b._h, A.data_ptr(), None, F.data_ptr(), J.data_ptr(), None, None) == -1
    assert lib.mf_hip_batch_eval(b._h, A.data_ptr(), None, F.data_ptr(), None, H.data_ptr(), None) == -1
    assert lib.mf_hip_batch_eval(None, A.data_ptr(), None, F.data_ptr(), J.data_ptr(), H.data_ptr(), None) == -1
    for bad in ((4, 0, 0), (-1, 0, 0), (1, 65, 0), (0, 0, 2)):
        with pytest.raises(ValueError):
            b.set_strategy(*bad)
    b.set_strategy(3, 0, 0)                          # (no staged class in this plan: no workspace needed)
    assert lib.mf_hip_batch_eval(b._h, A.data_ptr(), None, F.data_ptr(), J.data_ptr(), H.data_ptr(), None) == 0
    torch.cuda.synchronize()
    empty = matfun_amd.Batch([])
    assert empty.n == 0
    empty.eval(None, out=(torch.empty(0), torch.empty(0), torch.empty(0)))
    empty.set_strategy(3, 0, 0)
    b.close()
    empty.close()


def test_python_checks_reject_bad_tensors_before_launch():
    """a wrong dtype, a short or strided tensor, or a host tensor never reaches the C ABI"""
    import torch
    import matfun_amd
    b = matfun_amd.Batch([("smf", "log", 2, 10), ("gmf", "pow", 1, 3, 0.5)])
    A = torch.from_numpy(W.spd(13, seed=613)).cuda()
    F, J, H = b.outputs()
    for bad_call in (lambda: b.eval(A.float()), lambda: b.eval(A[:-1]), lambda: b.eval(A.cpu()),
                     lambda: b.eval(A, alpha=torch.zeros(12, dtype=torch.float64, device="cuda")),
                     lambda: b.eval(A, out=(F, J[:-1], H)), lambda: b.eval(A, out=(F, J, H), precision="mixed"),
                     lambda: b.perturb(A.reshape(13, 9).t(), 1e-3, 1, 1), lambda: b.tune(A, out=(F, None, H))):
        with pytest.raises(ValueError):
            bad_call()
    b.close()


def test_plan_outlives_its_sequences():
    """mf_hip_plan_destroy refuses (-1, nothing freed) while a sequence captured from the plan is alive, and
    Batch.close raises; once the sequence is destroyed the plan goes.  Nothing is launched."""
    import torch
    import matfun_amd
    gs = [("smf", "log", 2, 5), ("gmf", "exp", 1, 7)]
    b = matfun_amd.Batch(gs)
    dA0 = torch.from_numpy(inputs([(g[0], g[1], g[2], g[3], 0.0) for g in gs], 614)).cuda()
    dA = torch.empty_like(dA0)
    out = b.outputs()
    seq = b.sequence(dA0, dA, out, 2, 1e-3, 5)
    lib = matfun_amd.library()
    assert lib.mf_hip_plan_destroy(b._h) == -1            # (the C ABI itself refuses)
    with pytest.raises(RuntimeError):
        b.close()
    assert b._h                                           # (still usable)
    seq.close()
    b.close()
    assert not b._h
    assert lib.mf_hip_plan_destroy(None) == 0
