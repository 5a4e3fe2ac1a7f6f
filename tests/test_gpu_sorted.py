"""gmf_hip_eval_sorted* (general matrices evaluated in the order of their series lengths, an extension without
counterpart in the reference API) on the GPU, for each strategy (gmf_hip_set_sort_strategy): 0 (unsorted) and 1 (the
tile code of gmf_hip_eval over the length-sorted permutation) must give gmf_hip_eval / gmf_hip_eval_pow's results
bitwise (the other GPU tests pin those to the oracle); 2 (staged: nodes, series in length order -- in real arithmetic
for wavefronts whose group offsets are all real --, the rest) the same up to rounding -- per matrix normwise <= 1e-13
plus 8 times gmf_hip_eval's own distance to the oracle on that matrix (the two forms round differently, and a matrix
whose derivatives cancel amplifies both alike), and NaN exactly where gmf_hip_eval has NaN -- for every kind, order
and precision, at tile edges, with workspaces that hold only part of the batch, and with NaN results only for the
matrices outside the domain."""
import numpy as np
import pytest

from matfun_amd import workloads as W

pytestmark = pytest.mark.gpu
KINDS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 1.7)]


def run(X, kind, alpha, order, mixed, sorted_, work_bytes=None, alphas=None):
    """one call on host array X (n, 9); returns F, J, H as NumPy (J / H None below the order)"""
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    dt = torch.float32 if mixed else torch.float64
    n = len(X)
    dX = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    outs = [torch.full((n, k), 7.0, dtype=dt, device="cuda") if order >= o else None
            for o, k in ((0, 9), (1, 81), (2, 729))]
    p = lambda t: None if t is None else t.data_ptr()
    st = torch.cuda.current_stream().cuda_stream
    dal = torch.from_numpy(np.ascontiguousarray(alphas)).cuda() if alphas is not None else None
    if sorted_:
        nb = lib.gmf_hip_sort_workspace(n) if work_bytes is None else work_bytes
        work = torch.empty(max(nb, 1), dtype=torch.uint8, device="cuda")
        if dal is not None:
            fn = lib.gmf_hip_eval_pow_sorted_mixed if mixed else lib.gmf_hip_eval_pow_sorted
            rc = fn(p(dal), n, p(dX), *map(p, outs), work.data_ptr(), nb, st)
        else:
            fn = lib.gmf_hip_eval_sorted_mixed if mixed else lib.gmf_hip_eval_sorted
            rc = fn(matfun_amd.KINDS[kind], alpha, n, p(dX), *map(p, outs), work.data_ptr(), nb, st)
    else:
        if dal is not None:
            fn = lib.gmf_hip_eval_pow_mixed if mixed else lib.gmf_hip_eval_pow
            rc = fn(p(dal), n, p(dX), *map(p, outs), st)
        else:
            fn = lib.gmf_hip_eval_mixed if mixed else lib.gmf_hip_eval
            rc = fn(matfun_amd.KINDS[kind], alpha, n, p(dX), *map(p, outs), st)
    torch.cuda.synchronize()
    assert rc == 0, rc
    return [None if t is None else t.cpu().numpy() for t in outs]


def oracle_of(X, kind, alpha, order, alphas=None):
    """the oracle's F, J, H (J / H None below the order) for the same call"""
    from oracle.pyoracle import Checker
    o = Checker("oracle")
    _, *r = o.gmf_alpha(alphas, X, order) if alphas is not None else o.gmf(kind, alpha, X, order)
    return r


def same(a, b, strategy=1, ref=None):
    """bitwise equal (strategies 0 and 1), or equal up to rounding (strategy 2: NaN in the same matrices, normwise
    <= 1e-13 per matrix, plus 8 times b's distance to the oracle's results ref on that matrix when ref is given)"""
    from tests.parity import normwise
    for k, (x, y) in enumerate(zip(a, b)):
        if y is None:
            continue
        if strategy != 2:
            assert np.array_equal(x, y, equal_nan=True)
        else:
            assert np.array_equal(np.isnan(x), np.isnan(y))
            e = normwise(x, y)
            bound = 1e-13 + (0.0 if ref is None else 8.0 * np.nan_to_num(normwise(y, ref[k]), posinf=0.0))
            worst = int(np.argmax(e - bound))
            assert (e <= bound).all(), ("FJH"[k], worst, e[worst], bound[worst])


STRATEGIES = [0, 1, 2]


@pytest.fixture(params=STRATEGIES)
def strategy(request):
    """every table entry set to one strategy for the test, the table restored afterwards"""
    import matfun_amd
    lib = matfun_amd.library()
    keys = [(k, o, x) for k in range(5) for o in range(3) for x in range(2)]
    saved = {key: lib.gmf_hip_get_sort_strategy(*key) for key in keys}
    for key in keys:
        assert lib.gmf_hip_set_sort_strategy(*key, request.param) == 0
    yield request.param
    for key, v in saved.items():
        lib.gmf_hip_set_sort_strategy(*key, v)


@pytest.fixture(scope="module")
def gen():
    return W.general(3000, seed=701)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("kind,alpha", KINDS)
def test_sorted_equals_gmf_hip_eval(gen, kind, alpha, order, mixed, strategy):
    print(f"sorted vs plain: {kind} {alpha} order {order} mixed {mixed} strategy {strategy}", flush=True)
    ref = oracle_of(gen, kind, alpha, order) if strategy == 2 else None
    same(run(gen, kind, alpha, order, mixed, True), run(gen, kind, alpha, order, mixed, False), strategy, ref)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 129, 1000])
def test_tile_edges(gen, n, strategy):
    X = gen[:n]
    for kind, alpha in (("log", 0.0), ("pow", 0.3)):
        ref = oracle_of(X, kind, alpha, 2) if strategy == 2 else None
        same(run(X, kind, alpha, 2, False, True), run(X, kind, alpha, 2, False, False), strategy, ref)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("order", [1, 2])
def test_per_matrix_exponents(gen, order, mixed, strategy):
    al = np.random.default_rng(702).uniform(-1.0, 1.0, len(gen))
    al[:4] = [0.0, 1.0, -1.0, 0.5]
    ref = oracle_of(gen, "pow", 0.0, order, alphas=al) if strategy == 2 else None
    same(run(gen, "pow", 0.0, order, mixed, True, alphas=al), run(gen, "pow", 0.0, order, mixed, False, alphas=al),
         strategy, ref)


def test_workspace_holding_part_of_the_batch(gen, strategy):
    """a workspace for 64, 100 or 1000 matrices runs the 3000 in parts (a staged strategy falls back to the permuted
    one there, bitwise): the same results as one pass"""
    import matfun_amd
    lib = matfun_amd.library()
    ref = run(gen, "log", 0.0, 2, False, False)
    for m in (64, 100, 1000):
        nb = lib.gmf_hip_sort_workspace(m)
        same(run(gen, "log", 0.0, 2, False, True, work_bytes=nb), ref, 1 if strategy == 2 else strategy)


def test_staged_parts_equal_one_pass():
    """the staged strategy in parts of 2^15 matrices (the smallest it takes) equals one pass bitwise: a matrix's results
    do not depend on the part it runs in"""
    import matfun_amd
    lib = matfun_amd.library()
    G = W.general(70000, seed=705)
    saved = lib.gmf_hip_get_sort_strategy(1, 2, 0)
    lib.gmf_hip_set_sort_strategy(1, 2, 0, 2)
    try:
        one = run(G, "log", 0.0, 2, False, True)
        per = 5 + 80 + 144 + 19 * 16
        parts = run(G, "log", 0.0, 2, False, True, work_bytes=lib.gmf_hip_sort_workspace(0) + (1 << 15) * per + 4096)
        same(parts, one, 1)
    finally:
        lib.gmf_hip_set_sort_strategy(1, 2, 0, saved)


def test_staged_results_do_not_depend_on_the_order_of_the_batch():
    """the staged strategy's results per matrix are bitwise the same whatever the batch's order (the order within a
    sort key comes from atomics, so it changes from run to run as well): the batch and a shuffled copy, every kind,
    orders 1 and 2, FP64 and mixed"""
    import matfun_amd
    lib = matfun_amd.library()
    G = W.general(6000, seed=706)
    perm = np.random.default_rng(707).permutation(len(G))
    keys = [(k, o, x) for k in range(5) for o in range(3) for x in range(2)]
    saved = {key: lib.gmf_hip_get_sort_strategy(*key) for key in keys}
    for key in keys:
        assert lib.gmf_hip_set_sort_strategy(*key, 2) == 0
    try:
        for kind, alpha in KINDS:
            for order in (1, 2):
                for mixed in (False, True):
                    print(f"staged, shuffled batch: {kind} {alpha} order {order} mixed {mixed}", flush=True)
                    a = run(G, kind, alpha, order, mixed, True)
                    b = run(G[perm], kind, alpha, order, mixed, True)
                    for x, y in zip(a, b):
                        if y is not None:
                            assert np.array_equal(x[perm], y, equal_nan=True), (kind, alpha, order, mixed)
    finally:
        for key, v in saved.items():
            lib.gmf_hip_set_sort_strategy(*key, v)


def test_strategy_table_rejects_invalid_entries():
    import matfun_amd
    lib = matfun_amd.library()
    assert lib.gmf_hip_set_sort_strategy(5, 2, 0, 1) == -1
    assert lib.gmf_hip_set_sort_strategy(1, 3, 0, 1) == -1
    assert lib.gmf_hip_set_sort_strategy(1, 2, 0, 3) == -1
    assert lib.gmf_hip_get_sort_strategy(-1, 0, 0) == -1
    for k in range(5):
        for o in range(3):
            for x in range(2):
                assert lib.gmf_hip_get_sort_strategy(k, o, x) in (0, 1, 2)


def test_nan_only_outside_the_domain(strategy):
    """real eigenvalues <= 0 give NaN for log / sqrt / pow in their matrix only, wherever the sort puts it"""
    A = 1.5 * np.eye(3) + 0.3 * np.random.default_rng(703).standard_normal((200, 3, 3))
    A[[4, 77, 150]] = np.diag([-1.0, 2.0, 3.0])
    A = A.reshape(-1, 9)
    for kind in ("log", "sqrt"):
        F, J, H = run(A, kind, 0.0, 2, False, True)
        bad = np.isnan(F).any(axis=1)
        assert np.nonzero(bad)[0].tolist() == [4, 77, 150]
        ref = oracle_of(A, kind, 0.0, 2) if strategy == 2 else None
        same((F, J, H), run(A, kind, 0.0, 2, False, False), strategy, ref)


def test_rejections():
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    X = torch.from_numpy(W.general(100, seed=704)).cuda()
    F = torch.empty((100, 9), dtype=torch.float64, device="cuda")
    small = lib.gmf_hip_sort_workspace(0)                       # (the fixed tables only: no matrix)
    w = torch.empty(lib.gmf_hip_sort_workspace(100), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    assert lib.gmf_hip_eval_sorted(1, 0.0, 100, X.data_ptr(), F.data_ptr(), None, None, None, 0, st) == -1
    assert lib.gmf_hip_eval_sorted(1, 0.0, 100, X.data_ptr(), F.data_ptr(), None, None, w.data_ptr(), small, st) == -1
    assert lib.gmf_hip_eval_sorted(2, -1.0, 100, X.data_ptr(), F.data_ptr(), None, None, w.data_ptr(), w.numel(),
                                   st) == -1
    assert lib.gmf_hip_eval_pow_sorted(None, 100, X.data_ptr(), F.data_ptr(), None, None, w.data_ptr(), w.numel(),
                                       st) == -1
    assert lib.gmf_hip_eval_sorted(1, 0.0, 0, None, None, None, None, None, 0, st) == 0


def test_python_front_end_sort_flag(gen, strategy):
    import matfun_amd
    ref = oracle_of(gen, "log", 0.0, 2) if strategy == 2 else None
    for prec in ("fp64", "mixed"):
        a = matfun_amd.gmf("log", gen.reshape(-1, 3, 3), order=2, precision=prec, sort=True)
        b = matfun_amd.gmf("log", gen.reshape(-1, 3, 3), order=2, precision=prec)
        same(a, b, strategy, ref)


def test_full_size_general_workload_against_the_oracle(strategy):
    """the general-matrix configuration at full size, n = 2^20, log and pow(0.3) at order 2 through the sorted path:
    4096 rows spread over the batch plus the last 64 checked against gmf_hip_eval on the same rows (bitwise) and
    against the oracle (tests.parity.spread_check: per matrix max(1e-12, 32 s), s the reference's own SIMD-vs-scalar
    difference on that matrix, which reaches 1e-11 .. 1e-10 on the general mix's ill-conditioned tail)"""
    import torch
    import matfun_amd
    from oracle.pyoracle import Checker, available
    from tests.parity import spread_check
    lib = matfun_amd.library()
    ref = Checker("ref") if available("ref") else None
    ref = None if ref is None or ref.name().endswith("(scalar)") else ref
    n = 1 << 20
    G = W.general(n, seed=0)
    dG = torch.from_numpy(G).cuda()
    rows = np.unique(np.concatenate([np.linspace(0, n - 1, 4096).astype(np.int64), np.arange(n - 64, n)]))
    drows = torch.from_numpy(rows).cuda()
    nb = lib.gmf_hip_sort_workspace(n)
    work = torch.empty(nb, dtype=torch.uint8, device="cuda")
    F = torch.empty((n, 9), dtype=torch.float64, device="cuda")
    J = torch.empty((n, 81), dtype=torch.float64, device="cuda")
    H = torch.empty((n, 729), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    oracle = Checker("oracle")
    for kind, alpha in (("log", 0.0), ("pow", 0.3)):
        print(f"full size sorted gmf {kind} order 2, n = {n}", flush=True)
        k = matfun_amd.KINDS[kind]
        assert lib.gmf_hip_eval_sorted(k, alpha, n, dG.data_ptr(), F.data_ptr(), J.data_ptr(), H.data_ptr(),
                                       work.data_ptr(), nb, st) == 0
        torch.cuda.synchronize()
        got = [t.index_select(0, drows).cpu().numpy() for t in (F, J, H)]
        # the same rows through gmf_hip_eval (the unsorted kernel at the same order): bitwise equal
        from tests.gpu_util import call
        rc, Fp, Jp, Hp = call("general", kind, alpha, G[rows], 2)
        assert rc == 0
        _, Fo, Jo, Ho = oracle.gmf(kind, alpha, G[rows], 2)
        same(got, (Fp, Jp, Hp), strategy, (Fo, Jo, Ho))
        rr = ref.gmf(kind, alpha, G[rows], 2) if ref is not None else (0, None, None, None)
        for name, x, y, z in zip("FJH", got, (Fo, Jo, Ho), rr[1:]):
            spread_check(f"full size sorted {kind} {alpha} o2 {name}", x, y, z)
