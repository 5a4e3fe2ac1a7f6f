"""The symmetric order-2 kernels' H store tile (smf_hip_set_tile: 1, 10, 8, 5) and the general-matrix kernels'
workgroup shape (gmf_hip_set_tiles: 1, 2, 4 wavefronts sharing out their group series) change no result: every
choice gives bitwise the same F, J, H, for both layouts and precisions, every order, uniform and per-matrix exponents;
and the tuners (matfun_amd.tune) pick one of the choices and keep the results."""
import numpy as np
import pytest

from matfun_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["matrix", "voigt"])
@pytest.mark.parametrize("mixed", [False, True])
def test_every_tile_gives_the_same_results(layout, mixed):
    import matfun_amd
    from tests.gpu_util import call, call_pow
    lib = matfun_amd.library()
    # 997 = 15 x 64 + 37: the last tile's J block and last H pass end in the middle of a vector store (tails)
    A = W.spd(997, seed=701, log10_range=(-2.0, 1.7))
    X = A if layout == "matrix" else W.voigt(A)
    al = np.random.default_rng(702).uniform(-1.0, 1.0, len(X))
    cases = [(k, s) for s, k in enumerate(["exp", "log", "pow", "sqrt"])] + [("pow_alpha", 4)]
    try:
        for kind, slot in cases:
            base = None
            for tile in (1, 10, 8, 5):
                assert lib.smf_hip_set_tile(slot, layout == "matrix", mixed, tile) == 0
                assert lib.smf_hip_get_tile(slot, layout == "matrix", mixed) == tile
                print(f"tile {tile} {kind} {layout} mixed={mixed}", flush=True)
                if kind == "pow_alpha":
                    rc, *out = call_pow(layout, al, X, 2, mixed=mixed)
                else:
                    rc, *out = call(layout, kind, 0.3, X, 2, mixed=mixed)
                assert rc == 0
                if base is None:
                    base = out
                for x, y in zip(out, base):
                    assert np.array_equal(x, y, equal_nan=True), (kind, tile)
    finally:
        for s in range(5):
            lib.smf_hip_set_tile(s, layout == "matrix", mixed, 8 if mixed else 1)


def test_tile_arguments_checked():
    import matfun_amd
    lib = matfun_amd.library()
    assert lib.smf_hip_set_tile(0, 1, 0, 7) == -1
    assert lib.smf_hip_set_tile(5, 1, 0, 8) == -1
    assert lib.smf_hip_get_tile(-1, 1, 0) == -1


def test_tuner_picks_a_tile():
    import matfun_amd
    import matfun_amd.tune as T
    res = T.tune_smf_tiles(funcs=("log",), layouts=("matrix",), precisions=("fp64",), n=1 << 16, reps=2)
    best, ms = res[("log", "matrix", "fp64")]
    print("tuned", res, flush=True)
    assert best in (1, 10, 8, 5) and matfun_amd.library().smf_hip_get_tile(1, 1, 0) == best and len(ms) == 4


def _gmf(lib, X, kind, order, mixed, al=None):
    """gmf_hip_eval* (al: per-matrix exponents through gmf_hip_eval_pow*) on host array X; F, J, H as NumPy"""
    import torch
    import matfun_amd
    n = len(X)
    dt = torch.float32 if mixed else torch.float64
    dX = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    outs = [torch.full((n, k), 7.0, dtype=dt, device="cuda") if order >= o else None
            for o, k in ((0, 9), (1, 81), (2, 729))]
    p = [None if t is None else t.data_ptr() for t in outs]
    st = torch.cuda.current_stream().cuda_stream
    if al is not None:
        dal = torch.from_numpy(al).cuda()
        fn = lib.gmf_hip_eval_pow_mixed if mixed else lib.gmf_hip_eval_pow
        rc = fn(dal.data_ptr(), n, dX.data_ptr(), *p, st)
    else:
        fn = lib.gmf_hip_eval_mixed if mixed else lib.gmf_hip_eval
        rc = fn(matfun_amd.KINDS[kind], 0.3, n, dX.data_ptr(), *p, st)
    torch.cuda.synchronize()
    assert rc == 0
    return [None if t is None else t.cpu().numpy() for t in outs]


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_every_gmf_workgroup_shape_gives_the_same_results(order, mixed):
    import matfun_amd
    lib = matfun_amd.library()
    # 997 = 15 x 64 + 37: a workgroup of 2 or 4 tiles ends in a partial tile, the last one with wavefronts past the end
    G = W.general(997, seed=711)
    al = np.random.default_rng(712).uniform(-1.0, 1.0, len(G))
    cases = [(k, s) for s, k in enumerate(["exp", "log", "pow", "sqrt"])] + [("pow_alpha", 4)]
    saved = {s: lib.gmf_hip_get_tiles(s, order, mixed) for _, s in cases}
    try:
        for kind, slot in cases:
            base = None
            for tiles in (1, 2, 4):
                assert lib.gmf_hip_set_tiles(slot, order, mixed, tiles) == 0
                assert lib.gmf_hip_get_tiles(slot, order, mixed) == tiles
                print(f"gmf tiles {tiles} {kind} order {order} mixed={mixed}", flush=True)
                out = _gmf(lib, G, kind, order, mixed, al if kind == "pow_alpha" else None)
                if base is None:
                    base = out
                for x, y in zip(out, base):
                    if y is not None:
                        assert np.array_equal(x, y, equal_nan=True), (kind, tiles)
    finally:
        for s, v in saved.items():
            lib.gmf_hip_set_tiles(s, order, mixed, v)


def test_gmf_tiles_arguments_checked():
    import matfun_amd
    lib = matfun_amd.library()
    assert lib.gmf_hip_set_tiles(1, 1, 0, 3) == -1
    assert lib.gmf_hip_set_tiles(5, 1, 0, 2) == -1
    assert lib.gmf_hip_set_tiles(1, 3, 0, 2) == -1
    assert lib.gmf_hip_get_tiles(-1, 1, 0) == -1
    assert lib.gmf_hip_get_tiles(1, 1, 1) in (1, 2, 4)


def test_gmf_tuner_picks_a_choice():
    import matfun_amd
    import matfun_amd.tune as T
    lib = matfun_amd.library()
    saved = (lib.gmf_hip_get_tiles(1, 1, 1), lib.gmf_hip_get_sort_strategy(1, 1, 1))
    try:
        res = T.tune_gmf(funcs=("log",), orders=(1,), precisions=("mixed",), n=1 << 15, reps=2)
        r = res[("log", 1, "mixed")]
        print("tuned", res, flush=True)
        assert r["tiles"][0] in (1, 2, 4) and lib.gmf_hip_get_tiles(1, 1, 1) == r["tiles"][0]
        assert r["strategy"][0] in (0, 1, 2) and lib.gmf_hip_get_sort_strategy(1, 1, 1) == r["strategy"][0]
        assert len(r["tiles"][1]) == 3 and len(r["strategy"][1]) == 3
    finally:
        lib.gmf_hip_set_tiles(1, 1, 1, saved[0])
        lib.gmf_hip_set_sort_strategy(1, 1, 1, saved[1])
