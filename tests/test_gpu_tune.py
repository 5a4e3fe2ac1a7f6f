"""The symmetric order-2 kernels' H store tile (smf_hip_set_tile: 1, 10, 8, 5) changes no result:
every tile gives bitwise the same F, J, H, for both layouts and precisions, uniform and per-matrix exponents; and
the tuner (matfun_amd.tune) picks one of the tiles and keeps the results."""
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
