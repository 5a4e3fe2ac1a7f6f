"""Mixed-precision form (FP64 inputs, FP32 outputs and assembly) against the FP64 oracle: per-matrix normwise
relative error <= 1e-5 for F and <= 1e-4 for J and H, per function, on the symmetric mix (SPD with exact / near
repeated eigenvalues, lam in [1e-2, 1e2]; for exp lam <= 60, since e^100 is beyond FP32) and the general mix
(similarity, complex pair, shear, 4 I + N)."""
import numpy as np
import pytest

from oracle.pyoracle import Checker
from matfun_amd import workloads as W
from tests.parity import normwise

pytestmark = pytest.mark.gpu
KINDS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 1.7)]
TOL = {"F": 1e-5, "J": 1e-4, "H": 1e-4}


@pytest.fixture(scope="module")
def oracle():
    return Checker("oracle")


@pytest.mark.parametrize("kind,alpha", KINDS)
@pytest.mark.parametrize("family", ["matrix", "voigt"])
def test_smf_mixed(oracle, kind, alpha, family):
    from tests.gpu_util import call
    # exp of eigenvalues up to 1e2 exceeds the FP32 range (e^88.7): the mixed form is tested on lam <= 60 for exp
    # (3997 = 62 x 64 + 29: partial last tile, vector-store tails)
    A = W.spd(3997, seed=303, log10_range=(-2.0, np.log10(60.0) if kind == "exp" else 2.0))
    X = A if family == "matrix" else W.voigt(A)
    rc, F, J, H = call(family, kind, alpha, X, 2, mixed=True)
    _, Fo, Jo, Ho = (oracle.smf_matrix if family == "matrix" else oracle.smf)(kind, alpha, X, 2)
    assert rc == 0
    for name, x, y in (("F", F, Fo), ("J", J, Jo), ("H", H, Ho)):
        e = normwise(x, y)
        assert e.max() <= TOL[name], (kind, alpha, name, e.max(), np.quantile(e, 0.999))


@pytest.mark.parametrize("kind,alpha", KINDS)
@pytest.mark.parametrize("order", [1, 2])
def test_gmf_mixed(oracle, kind, alpha, order):
    from tests.gpu_util import call
    G = W.general(4000, seed=404)
    rc, F, J, H = call("general", kind, alpha, G, order, mixed=True)
    _, Fo, Jo, Ho = oracle.gmf(kind, alpha, G, order)
    assert rc == 0
    for name, x, y in (("F", F, Fo), ("J", J, Jo), ("H", H, Ho))[:order + 1]:
        e = normwise(x, y)
        assert e.max() <= TOL[name], (kind, alpha, name, e.max(), np.quantile(e, 0.999))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 191, 255, 256, 257, 1000])
def test_gmf_mixed_order1_shared_series_equals_own_series(n):
    """The FP32 order-1 gmf kernels of log and powers run four tiles per workgroup and deal the group series out by
    length across the workgroup's lanes; a heterogeneous batch of one group computes every series in the lane of its
    matrix.  Same arithmetic, so bitwise the same F and J -- at sizes around the 64-matrix tile and the 256-matrix
    workgroup (wavefronts of a workgroup past the end take part with no matrices)."""
    import torch
    import matfun_amd
    from tests.gpu_util import call, call_pow
    G = W.general(n, seed=505 + n)
    for kind, alpha in (("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)):
        # (a batch's pow groups take per-matrix exponents: the counterpart is the per-matrix entry point)
        if kind == "pow":
            rc, F, J, _ = call_pow("general", np.full(n, alpha), G, 1, mixed=True)
        else:
            rc, F, J, _ = call("general", kind, alpha, G, 1, mixed=True)
        assert rc == 0
        b = matfun_amd.Batch([("gmf", kind, 1, n, alpha)])
        Fb, Jb, _ = b.eval(torch.from_numpy(G).cuda(), precision="mixed")
        torch.cuda.synchronize()
        for name, x, y in (("F", F, Fb), ("J", J, Jb)):
            y = y.reshape(n, -1).cpu().double().numpy()
            assert np.array_equal(x.reshape(n, -1), y, equal_nan=True), (kind, n, name)
        b.close()
