"""The HIP kernels' per-lane math compiled for the CPU (tests/lanecheck: the __host__ __device__ functions of
matfun_amd/csrc, assembled into F / J / H the way the kernels assemble them) against the CPU oracle on the seeded
workloads and the reference's golden vectors.  Pins the kernels' algorithm and index maps without a GPU; the GPU
tests then pin the device build (ocml math, LDS staging, stores).  Tolerance: per-matrix normwise relative error
<= 1e-12 (the host build uses the same libm as the oracle, so most results agree to a few ulp)."""
import ctypes
import os

import numpy as np
import pytest

from oracle.pyoracle import Checker
from matfun_amd import workloads as W
from tests.parity import normwise

LIB = os.path.join(os.path.dirname(__file__), "lanecheck", "liblanecheck.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="tests/lanecheck not built (__graft_entry__.build)")
KINDS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 1.7), ("pow", 60.5)]
KC = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3}


@pytest.fixture(scope="module")
def lane():
    lib = ctypes.CDLL(LIB)
    for f in (lib.lanecheck_smf, lib.lanecheck_smf_matrix, lib.lanecheck_gmf):
        f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 4
    return lib


def run(fn, kind, alpha, X, nf, order):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = len(X)
    F = np.empty((n, nf))
    J = np.empty((n, nf * nf)) if order >= 1 else None
    H = np.empty((n, nf ** 3)) if order >= 2 else None
    p = lambda a: None if a is None else a.ctypes.data
    fn(KC[kind] if isinstance(kind, str) else int(kind), float(alpha), n, X.ctypes.data, F.ctypes.data, p(J), p(H))
    return F, J, H


@pytest.mark.parametrize("kind,alpha", KINDS)
def test_smf_lane_math_vs_oracle(lane, kind, alpha):
    o = Checker("oracle")
    v = W.voigt(W.spd(2000, seed=31))
    for order in (0, 1, 2):
        got = run(lane.lanecheck_smf, kind, alpha, v, 6, order)
        _, *ref = o.smf(kind, alpha, v, order)
        for name, x, y in zip("FJH", got[:order + 1], ref[:order + 1]):
            e = normwise(x, y)
            assert e.max() <= 1e-12, (kind, alpha, order, name, e.max(), int(np.argmax(e)))


@pytest.mark.parametrize("kind,alpha", KINDS)
def test_smf_matrix_layout_maps_vs_oracle(lane, kind, alpha):
    """the kernels' output index maps (distinct values -> matrix layout with weights 1, 1/2, 1/4)"""
    o = Checker("oracle")
    A = W.spd(500, seed=33)
    A[7, 1] += 0.1                                   # not symmetric: NaN results
    got = run(lane.lanecheck_smf_matrix, kind, alpha, A, 9, 2)
    _, *ref = o.smf_matrix(kind, alpha, A, 2)
    for name, x, y in zip("FJH", got, ref):
        e = normwise(x, y)
        assert e.max() <= 1e-12, (kind, alpha, name, e.max(), int(np.argmax(e)))
        assert np.isnan(x[7]).all()


@pytest.mark.parametrize("kind,alpha", KINDS)
def test_gmf_lane_math_vs_oracle(lane, kind, alpha):
    o = Checker("oracle")
    G = W.general(2000, seed=32)
    for order in (0, 1, 2):
        got = run(lane.lanecheck_gmf, kind, alpha, G, 9, order)
        _, *ref = o.gmf(kind, alpha, G, order)
        for name, x, y in zip("FJH", got[:order + 1], ref[:order + 1]):
            e = normwise(x, y)
            assert e.max() <= 1e-12, (kind, alpha, order, name, e.max(), int(np.argmax(e)))


def test_golden_vectors(lane):
    gold = os.path.join(os.path.dirname(__file__), "golden")
    d = np.load(os.path.join(gold, "cases_smf.npz"))
    for i in range(len(d["kind"])):
        F, J, H = run(lane.lanecheck_smf, d["kind"][i], d["alpha"][i], d["v"][i][None], 6, 2)
        for x, ref in ((F[0], d["F"][i]), (J[0], d["J"][i]), (H[0], d["H"][i])):
            assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) <= d["tol"][i], (i, str(d["label"][i]))
    g = np.load(os.path.join(gold, "cases_gmf.npz"))
    for i in range(len(g["kind"])):
        F, J, H = run(lane.lanecheck_gmf, g["kind"][i], g["alpha"][i], g["A"][i][None], 9, 2)
        for x, ref in ((F[0], g["F"][i]), (J[0], g["J"][i]), (H[0], g["H"][i])):
            assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) <= g["tol"][i], (i, str(g["label"][i]))
