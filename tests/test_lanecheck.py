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
    for f in (lib.lanecheck_smf_pow, lib.lanecheck_smf_matrix_pow, lib.lanecheck_gmf_pow):
        f.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 4
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


def run_pow(fn, alphas, X, nf, order):
    X = np.ascontiguousarray(X, dtype=np.float64)
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    n = len(X)
    F = np.empty((n, nf))
    J = np.empty((n, nf * nf)) if order >= 1 else None
    H = np.empty((n, nf ** 3)) if order >= 2 else None
    p = lambda a: None if a is None else a.ctypes.data
    fn(al.ctypes.data, n, X.ctypes.data, F.ctypes.data, p(J), p(H))
    return F, J, H


def mixed_alphas(n, seed):
    """exponents uniform in [-1, 1] with the special values 0, +-1, 1/2 and a few large ones (Taylor radius 3/|al|)"""
    al = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    al[:8] = [0.0, 1.0, -1.0, 0.5, -0.5, 2.0, 40.0, -40.0]
    return al


def test_smf_per_matrix_exponents_vs_oracle(lane):
    """per-lane series coefficients (PA) against the oracle's per-matrix exponents, matrix layout, order 2"""
    o = Checker("oracle")
    A = W.spd(3000, seed=34)
    al = mixed_alphas(len(A), 35)
    al[40:] = np.where(np.abs(al[40:]) < 1e-3, 1e-3, al[40:])
    got = run_pow(lane.lanecheck_smf_matrix_pow, al, A, 9, 2)
    _, *ref = o.smf_matrix_alpha(al, A, 2)
    for name, x, y in zip("FJH", got, ref):
        e = normwise(x, y)
        assert e.max() <= 1e-12, (name, e.max(), int(np.argmax(e)), al[int(np.argmax(e))])


def test_smf_voigt_per_matrix_exponents_equal_uniform_calls(lane):
    """Voigt layout: each matrix's result with its own exponent equals the uniform-exponent path at that exponent
    (the two differ only in how the series coefficients binom(al, k) are rounded)"""
    v = W.voigt(W.spd(400, seed=36))
    al = np.repeat(np.array([0.3, -0.7, 0.0, 1.0, 2.5, 40.0, -1.0, 0.9]), 50)
    got = run_pow(lane.lanecheck_smf_pow, al, v, 6, 2)
    for a in np.unique(al):
        sel = al == a
        ref = run(lane.lanecheck_smf, "pow", a, v[sel], 6, 2)
        for name, x, y in zip("FJH", got, ref):
            e = normwise(x[sel], y)
            assert e.max() <= 1e-13, (a, name, e.max())


def test_gmf_per_matrix_exponents_vs_oracle(lane):
    o = Checker("oracle")
    G = W.general(3000, seed=37)
    al = mixed_alphas(len(G), 38)
    for order in (0, 1, 2):
        got = run_pow(lane.lanecheck_gmf_pow, al, G, 9, order)
        _, *ref = o.gmf_alpha(al, G, order)
        for name, x, y in zip("FJH", got[:order + 1], ref[:order + 1]):
            e = normwise(x, y)
            assert e.max() <= 1e-12, (order, name, e.max(), int(np.argmax(e)), al[int(np.argmax(e))])


def test_non_finite_exponent_gives_nan_for_its_matrix_only(lane):
    A = W.spd(64, seed=39)
    G = W.general(64, seed=40)
    al = np.full(64, 0.4)
    al[[3, 17, 50]] = [np.nan, np.inf, -np.inf]
    bad = np.zeros(64, bool)
    bad[[3, 17, 50]] = True
    for fn, X in ((lane.lanecheck_smf_matrix_pow, A), (lane.lanecheck_gmf_pow, G)):
        F, J, H = run_pow(fn, al, X, 9, 2)
        for x in (F, J, H):
            assert np.isnan(x[bad]).all()
        ref = run_pow(fn, np.full(64, 0.4), X, 9, 2)
        for x, y in zip((F, J, H), ref):
            assert np.array_equal(x[~bad], y[~bad], equal_nan=True)


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


def test_philox4x32_10_known_answers(lane):
    """the generator of the Newton driver's perturbations against the known-answer vectors published with Philox
    (Salmon et al., SC'11; Random123 kat_vectors: philox4x32 10 rounds)"""
    lane.lanecheck_philox.argtypes = [ctypes.c_void_p] * 3
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, want in kat:
        c, k, o = (np.array(ctr, np.uint32), np.array(key, np.uint32), np.zeros(4, np.uint32))
        lane.lanecheck_philox(c.ctypes.data, k.ctypes.data, o.ctypes.data)
        assert [int(x) for x in o] == want


def test_perturbation_factors_are_normal_and_symmetric(lane):
    lane.lanecheck_perturb.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_void_p]
    n = 200000
    sym = np.zeros(n, np.int32)
    sym[::2] = 1
    f = np.empty((n, 9))
    lane.lanecheck_perturb(12345, 3, n, sym.ctypes.data, 1.0, f.ctypes.data)
    z = f - 1.0
    zs = z[1::2].ravel()                                 # general matrices: 9 independent normals each
    assert abs(zs.mean()) < 0.01 and abs(zs.var() - 1.0) < 0.01
    assert abs(np.mean(zs ** 4) - 3.0) < 0.05           # (kurtosis of a normal)
    Z = z[::2].reshape(-1, 3, 3)                        # symmetric ones: (i, j) and (j, i) share the normal
    assert np.array_equal(Z, Z.transpose(0, 2, 1))
    f2 = np.empty((n, 9))
    lane.lanecheck_perturb(12345, 4, n, sym.ctypes.data, 1.0, f2.ctypes.data)
    assert abs(np.corrcoef(f[1::2, 0], f2[1::2, 0])[0, 1]) < 0.01    # iterations draw independent streams


def test_fused_hencky_lane_math_vs_oracle(lane):
    """the fused symmetric Hencky step (hencky_device.hpp: contraction in the Voigt basis) against the oracle's Hencky
    step (pinned by finite differences in tests/test_hencky_oracle.py) on the oracle's smf log of the same matrices"""
    from oracle.pyoracle import hencky
    f = lane.lanecheck_smf_hencky
    f.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 5
    A = W.spd(1500, seed=35)
    A[11, 5] += 0.1                                  # not symmetric: NaN results
    n = len(A)
    L, psi, g, K = np.empty((n, 9)), np.empty(n), np.empty((n, 9)), np.empty((n, 81))
    f(n, np.ascontiguousarray(A).ctypes.data, L.ctypes.data, psi.ctypes.data, g.ctypes.data, K.ctypes.data)
    _, rL, rJ, rH = Checker("oracle").smf_matrix("log", 0.0, A, 2)
    rpsi, rg, rK = hencky(rL, rJ, rH)
    for name, x, y in (("L", L, rL), ("psi", psi[:, None], rpsi[:, None]), ("g", g, rg), ("K", K, rK)):
        e = normwise(x, y)
        assert e.max() <= 1e-12, (name, e.max(), int(np.argmax(e)))
    assert np.isnan(K[11]).all() and not np.isnan(np.delete(K, 11, axis=0)).any()



@pytest.mark.parametrize("kind,alpha", [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 60.5)])
def test_gmf_real_series_loops_equal_the_complex_loop(lane, kind, alpha):
    """the real-arithmetic group series (gmf_series_real, run by the staged evaluation for wavefronts whose offsets
    are all real) against the complex loop on the general mix: sums within 1e-14 (normwise), and the complex loop's
    imaginary parts exactly zero for those lanes, so the real loop drops nothing"""
    lane.lanecheck_gmf_series_real.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    A = np.ascontiguousarray(W.general(20000, seed=41))
    for order in (0, 1, 2):
        cls = np.empty(len(A), dtype=np.int32)
        err = np.empty((len(A), 3))
        lane.lanecheck_gmf_series_real(KC[kind], alpha, order, len(A), A.ctypes.data, cls.ctypes.data,
                                       err.ctypes.data)
        counts = np.bincount(cls, minlength=3)
        assert counts[0] > 1000 and counts[1] > 100 and counts[2] > 100, counts     # (every class occurs)
        real = cls <= 1
        assert err[real, 2].max() == 0.0, (kind, order)
        assert err[cls == 0, 0].max() <= 1e-14, (kind, order, err[cls == 0, 0].max())
        assert err[real, 1].max() <= 1e-14, (kind, order, err[real, 1].max())


def test_gmf_staged_nodes_come_back_bitwise(lane):
    """the staged evaluation keeps the nodes in a real record plus the imaginary parts only where one is not +0
    (gmf_stage_nodes / gmf_unstage_nodes): every matrix of the general mix, and matrices with a negative-zero or NaN
    entry, get their nodes back bit for bit, with the series class in the flags"""
    lane.lanecheck_gmf_stage.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    A = np.ascontiguousarray(W.general(20000, seed=42))
    A[5] = [-0.0, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 3.0]
    A[6] = [np.nan, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 3.0]
    nbad, ncplx = ctypes.c_int64(), ctypes.c_int64()
    lane.lanecheck_gmf_stage(len(A), A.ctypes.data, ctypes.byref(nbad), ctypes.byref(ncplx))
    assert nbad.value == 0
    assert 1000 < ncplx.value < len(A) - 1000, ncplx.value          # (both layouts occur)
