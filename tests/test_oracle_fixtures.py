"""The CPU oracle (oracle/, a scalar C restatement of the reference kernels) against the reference's own golden
vectors: testdata/cases.txt (smf, Voigt layout, 210 cases) and testdata/cases_general.txt (gmf, 204 cases), converted
losslessly to tests/golden/cases_*.npz.  Metric and per-case tolerances as the reference's python/tests/test_capi.py
:33-43 and test_gmf.py:27-37: max|x - ref| / max|ref| per case and output <= tol (1e-12, 1e-11, 1e-9)."""
import os

import numpy as np
import pytest

from oracle.pyoracle import Checker, available

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def maxrel(x, ref):
    return np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300)


@pytest.fixture(scope="module")
def oracle():
    return Checker("oracle")


def test_smf_fixtures(oracle):
    d = np.load(os.path.join(GOLD, "cases_smf.npz"))
    assert len(d["kind"]) == 210
    for i in range(len(d["kind"])):
        rc, F, J, H = oracle.smf(d["kind"][i], d["alpha"][i], d["v"][i][None], order=2)
        assert rc == 0
        for name, x, ref in (("F", F[0], d["F"][i]), ("J", J[0], d["J"][i]), ("H", H[0], d["H"][i])):
            assert maxrel(x, ref) <= d["tol"][i], (i, str(d["label"][i]), name, maxrel(x, ref))


def test_smf_matrix_layout_from_fixtures(oracle):
    """matrix layout == Voigt results times the exact weights 1 / 0.5 (c/tests/test_c.c:106-123, tolerance 0)"""
    d = np.load(os.path.join(GOLD, "cases_smf.npz"))
    vi = [0, 3, 4, 3, 1, 5, 4, 5, 2]
    w = np.array([1.0 if m % 4 == 0 else 0.5 for m in range(9)])
    for i in range(0, len(d["kind"]), 7):
        v = d["v"][i]
        A = v[vi]
        rc, Fm, Jm, Hm = oracle.smf_matrix(d["kind"][i], d["alpha"][i], A[None])
        _, Fv, Jv, Hv = oracle.smf(d["kind"][i], d["alpha"][i], v[None])
        Jv, Hv = Jv.reshape(6, 6), Hv.reshape(6, 6, 6)
        assert rc == 0
        assert np.array_equal(Fm[0], Fv[0][vi])
        assert np.array_equal(Jm[0].reshape(9, 9), Jv[np.ix_(vi, vi)] * w[None, :])
        assert np.array_equal(Hm[0].reshape(9, 9, 9), Hv[np.ix_(vi, vi, vi)] * w[None, :, None] * w[None, None, :])


def test_gmf_fixtures(oracle):
    g = np.load(os.path.join(GOLD, "cases_gmf.npz"))
    assert len(g["kind"]) == 204
    for i in range(len(g["kind"])):
        rc, F, J, H = oracle.gmf(g["kind"][i], g["alpha"][i], g["A"][i][None], order=2)
        assert rc == 0
        for name, x, ref in (("F", F[0], g["F"][i]), ("J", J[0], g["J"][i]), ("H", H[0], g["H"][i])):
            assert maxrel(x, ref) <= g["tol"][i], (i, str(g["label"][i]), name, maxrel(x, ref))


@pytest.mark.skipif(not available("ref"), reason="reference C library not built (oracle/_ref)")
def test_oracle_bitwise_equals_reference_scalar_on_config_samples():
    """the restatement is the same computation as the reference's scalar kernel: bitwise equal on seeded samples of
    the symmetric (SPD, with exact / near repeats) and general (pairs, shear, similarity) workload distributions"""
    import subprocess
    import sys
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r)
from oracle.pyoracle import Checker
from matfun_amd import workloads as W
o, r = Checker('oracle'), Checker('ref')
assert r.lib.smf_simd_name() == b'scalar'
A, G = W.spd(3000, seed=11), W.general(3000, seed=12)
for kind, al in (('exp', 0), ('log', 0), ('sqrt', 0), ('pow', 0.3), ('pow', 1.7)):
    for order in (0, 1, 2):
        for f, X in ((o.smf_matrix, A), (o.gmf, G)):
            g = r.smf_matrix if f == o.smf_matrix else r.gmf
            a, b = f(kind, al, X, order), g(kind, al, X, order)
            assert a[0] == b[0]
            for x, y in zip(a[1:], b[1:]):
                if x is not None:
                    assert np.array_equal(x, y, equal_nan=True), (kind, al, order)
print('ok')
""" % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MATFUN_SIMD="scalar", OMP_NUM_THREADS="4")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "ok" in out.stdout, out.stderr[-2000:]
