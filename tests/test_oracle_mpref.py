"""The oracle against an independent high-precision reference (tests/golden/mpref.npz, written by
tests/golden/make_mpref.py with mpmath: eigen-decompositions and central differences in 60 / 120 digits, none of the
divided-difference formulas involved) on samples of each class of the benchmark distributions: symmetric matrices
(generic, two equal eigenvalues, three equal, relative gap 1e-8) and general ones (similarity with cond(V) <= 1e3,
complex pair, defective simple shear, 4 I + N(0, 1)), exp / log / sqrt / pow 0.3, F / J / H.  This pins the
reference's algorithm itself (an accuracy defect shared by the reference and the oracle would show here); measured:
<= 3e-14 normwise per matrix on every class.  Bound 1e-13."""
import os

import numpy as np
import pytest

from oracle.pyoracle import Checker
from tests.parity import normwise

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mpref.npz")
NAMES = {0: "exp", 1: "log", 2: "pow", 3: "sqrt"}
TOL = 1e-13


def errors(d, fam, evalf):
    """per case (label, kind) -> [F, J, H] normwise errors of evalf(kind, alpha, x) against the fixture"""
    out = []
    for i in range(len(d[f"{fam}_kind"])):
        k, al = NAMES[int(d[f"{fam}_kind"][i])], float(d[f"{fam}_alpha"][i])
        x = d[f"{fam}_v" if fam == "smf" else "gmf_A"][i]
        F, J, H = evalf(k, al, x[None])
        e = [normwise(F, d[f"{fam}_F"][i][None])[0], normwise(J, d[f"{fam}_J"][i][None])[0],
             normwise(H, d[f"{fam}_H"][i][None])[0]]
        out.append((str(d[f"{fam}_label"][i]), k, e))
    return out


def test_oracle_against_high_precision():
    d = np.load(GOLD)
    o = Checker("oracle")
    assert len(d["smf_kind"]) == 96 and len(d["gmf_kind"]) == 80
    for fam, f in (("smf", lambda k, al, x: o.smf(k, al, x, 2)[1:]), ("gmf", lambda k, al, x: o.gmf(k, al, x, 2)[1:])):
        for label, k, e in errors(d, fam, f):
            assert max(e) <= TOL, (fam, label, k, e)


def test_fixture_reproduces_from_mpmath():
    """two fixture entries (a symmetric one with a relative eigenvalue gap of 1e-8, a defective general one)
    recomputed with the generator's own functions"""
    pytest.importorskip("mpmath")
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_mpref as M
    d = np.load(GOLD)
    i = int(np.nonzero((d["smf_label"] == "gap1e-8") & (d["smf_kind"] == 1))[0][0])
    F, J, H = M.smf_reference(("log", 0.0, d["smf_v"][i]))
    for x, y in ((F, d["smf_F"][i]), (J, d["smf_J"][i]), (H, d["smf_H"][i])):
        assert np.array_equal(x, y)
    j = int(np.nonzero((d["gmf_label"] == "shear") & (d["gmf_kind"] == 3))[0][0])
    F, J, H = M.gmf_reference(("sqrt", 0.0, d["gmf_A"][j]))
    for x, y in ((F, d["gmf_F"][j]), (J, d["gmf_J"][j]), (H, d["gmf_H"][j])):
        assert np.array_equal(x, y)
