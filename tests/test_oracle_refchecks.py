"""The reference's edge-case checks (tests/refchecks.py: c/tests/test_c.c, python/tests/test_gmf.py, the large-exponent
cases of python/tests/test_mpref.py against a 60-digit mpmath reference) on the CPU oracle."""
import numpy as np
import pytest

from oracle.pyoracle import Checker
from tests import refchecks as R


class OracleEval:
    gpu = False

    def __init__(self):
        self.c = Checker("oracle")

    def smf(self, kind, alpha, V, order):
        rc, F, J, H = self.c.smf(kind, alpha, V, order)
        assert rc == 0
        return F, J, H

    def smf_matrix(self, kind, alpha, A, order):
        rc, F, J, H = self.c.smf_matrix(kind, alpha, A, order)
        assert rc in (0, -3)
        return F, J, H, rc

    def gmf(self, kind, alpha, A, order):
        rc, F, J, H = self.c.gmf(kind, alpha, A, order)
        assert rc == 0
        return F, J, H


@pytest.fixture(scope="module")
def ev():
    return OracleEval()


@pytest.fixture(scope="module")
def mp_refs():
    pytest.importorskip("mpmath")
    return {(al, j): R.mp_reference_voigt("pow", al, v)
            for j, v in enumerate(R.large_alpha_inputs()) for al in R.LARGE_ALPHA}


@pytest.mark.parametrize("check", R.SMF_CHECKS + R.GMF_CHECKS, ids=lambda f: f.__name__)
def test_reference_check(ev, check):
    check(ev)


@pytest.mark.parametrize("deg", R.ROTATION_DEGREES)
def test_rotations_near_180(ev, deg):
    R.gmf_rotations_near_180(ev, deg)


def test_pow_large_alpha_against_mpmath(ev, mp_refs):
    R.pow_large_alpha(ev, mp_refs)


def test_a_planted_mistake_fails():
    """the checks detect a wrong result: the oracle's log with J transposed in its index pairs fails the central
    differences (gmf) and the exp/log inverse check fails for a perturbed log"""
    class Broken(OracleEval):
        def gmf(self, kind, alpha, A, order):
            F, J, H = super().gmf(kind, alpha, A, order)
            if J is not None and kind == "log":
                J = J.reshape(-1, 9, 9).transpose(0, 2, 1).reshape(-1, 81)
            return F, J, H

        def smf(self, kind, alpha, V, order):
            F, J, H = super().smf(kind, alpha, V, order)
            return (F * (1 + 1e-12) if kind == "log" else F), J, H
    b = Broken()
    with pytest.raises(AssertionError):
        R.gmf_central_differences(b)
    with pytest.raises(AssertionError):
        R.smf_exp_log(b)
