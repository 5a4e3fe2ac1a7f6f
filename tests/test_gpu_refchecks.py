"""The reference's edge-case checks (tests/refchecks.py: c/tests/test_c.c, python/tests/test_gmf.py, the large-exponent
cases of python/tests/test_mpref.py against a 60-digit mpmath reference) on the GPU, through the C ABI."""
import pytest

from tests import refchecks as R

pytestmark = pytest.mark.gpu


class GpuEval:
    gpu = True

    def __init__(self, mixed=False):
        self.mixed = mixed

    def smf(self, kind, alpha, V, order):
        from tests.gpu_util import call
        rc, F, J, H = call("voigt", kind, alpha, V, order, mixed=self.mixed)
        assert rc == 0
        return F, J, H

    def smf_matrix(self, kind, alpha, A, order):
        from tests.gpu_util import call
        rc, F, J, H = call("matrix", kind, alpha, A, order, mixed=self.mixed)
        return F, J, H, rc

    def gmf(self, kind, alpha, A, order):
        from tests.gpu_util import call
        rc, F, J, H = call("general", kind, alpha, A, order, mixed=self.mixed)
        assert rc == 0
        return F, J, H


@pytest.fixture(scope="module")
def ev():
    return GpuEval()


@pytest.fixture(scope="module")
def mp_refs():
    pytest.importorskip("mpmath")
    return {(al, j): R.mp_reference_voigt("pow", al, v)
            for j, v in enumerate(R.large_alpha_inputs()) for al in R.LARGE_ALPHA}


@pytest.mark.parametrize("check", R.SMF_CHECKS + R.GMF_CHECKS, ids=lambda f: f.__name__)
def test_reference_check(ev, check):
    print(f"reference check on the GPU: {check.__name__}", flush=True)
    check(ev)


@pytest.mark.parametrize("deg", R.ROTATION_DEGREES)
def test_rotations_near_180(ev, deg):
    R.gmf_rotations_near_180(ev, deg)


def test_pow_large_alpha_against_mpmath(ev, mp_refs):
    R.pow_large_alpha(ev, mp_refs)


def test_against_high_precision_fixture(ev):
    """the GPU against the independent 60 / 120-digit reference (tests/golden/mpref.npz; tests/test_oracle_mpref.py
    pins the oracle to it at 1e-13): every class of both benchmark distributions, F / J / H, bound 1e-13 per
    matrix; the per-class maxima are printed"""
    import os
    import numpy as np
    from tests.test_oracle_mpref import GOLD, errors
    d = np.load(GOLD)
    worst = {}
    for fam, f in (("smf", lambda k, al, x: ev.smf(k, al, x, 2)), ("gmf", lambda k, al, x: ev.gmf(k, al, x, 2))):
        for label, k, e in errors(d, fam, f):
            key = f"{fam} {k} {label}"
            worst[key] = [max(a, b) for a, b in zip(worst.get(key, [0, 0, 0]), e)]
    for key, e in sorted(worst.items()):
        print(f"MPREF {key}: F {e[0]:.1e} J {e[1]:.1e} H {e[2]:.1e}", flush=True)
    bad = {k: e for k, e in worst.items() if max(e) > 1e-13}
    assert not bad, bad
