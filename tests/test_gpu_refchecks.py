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
