"""expand_packed (the Voigt / matrix layouts of the packed symmetric outputs) against the oracle's layouts: the packed
values taken from the oracle's Voigt results (J's upper triangle divided by the exact weights, H's p <= q entries)
and expanded again give the oracle's Voigt and matrix results -- bitwise for F and H; J to a few ulp, because the
oracle (like the reference) computes J_voigt[i][k] and J_voigt[k][i] separately, so its two triangles agree only to
rounding (the GPU kernel writes both from the same 21 values: bitwise there, tests/test_gpu_packed.py)."""
import numpy as np
import pytest

import matfun_amd
from oracle.pyoracle import Checker
from matfun_amd import workloads as W
from tests.parity import normwise


@pytest.mark.parametrize("kind,alpha", [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)])
def test_expand_packed_gives_the_api_layouts(kind, alpha):
    o = Checker("oracle")
    A = W.spd(200, seed=801)
    v = W.voigt(A)
    _, Fv, Jv, Hv = o.smf(kind, alpha, v, 2)
    _, Fm, Jm, Hm = o.smf_matrix(kind, alpha, A, 2)
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    iu = [(i, k) for i in range(6) for k in range(i, 6)]
    Jp = np.stack([Jv.reshape(-1, 6, 6)[:, i, k] / w[k] for i, k in iu], axis=1)
    Hp = np.concatenate([np.stack([Hv.reshape(-1, 6, 6, 6)[:, r, p, q] for p, q in iu], axis=1) for r in range(6)],
                        axis=1)
    F2, J2, H2 = matfun_amd.expand_packed(Fv, Jp, Hp, "voigt")
    assert np.array_equal(F2, Fv) and np.array_equal(H2.reshape(-1, 216), Hv)
    assert normwise(J2.reshape(-1, 36), Jv).max() <= 1e-15
    F3, J3, H3 = matfun_amd.expand_packed(Fv, Jp, Hp, "matrix")
    assert np.array_equal(F3.reshape(-1, 9), Fm) and np.array_equal(H3.reshape(-1, 729), Hm)
    assert normwise(J3.reshape(-1, 81), Jm).max() <= 1e-15
