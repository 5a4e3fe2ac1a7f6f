"""smf_hip_eval / smf_hip_eval_matrix on the GPU against the CPU oracle (oracle/, bitwise equal to the reference C
kernel, see test_oracle_fixtures.py) and against the reference's golden vectors.

Tolerances: FP64, per matrix normwise relative error (||x - ref||_2 / ||ref||_2 over F, J, H separately)
<= 1e-12 on the seeded symmetric workload (SPD, 15% exact double, 5% triple, 10% 1e-8 gap); the reference's GPU
test bound max|x - ref| / max|ref| <= 1e-13 over a batch (cuda/tests/test_cuda_api.cu:76) on its SPD sample;
the golden vectors at their own per-case tolerances (1e-12 / 1e-11).
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import Checker
from matfun_amd import workloads as W
from tests.parity import maxrel, normwise

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
KINDS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 1.7)]
TOL = 1e-12


@pytest.fixture(scope="module")
def oracle():
    return Checker("oracle")


@pytest.fixture(scope="module")
def spd():
    return W.spd(3000, seed=101)


def _check(name, got, ref, tol):
    e = normwise(got, ref)
    worst = int(np.argmax(e))
    assert e.max() <= tol, f"{name}: max normwise error {e.max():.2e} at matrix {worst} (p99.9 {np.quantile(e, 0.999):.2e})"


@pytest.mark.parametrize("kind,alpha", KINDS)
@pytest.mark.parametrize("order", [0, 1, 2])
def test_matrix_layout_vs_oracle(oracle, spd, kind, alpha, order):
    from tests.gpu_util import call
    rc, F, J, H = call("matrix", kind, alpha, spd, order)
    orc, Fo, Jo, Ho = oracle.smf_matrix(kind, alpha, spd, order)
    assert rc == 0 and orc == 0
    _check(f"{kind} F", F, Fo, TOL)
    if order >= 1:
        _check(f"{kind} J", J, Jo, TOL)
    if order >= 2:
        _check(f"{kind} H", H, Ho, TOL)


@pytest.mark.parametrize("kind,alpha", KINDS)
@pytest.mark.parametrize("order", [0, 1, 2])
def test_voigt_layout_vs_oracle(oracle, spd, kind, alpha, order):
    from tests.gpu_util import call
    v = W.voigt(spd)
    rc, F, J, H = call("voigt", kind, alpha, v, order)
    orc, Fo, Jo, Ho = oracle.smf(kind, alpha, v, order)
    assert rc == 0 and orc == 0
    _check(f"{kind} F", F, Fo, TOL)
    if order >= 1:
        _check(f"{kind} J", J, Jo, TOL)
    if order >= 2:
        _check(f"{kind} H", H, Ho, TOL)


def test_golden_vectors():
    """the reference's 210 smf golden cases (Voigt, F/J/H) through the GPU, at their per-case tolerances
    (metric max|x - ref| / max|ref| per case and output, python/tests/test_capi.py:33-43)"""
    from tests.gpu_util import call
    d = np.load(os.path.join(GOLD, "cases_smf.npz"))
    for kind in (0, 1, 2, 3):
        for alpha in np.unique(d["alpha"][d["kind"] == kind]):
            sel = np.nonzero((d["kind"] == kind) & (d["alpha"] == alpha))[0]
            rc, F, J, H = call("voigt", int(kind), float(alpha), d["v"][sel], 2)
            assert rc == 0
            for j, i in enumerate(sel):
                for name, x, ref in (("F", F[j], d["F"][i]), ("J", J[j], d["J"][i]), ("H", H[j], d["H"][i])):
                    e = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
                    assert e <= d["tol"][i], (str(d["label"][i]), int(kind), float(alpha), name, e)


def test_reference_gpu_api_sample(oracle):
    """as cuda/tests/test_cuda_api.cu:25-80: n = 1000 (not a multiple of the block), A = B B^T + 0.5 I, matrix 17
    not symmetric -> NaN there only; all kinds, orders, both layouts; batch max-norm relative difference <= 1e-13"""
    from tests.gpu_util import call
    rng = np.random.default_rng(7)
    n = 1000
    B = rng.standard_normal((n, 3, 3))
    A = B @ np.swapaxes(B, 1, 2) + 0.5 * np.eye(3)
    A = A.reshape(n, 9).copy()
    v = W.voigt(A)
    A[17, 1] += 0.1
    for kind, alpha in KINDS:
        for order in (0, 1, 2):
            rc, F, J, H = call("matrix", kind, alpha, A, order)
            orc, Fo, Jo, Ho = oracle.smf_matrix(kind, alpha, A, order)
            assert rc == 0 and orc == -3
            assert np.isnan(F[17]).all() and np.isnan(Fo[17]).all()
            assert not np.isnan(np.delete(F, 17, axis=0)).any()
            e = maxrel(F, Fo)
            if order >= 1:
                assert np.isnan(J[17]).all()
                e = max(e, maxrel(J, Jo))
            if order >= 2:
                assert np.isnan(H[17]).all()
                e = max(e, maxrel(H, Ho))
            assert e <= 1e-13, (kind, alpha, order, e)
            rc, F, J, H = call("voigt", kind, alpha, v, order)
            orc, Fo, Jo, Ho = oracle.smf(kind, alpha, v, order)
            e = maxrel(F, Fo)
            if order >= 1:
                e = max(e, maxrel(J, Jo))
            if order >= 2:
                e = max(e, maxrel(H, Ho))
            assert rc == 0 and e <= 1e-13, (kind, alpha, order, e)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 1000])
def test_batch_sizes_and_tails(oracle, n):
    """partial wavefronts: the same results for every batch size (and nothing written past the batch)"""
    import torch
    from tests.gpu_util import call
    A = W.spd(n, seed=5 + n)
    rc, F, J, H = call("matrix", "log", 0.0, A, 2)
    _, Fo, Jo, Ho = oracle.smf_matrix("log", 0.0, A, 2)
    assert rc == 0
    for x, y in ((F, Fo), (J, Jo), (H, Ho)):
        assert normwise(x, y).max() <= TOL
    # guard region after the outputs stays untouched
    import matfun_amd
    lib = matfun_amd.library()
    dA = torch.from_numpy(A).cuda()
    dF = torch.full((n + 64, 9), 7.0, dtype=torch.float64, device="cuda")
    dJ = torch.full((n + 64, 81), 7.0, dtype=torch.float64, device="cuda")
    dH = torch.full((n + 64, 729), 7.0, dtype=torch.float64, device="cuda")
    assert lib.smf_hip_eval_matrix(1, 0.0, n, dA.data_ptr(), dF.data_ptr(), dJ.data_ptr(), dH.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    for t in (dF, dJ, dH):
        assert (t[n:] == 7.0).all()


def test_derivatives_vs_central_differences():
    """J and H of the GPU against central differences of the GPU's own F and J (c/tests/test_c.c:60-84: h = 1e-5,
    tol 1e-8), distinct eigenvalues and the double eigenvalue (2, 2, 5)"""
    from tests.gpu_util import call
    xs = np.array([[2.0, 1.5, 1.2, 0.3, -0.2, 0.1], [3.0, 3.0, 3.0, 1.0, 1.0, 1.0]])
    h = 1e-5
    for kind, alpha in KINDS:
        for x in xs:
            P = np.repeat(x[None], 13, axis=0)
            for p in range(6):
                P[1 + 2 * p, p] += h
                P[2 + 2 * p, p] -= h
            rc, F, J, H = call("voigt", kind, alpha, P, 2)
            assert rc == 0
            Jfd = np.stack([(F[1 + 2 * p] - F[2 + 2 * p]) / (2 * h) for p in range(6)], axis=1)
            Hfd = np.stack([(J[1 + 2 * p] - J[2 + 2 * p]).reshape(6, 6) / (2 * h) for p in range(6)], axis=2)
            assert maxrel(J[0].reshape(6, 6), Jfd) <= 1e-8, kind
            assert maxrel(H[0].reshape(6, 6, 6), Hfd) <= 1e-8, kind


def test_matrix_layout_is_voigt_times_weights():
    """smf_hip_eval_matrix == smf_hip_eval with the weights 1 / 0.5 of d v / d A_cd (c/tests/test_c.c:106-123) and the
    symmetries J_abcd = J_bacd = J_abdc = J_cdab, H likewise and (cd) <-> (ef) (:124-138)"""
    from tests.gpu_util import call
    A = W.spd(200, seed=9)
    vi = [0, 3, 4, 3, 1, 5, 4, 5, 2]
    w = np.array([1.0 if m % 4 == 0 else 0.5 for m in range(9)])
    for kind, alpha in KINDS:
        _, Fm, Jm, Hm = call("matrix", kind, alpha, A, 2)
        _, Fv, Jv, Hv = call("voigt", kind, alpha, W.voigt(A), 2)
        Jv, Hv = Jv.reshape(-1, 6, 6), Hv.reshape(-1, 6, 6, 6)
        assert normwise(Fm, Fv[:, vi]).max() <= 1e-15
        assert normwise(Jm, (Jv[:, vi][:, :, vi] * w).reshape(-1, 81)).max() <= 1e-15
        assert normwise(Hm, (Hv[:, vi][:, :, vi][:, :, :, vi] * w[:, None] * w[None, :]).reshape(-1, 729)).max() <= 1e-15
        J4 = Jm.reshape(-1, 3, 3, 3, 3)
        H6 = Hm.reshape(-1, 3, 3, 3, 3, 3, 3)
        s = np.abs(J4).max()
        assert np.abs(J4 - J4.transpose(0, 2, 1, 3, 4)).max() <= 1e-15 * s
        assert np.abs(J4 - J4.transpose(0, 1, 2, 4, 3)).max() <= 1e-15 * s
        assert np.abs(J4 - J4.transpose(0, 3, 4, 1, 2)).max() <= 4e-15 * s
        s = np.abs(H6).max()
        assert np.abs(H6 - H6.transpose(0, 2, 1, 3, 4, 5, 6)).max() <= 1e-15 * s
        assert np.abs(H6 - H6.transpose(0, 1, 2, 5, 6, 3, 4)).max() <= 1e-15 * s


def test_extreme_scales():
    """A scaled by 1e+-100, 1e+-250 (exact power-of-two scaling inside): sqrt(sA) = sqrt(s) sqrt(A), log(sA) =
    log(s) I + log(A), J_sqrt / sqrt(s), J_log / s, H_sqrt / s^1.5 (c/tests/test_c.c:152-174, tol 1e-14)"""
    from tests.gpu_util import call
    x = np.array([2.0, 1.5, 1.2, 0.3, -0.2, 0.1])
    _, F0, J0, H0 = call("voigt", "sqrt", 0.0, x[None], 2)
    _, L0, JL0, _ = call("voigt", "log", 0.0, x[None], 1)
    for k, s in enumerate((1e100, 1e-100, 1e250, 1e-250)):
        _, F, J, H = call("voigt", "sqrt", 0.0, (s * x)[None], 2)
        _, L, JL, _ = call("voigt", "log", 0.0, (s * x)[None], 1)
        e = max(maxrel(F, np.sqrt(s) * F0), maxrel(J, J0 / np.sqrt(s)), maxrel(JL, JL0 / s),
                maxrel(L, L0 + np.log(s) * np.array([1, 1, 1, 0, 0, 0.0])))
        if k < 2:
            e = max(e, maxrel(H, H0 / (s * np.sqrt(s))))
        assert e <= 1e-14, (s, e)


@pytest.mark.parametrize("mixed", [False, True])
def test_unaligned_output_pointers(oracle, mixed):
    """outputs that are not aligned to two elements take the one-element store path: same results"""
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    n = 130
    A = W.spd(n, seed=77)
    dt = torch.float32 if mixed else torch.float64
    fn = lib.smf_hip_eval_matrix_mixed if mixed else lib.smf_hip_eval_matrix
    dA = torch.from_numpy(A).cuda()
    bufs = [torch.full((n * k + 1,), 7.0, dtype=dt, device="cuda") for k in (9, 81, 729)]
    views = [b[1:] for b in bufs]                   # one element past an aligned allocation
    assert fn(1, 0.0, n, dA.data_ptr(), *[v.data_ptr() for v in views], torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    _, Fo, Jo, Ho = oracle.smf_matrix("log", 0.0, A, 2)
    tol = 1e-4 if mixed else 1e-12
    for v, ref in zip(views, (Fo, Jo, Ho)):
        assert normwise(v.double().cpu().numpy().reshape(n, -1), ref).max() <= tol
    for b in bufs:
        assert b[0].item() == 7.0
