"""The benchmark configurations at their full sizes on the GPU, checked through a sample of rows against the oracle:
4096 rows spread evenly over the batch plus the last 64 (at n = 2^22 the order-2 H rows of the last 1.2 M matrices
start past 2^31 elements, and every output region is larger than 2^32 bytes).  The sizes are those bench.py
measures: symmetric matrices of the headline distribution at n = 2^22 (FP64, order 2, matrix layout), the mixed form
with one exponent per matrix at n = 2^22, general matrices at n = 2^20 (FP64, order 2).  Tolerances as the smaller
GPU tests: smf FP64 <= 1e-12 per matrix; mixed <= 1e-5 (F) / 1e-4 (J, H); gmf FP64 as tests/test_gpu_gmf.py (the
general mix's ill-conditioned tail, where two builds of the reference itself differ by up to 8e-11: per matrix
max(1e-12, 32 s) with s that difference, tests.parity.spread_check)."""
import numpy as np
import pytest

from matfun_amd import workloads as W
from tests.parity import normwise

pytestmark = pytest.mark.gpu


def sample_rows(n):
    return np.unique(np.concatenate([np.linspace(0, n - 1, 4096).astype(np.int64), np.arange(n - 64, n)]))


def picked(tensors, rows):
    import torch
    d = torch.from_numpy(rows).cuda()
    return [t.index_select(0, d).double().cpu().numpy() for t in tensors]


def buffers(n, dt):
    import torch
    return [torch.empty((n, k), dtype=dt, device="cuda") for k in (9, 81, 729)]


@pytest.mark.parametrize("kind,alpha", [("log", 0.0), ("pow", 0.3)])
def test_smf_fp64_order2_at_2_22(kind, alpha):
    import torch
    import matfun_amd
    from oracle.pyoracle import Checker
    n = 1 << 22
    print(f"full size smf {kind} FP64 order 2, n = {n}", flush=True)
    A = W.spd(n, seed=0)
    dA = torch.from_numpy(A).cuda()
    out = buffers(n, torch.float64)
    lib = matfun_amd.library()
    assert lib.smf_hip_eval_matrix(matfun_amd.KINDS[kind], alpha, n, dA.data_ptr(), *[t.data_ptr() for t in out],
                                   torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    rows = sample_rows(n)
    got = picked(out, rows)
    _, Fo, Jo, Ho = Checker("oracle").smf_matrix(kind, alpha, A[rows], 2)
    for name, x, y in zip("FJH", got, (Fo, Jo, Ho)):
        e = normwise(x, y)
        assert e.max() <= 1e-12, (name, e.max(), rows[np.argmax(e)])


def test_smf_mixed_per_matrix_exponents_at_2_22():
    import torch
    import matfun_amd
    from oracle.pyoracle import Checker
    n = 1 << 22
    print(f"full size smf pow (exponent per matrix) mixed order 2, n = {n}", flush=True)
    A = W.spd(n, seed=1)
    al = np.random.default_rng(2).uniform(-1.0, 1.0, n)
    dA, dal = torch.from_numpy(A).cuda(), torch.from_numpy(al).cuda()
    out = buffers(n, torch.float32)
    lib = matfun_amd.library()
    assert lib.smf_hip_eval_matrix_pow_mixed(dal.data_ptr(), n, dA.data_ptr(), *[t.data_ptr() for t in out],
                                             torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    rows = sample_rows(n)
    got = picked(out, rows)
    _, Fo, Jo, Ho = Checker("oracle").smf_matrix_alpha(al[rows], A[rows], 2)
    nA = np.linalg.norm(A[rows], axis=1)
    for name, x, y, tol, o in zip("FJH", got, (Fo, Jo, Ho), (1e-5, 1e-4, 1e-4), (0, 1, 2)):
        # outputs that vanish identically (J and H at alpha = 0, H at alpha = 1) absolutely against |F| / |A|^order
        den = np.maximum(np.linalg.norm(y, axis=1), np.linalg.norm(Fo, axis=1) / nA ** o)
        e = np.linalg.norm(x - y, axis=1) / den
        assert e.max() <= tol, (name, e.max(), rows[np.argmax(e)], al[rows][np.argmax(e)])


def test_gmf_fp64_order2_at_2_20():
    import torch
    import matfun_amd
    from oracle.pyoracle import Checker, available
    from tests.parity import spread_check
    n = 1 << 20
    print(f"full size gmf sqrt FP64 order 2, n = {n}", flush=True)
    G = W.general(n, seed=0)
    dG = torch.from_numpy(G).cuda()
    out = buffers(n, torch.float64)
    lib = matfun_amd.library()
    assert lib.gmf_hip_eval(3, 0.0, n, dG.data_ptr(), *[t.data_ptr() for t in out],
                            torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    rows = sample_rows(n)
    got = picked(out, rows)
    _, Fo, Jo, Ho = Checker("oracle").gmf("sqrt", 0.0, G[rows], 2)
    ref = Checker("ref") if available("ref") else None
    ref = None if ref is None or ref.name().endswith("(scalar)") else ref
    rr = ref.gmf("sqrt", 0.0, G[rows], 2) if ref is not None else (0, None, None, None)
    for name, x, y, z in zip("FJH", got, (Fo, Jo, Ho), rr[1:]):
        spread_check(f"full size gmf sqrt o2 {name}", x, y, z)
