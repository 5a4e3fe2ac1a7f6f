"""Packed symmetric outputs (smf_hip_eval_packed / _mixed: F 6, J 21, H 126 values per matrix; an extension whose H
packing is the reference kernel's HPACK layout) on the GPU: expanded with the exact weights (matfun_amd.expand_packed)
they equal, bitwise, the Voigt results of smf_hip_eval and the matrix-layout results of smf_hip_eval_matrix on the
same matrices (both pinned to the oracle by tests/test_gpu_smf.py), FP64 and mixed, all functions and orders,
including a partial last wavefront and unaligned output pointers."""
import numpy as np
import pytest

from matfun_amd import workloads as W

pytestmark = pytest.mark.gpu
KINDS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)]


def run_packed(kind, alpha, v, order, mixed, offset=0):
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    fn = lib.smf_hip_eval_packed_mixed if mixed else lib.smf_hip_eval_packed
    dt = torch.float32 if mixed else torch.float64
    n = len(v)
    dv = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    bufs = [torch.full((n * s + offset,), 7.0, dtype=dt, device="cuda") for s in (6, 21, 126)]
    outs = [b[offset:] for b in bufs]
    p = [t.data_ptr() for t in outs]
    rc = fn(matfun_amd.KINDS[kind], alpha, n, dv.data_ptr(), p[0], p[1] if order >= 1 else None,
            p[2] if order >= 2 else None, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == 0
    return [t.reshape(n, -1).cpu().double().numpy() for t in outs]


@pytest.mark.parametrize("kind,alpha", KINDS)
@pytest.mark.parametrize("mixed", [False, True])
def test_packed_expands_to_the_api_layouts(kind, alpha, mixed):
    import matfun_amd
    from tests.gpu_util import call
    A = W.spd(997, seed=802, log10_range=(-2.0, 1.7))       # (997: partial last wavefront; exp inside FP32)
    v = W.voigt(A)
    for order in (0, 1, 2):
        print(f"packed {kind} order {order} mixed={mixed}", flush=True)
        Fp, Jp, Hp = run_packed(kind, alpha, v, order, mixed)
        _, Fv, Jv, Hv = call("voigt", kind, alpha, v, order, mixed=mixed)
        _, Fm, Jm, Hm = call("matrix", kind, alpha, A, order, mixed=mixed)
        ev = matfun_amd.expand_packed(Fp, Jp if order >= 1 else None, Hp if order >= 2 else None, "voigt")
        em = matfun_amd.expand_packed(Fp, Jp if order >= 1 else None, Hp if order >= 2 else None, "matrix")
        ev, em = (ev,) if order == 0 else ev, (em,) if order == 0 else em
        for x, y in zip(ev, (Fv, Jv, Hv)[:order + 1]):
            assert np.array_equal(x.reshape(len(A), -1), y, equal_nan=True)
        for x, y in zip(em, (Fm, Jm, Hm)[:order + 1]):
            assert np.array_equal(x.reshape(len(A), -1), y, equal_nan=True)


@pytest.mark.parametrize("mixed", [False, True])
def test_packed_unaligned_outputs(mixed):
    """outputs shifted by one element: the single-element store path gives the same values"""
    A = W.spd(333, seed=803)
    v = W.voigt(A)
    a = run_packed("log", 0.0, v, 2, mixed)
    b = run_packed("log", 0.0, v, 2, mixed, offset=1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_packed_python_api():
    import matfun_amd
    A = W.spd(100, seed=804)
    v = W.voigt(A).reshape(10, 10, 6)
    F, J, H = matfun_amd.smf_packed("sqrt", v, order=2)
    assert F.shape == (10, 10, 6) and J.shape == (10, 10, 21) and H.shape == (10, 10, 126)
    Fm, Jm, Hm = matfun_amd.smf("sqrt", A.reshape(10, 10, 3, 3), order=2)
    em = matfun_amd.expand_packed(F, J, H, "matrix")
    for x, y in zip(em, (Fm, Jm, Hm)):
        assert np.array_equal(x, y)
    with pytest.raises(ValueError):
        matfun_amd.smf_packed("log", A.reshape(-1, 3, 3), order=1)
