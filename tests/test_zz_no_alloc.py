"""gmf_hip_eval_sorted_noalloc (the length-sorted general-matrix evaluation with its workspace in a static device array,
for callers that cannot allocate) against gmf_hip_eval_sorted with a caller workspace: bitwise equal results on two
calls in a row, the second on fewer matrices (the array keeps the first call's contents), for each order; n beyond
the array's 2^22 matrices is refused before anything is launched."""
import numpy as np
import pytest

from matfun_amd import workloads as W

pytestmark = pytest.mark.gpu


def outputs(n, order):
    import torch
    return [torch.full((n, k), 7.0, dtype=torch.float64, device="cuda") if order >= o else None
            for o, k in ((0, 9), (1, 81), (2, 729))]


def test_noalloc_equals_workspace_entry_point_on_two_calls():
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    st = torch.cuda.current_stream().cuda_stream
    p = lambda t: None if t is None else t.data_ptr()
    G = W.general(5000, seed=801)
    for kind, alpha in (("log", 0.0), ("pow", 0.3), ("exp", 0.0)):
        for order in (2, 1, 0):
            for n in (5000, 333):                         # (the second call on fewer matrices)
                print(f"noalloc vs workspace: {kind} order {order} n {n}", flush=True)
                dX = torch.from_numpy(np.ascontiguousarray(G[:n])).cuda()
                a, b = outputs(n, order), outputs(n, order)
                k = matfun_amd.KINDS[kind]
                assert lib.gmf_hip_eval_sorted_noalloc(k, alpha, n, dX.data_ptr(), *map(p, a), st) == 0
                nb = lib.gmf_hip_sort_workspace(n)
                work = torch.empty(nb, dtype=torch.uint8, device="cuda")
                assert lib.gmf_hip_eval_sorted(k, alpha, n, dX.data_ptr(), *map(p, b), work.data_ptr(), nb, st) == 0
                torch.cuda.synchronize()
                for x, y in zip(a, b):
                    if y is not None:
                        assert torch.equal(torch.nan_to_num(x, 7.0), torch.nan_to_num(y, 7.0)), (kind, order, n)


def test_noalloc_in_parts_equals_one_pass():
    """a batch larger than the static array's staged capacity (about 39000 matrices at order 2, 56000 at order 1) runs in
    parts there: bitwise the workspace entry point's one pass"""
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    st = torch.cuda.current_stream().cuda_stream
    p = lambda t: None if t is None else t.data_ptr()
    n = 70000
    dX = torch.from_numpy(np.ascontiguousarray(W.general(n, seed=802))).cuda()
    nb = lib.gmf_hip_sort_workspace(n)
    work = torch.empty(nb, dtype=torch.uint8, device="cuda")
    for kind, alpha, order in (("log", 0.0, 2), ("pow", 0.3, 1)):
        print(f"noalloc in parts vs workspace: {kind} order {order} n {n}", flush=True)
        a, b = outputs(n, order), outputs(n, order)
        k = matfun_amd.KINDS[kind]
        assert lib.gmf_hip_eval_sorted_noalloc(k, alpha, n, dX.data_ptr(), *map(p, a), st) == 0
        assert lib.gmf_hip_eval_sorted(k, alpha, n, dX.data_ptr(), *map(p, b), work.data_ptr(), nb, st) == 0
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            if y is not None:
                assert torch.equal(torch.nan_to_num(x, 7.0), torch.nan_to_num(y, 7.0)), (kind, order)


def test_noalloc_rejects_before_any_launch():
    import matfun_amd
    lib = matfun_amd.library()
    dummy = 0x1000
    assert lib.gmf_hip_eval_sorted_noalloc(1, 0.0, (1 << 22) + 1, dummy, dummy, None, None, None) == -1
    assert lib.gmf_hip_eval_sorted_noalloc(9, 0.0, 10, dummy, dummy, None, None, None) == -1
    assert lib.gmf_hip_eval_sorted_noalloc(1, 0.0, 10, dummy, dummy, None, dummy, None) == -1
    assert lib.gmf_hip_eval_sorted_noalloc(2, -1.0, 10, dummy, dummy, None, None, None) == -1
    assert lib.gmf_hip_eval_sorted_noalloc(1, 0.0, 0, None, None, None, None, None) == 0
