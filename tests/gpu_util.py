"""Helpers of the GPU tests: calls of the HIP library's C ABI on torch device buffers (no test functions)."""
import numpy as np

import matfun_amd

SIZES = {"voigt": (6, 6), "matrix": (9, 9), "general": (9, 9)}


def call(family, kind, alpha, X, order, stream=None, mixed=False):
    """run smf_hip_eval / smf_hip_eval_matrix / gmf_hip_eval (or their _mixed forms) on host array X (n, 6|9);
    returns rc, F, J, H (NumPy)"""
    import torch
    lib = matfun_amd.library()
    if mixed:
        fn = {"voigt": lib.smf_hip_eval_mixed, "matrix": lib.smf_hip_eval_matrix_mixed,
              "general": lib.gmf_hip_eval_mixed}[family]
    else:
        fn = {"voigt": lib.smf_hip_eval, "matrix": lib.smf_hip_eval_matrix, "general": lib.gmf_hip_eval}[family]
    dt = torch.float32 if mixed else torch.float64
    nin, nf = SIZES[family]
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, nin)
    n = X.shape[0]
    dX = torch.from_numpy(X).to("cuda")
    dF = torch.full((n, nf), 7.0, dtype=dt, device="cuda")        # (7: never a result by accident)
    dJ = torch.full((n, nf * nf), 7.0, dtype=dt, device="cuda") if order >= 1 else None
    dH = torch.full((n, nf ** 3), 7.0, dtype=dt, device="cuda") if order >= 2 else None
    p = lambda t: None if t is None else t.data_ptr()
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    kind = matfun_amd.KINDS[kind] if isinstance(kind, str) else int(kind)
    rc = fn(kind, float(alpha), n, p(dX), p(dF), p(dJ), p(dH), st)
    torch.cuda.synchronize()
    g = lambda t: None if t is None else t.cpu().double().numpy()
    return rc, g(dF), g(dJ), g(dH)


def call_pow(family, alphas, X, order, mixed=False):
    """run smf_hip_eval_pow / smf_hip_eval_matrix_pow / gmf_hip_eval_pow (or their _mixed forms) with one exponent
    per matrix; returns rc, F, J, H (NumPy)"""
    import torch
    lib = matfun_amd.library()
    fn = {("voigt", False): lib.smf_hip_eval_pow, ("matrix", False): lib.smf_hip_eval_matrix_pow,
          ("general", False): lib.gmf_hip_eval_pow, ("voigt", True): lib.smf_hip_eval_pow_mixed,
          ("matrix", True): lib.smf_hip_eval_matrix_pow_mixed, ("general", True): lib.gmf_hip_eval_pow_mixed}[
        (family, mixed)]
    dt = torch.float32 if mixed else torch.float64
    nin, nf = SIZES[family]
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, nin)
    n = X.shape[0]
    assert len(alphas) == n
    dX = torch.from_numpy(X).to("cuda")
    dA = torch.from_numpy(np.ascontiguousarray(alphas, dtype=np.float64)).to("cuda")
    dF = torch.full((n, nf), 7.0, dtype=dt, device="cuda")
    dJ = torch.full((n, nf * nf), 7.0, dtype=dt, device="cuda") if order >= 1 else None
    dH = torch.full((n, nf ** 3), 7.0, dtype=dt, device="cuda") if order >= 2 else None
    p = lambda t: None if t is None else t.data_ptr()
    rc = fn(p(dA), n, p(dX), p(dF), p(dJ), p(dH), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    g = lambda t: None if t is None else t.cpu().double().numpy()
    return rc, g(dF), g(dJ), g(dH)


def call_sorted(kind, alpha, X, order, mixed=False, alphas=None):
    """run gmf_hip_eval_sorted / gmf_hip_eval_pow_sorted (alphas: one exponent per matrix), or their _mixed forms, on
    host array X (n, 9) with a workspace of gmf_hip_sort_workspace(n) bytes, at the strategy the library's table
    holds; returns rc, F, J, H (NumPy)"""
    import torch
    lib = matfun_amd.library()
    dt = torch.float32 if mixed else torch.float64
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 9)
    n = X.shape[0]
    dX = torch.from_numpy(X).to("cuda")
    dF = torch.full((n, 9), 7.0, dtype=dt, device="cuda")
    dJ = torch.full((n, 81), 7.0, dtype=dt, device="cuda") if order >= 1 else None
    dH = torch.full((n, 729), 7.0, dtype=dt, device="cuda") if order >= 2 else None
    p = lambda t: None if t is None else t.data_ptr()
    nb = lib.gmf_hip_sort_workspace(n)
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    if alphas is not None:
        dA = torch.from_numpy(np.ascontiguousarray(alphas, dtype=np.float64)).to("cuda")
        fn = lib.gmf_hip_eval_pow_sorted_mixed if mixed else lib.gmf_hip_eval_pow_sorted
        rc = fn(p(dA), n, p(dX), p(dF), p(dJ), p(dH), work.data_ptr(), nb, st)
    else:
        fn = lib.gmf_hip_eval_sorted_mixed if mixed else lib.gmf_hip_eval_sorted
        kind = matfun_amd.KINDS[kind] if isinstance(kind, str) else int(kind)
        rc = fn(kind, float(alpha), n, p(dX), p(dF), p(dJ), p(dH), work.data_ptr(), nb, st)
    torch.cuda.synchronize()
    g = lambda t: None if t is None else t.cpu().double().numpy()
    return rc, g(dF), g(dJ), g(dH)
