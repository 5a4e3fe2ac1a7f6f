"""matfun_amd -- exp, log, sqrt and A^alpha of 3x3 matrices with first and second derivatives on an AMD MI355X
(gfx950), through the HIP library libmatfun_hip.so (C ABI: include/matfun_hip.h).

    F, J, H = matfun_amd.smf("log", A, order=2)      symmetric matrices (..., 3, 3) or Voigt vectors (..., 6)
    F, J, H = matfun_amd.gmf("log", A, order=2)      general real matrices (..., 3, 3)

A is a float64 torch tensor on the GPU (a NumPy array or CPU tensor is copied to the current device and the results
come back as NumPy arrays).  The mirror of the reference's Python API matfun.smf / matfun.gmf
(python/src/matfun/_capi.py:62-113): same argument meaning, shapes and ValueErrors, with one difference inherited
from the GPU API (cuda/include/matfun/matfun_cuda.h:1-3): a non-symmetric matrix given to smf yields NaN results for
that matrix instead of an error.  The work runs on torch's current stream, asynchronously.

Extension: for kind "pow", alpha may be an array of the batch shape (one exponent per matrix, any finite value;
alpha = 0 gives I, alpha < 0 powers of the inverse, a non-finite exponent NaN results for its matrix), which runs the
*_pow entry points of the C ABI.
"""
import ctypes
import math
import os

__version__ = "0.1.3"
__all__ = ["smf", "gmf", "library", "library_path", "backends", "KINDS"]

KINDS = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3}
_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def library_path():
    """path of libmatfun_hip.so (MATFUN_HIP_LIB overrides the in-tree build)"""
    return os.environ.get("MATFUN_HIP_LIB") or os.path.join(_HERE, "libmatfun_hip.so")


def library():
    """the loaded HIP library (ctypes); raises if it was not built -- there is no CPU fallback"""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"matfun_amd: HIP library {path} not built (python -m matfun_amd.build)")
        lib = ctypes.CDLL(path)
        args = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 5
        for fn in (lib.smf_hip_eval, lib.smf_hip_eval_matrix, lib.gmf_hip_eval, lib.smf_hip_eval_mixed,
                   lib.smf_hip_eval_matrix_mixed, lib.gmf_hip_eval_mixed):
            fn.argtypes, fn.restype = args, ctypes.c_int
        pargs = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
        for fn in (lib.smf_hip_eval_pow, lib.smf_hip_eval_matrix_pow, lib.gmf_hip_eval_pow,
                   lib.smf_hip_eval_pow_mixed, lib.smf_hip_eval_matrix_pow_mixed, lib.gmf_hip_eval_pow_mixed):
            fn.argtypes, fn.restype = pargs, ctypes.c_int
        lib.smf_hip_version.restype = ctypes.c_char_p
        lib.smf_hip_target.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def backends():
    """what is loaded: library path, version, device-code target"""
    lib = library()
    return {"version": __version__, "hip": library_path(), "hip_version": lib.smf_hip_version().decode(),
            "target": lib.smf_hip_target().decode()}


def _stream_handle(torch):
    return torch.cuda.current_stream().cuda_stream


def _per_matrix(alpha):
    """True for an exponent array (one per matrix), False for a scalar"""
    return hasattr(alpha, "shape") and len(alpha.shape) > 0 or isinstance(alpha, (list, tuple))


def _prepare(kind, alpha, order):
    if kind not in KINDS:
        raise ValueError(f"unknown function {kind!r}, expected one of {sorted(KINDS)}")
    if kind == "pow" and _per_matrix(alpha):
        pass                                           # (shape checked against the batch in _run)
    elif kind == "pow" and not (alpha is not None and alpha > 0):
        raise ValueError("pow needs an exponent alpha > 0")
    if order not in (0, 1, 2):
        raise ValueError("order must be 0, 1 or 2")


def _run(fn, kind, alpha, A, batch_shape, shapes, order, out_dtype, per_matrix=False):
    import torch
    to_numpy = not isinstance(A, torch.Tensor) or not A.is_cuda
    X = torch.as_tensor(A, dtype=torch.float64)
    if not X.is_cuda:
        X = X.to("cuda")
    X = X.contiguous()
    n = math.prod(batch_shape)
    if per_matrix:
        al = torch.as_tensor(alpha, dtype=torch.float64).to(X.device).contiguous()
        if tuple(al.shape) != tuple(batch_shape):
            raise ValueError(f"alpha has shape {tuple(al.shape)}, expected the batch shape {tuple(batch_shape)}")
    out = [torch.empty(batch_shape + s, dtype=out_dtype, device=X.device) for s in shapes[:order + 1]]
    ptrs = [o.data_ptr() for o in out] + [None] * (2 - order)
    with torch.cuda.device(X.device):
        if per_matrix:
            r = fn(al.data_ptr() if n else None, n, X.data_ptr() if n else None, *ptrs, _stream_handle(torch))
        else:
            r = fn(KINDS[kind], float(alpha or 0.0), n, X.data_ptr() if n else None, *ptrs, _stream_handle(torch))
    if r != 0:
        raise ValueError(f"{fn.__name__} returned {r}")
    if to_numpy:
        out = [o.cpu().numpy() for o in out]
    return tuple(out) if order else out[0]


def _precision(precision):
    import torch
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return precision == "mixed", (torch.float32 if precision == "mixed" else torch.float64)


def smf(kind, A, alpha=None, order=0, precision="fp64"):
    """f(A) and its derivatives for symmetric 3x3 matrices A (..., 3, 3) or Voigt vectors (..., 6) =
    (A11, A22, A33, A12, A13, A23); results in the same form: F (..., 3, 3), J (..., 3, 3, 3, 3) = dF_ab/dA_cd,
    H (..., 3^6) resp. F (..., 6), J (..., 6, 6), H (..., 6, 6, 6).  order 0: F, 1: (F, J), 2: (F, J, H).
    precision "mixed": FP32 results (FP64 eigen-decomposition and divided differences, FP32 assembly)."""
    _prepare(kind, alpha, order)
    mixed, dt = _precision(precision)
    lib = library()
    shape = tuple(A.shape)
    pa = kind == "pow" and _per_matrix(alpha)
    if shape[-2:] == (3, 3):
        fn = ((lib.smf_hip_eval_matrix_pow_mixed if mixed else lib.smf_hip_eval_matrix_pow) if pa else
              (lib.smf_hip_eval_matrix_mixed if mixed else lib.smf_hip_eval_matrix))
        return _run(fn, kind, alpha, A, shape[:-2], ((3, 3), (3, 3, 3, 3), (3, 3, 3, 3, 3, 3)), order, dt, pa)
    if shape[-1:] == (6,):
        fn = ((lib.smf_hip_eval_pow_mixed if mixed else lib.smf_hip_eval_pow) if pa else
              (lib.smf_hip_eval_mixed if mixed else lib.smf_hip_eval))
        return _run(fn, kind, alpha, A, shape[:-1], ((6,), (6, 6), (6, 6, 6)), order, dt, pa)
    raise ValueError(f"expected matrices (..., 3, 3) or Voigt vectors (..., 6), got shape {shape}")


def gmf(kind, A, alpha=None, order=0, precision="fp64"):
    """f(A) and its derivatives for general real 3x3 matrices A (..., 3, 3), principal branch; NaN results for a
    matrix with a real eigenvalue <= 0 (log, sqrt, pow).  order 0: F, 1: (F, J), 2: (F, J, H).
    precision "mixed": FP32 results (FP64 per-matrix spectral part, FP32 assembly of J and H)."""
    _prepare(kind, alpha, order)
    mixed, dt = _precision(precision)
    lib = library()
    shape = tuple(A.shape)
    if shape[-2:] != (3, 3):
        raise ValueError(f"expected matrices (..., 3, 3), got shape {shape}")
    pa = kind == "pow" and _per_matrix(alpha)
    fn = ((lib.gmf_hip_eval_pow_mixed if mixed else lib.gmf_hip_eval_pow) if pa else
          (lib.gmf_hip_eval_mixed if mixed else lib.gmf_hip_eval))
    return _run(fn, kind, alpha, A, shape[:-2], ((3, 3), (3, 3, 3, 3), (3, 3, 3, 3, 3, 3)), order, dt, pa)
