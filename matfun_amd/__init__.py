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
import weakref

__version__ = "0.1.3"
__all__ = ["smf", "smf_packed", "expand_packed", "gmf", "hencky", "smf_hencky", "gmf_hencky", "Batch", "Group",
           "library", "library_path", "backends", "KINDS"]

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
        for fn in (lib.smf_hip_eval_packed, lib.smf_hip_eval_packed_mixed):
            fn.argtypes, fn.restype = args, ctypes.c_int
        pargs = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
        for fn in (lib.smf_hip_eval_pow, lib.smf_hip_eval_matrix_pow, lib.gmf_hip_eval_pow,
                   lib.smf_hip_eval_pow_mixed, lib.smf_hip_eval_matrix_pow_mixed, lib.gmf_hip_eval_pow_mixed):
            fn.argtypes, fn.restype = pargs, ctypes.c_int
        sargs = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t,
                                                                                           ctypes.c_void_p]
        for fn in (lib.gmf_hip_eval_sorted, lib.gmf_hip_eval_sorted_mixed):
            fn.argtypes, fn.restype = sargs, ctypes.c_int
        spargs = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t, ctypes.c_void_p]
        for fn in (lib.gmf_hip_eval_pow_sorted, lib.gmf_hip_eval_pow_sorted_mixed):
            fn.argtypes, fn.restype = spargs, ctypes.c_int
        lib.gmf_hip_sort_workspace.argtypes, lib.gmf_hip_sort_workspace.restype = [ctypes.c_int64], ctypes.c_size_t
        lib.gmf_hip_eval_sorted_noalloc.argtypes, lib.gmf_hip_eval_sorted_noalloc.restype = args, ctypes.c_int
        lib.gmf_hip_set_sort_strategy.argtypes = [ctypes.c_int] * 4
        lib.gmf_hip_get_sort_strategy.argtypes = [ctypes.c_int] * 3
        lib.gmf_hip_set_tiles.argtypes = [ctypes.c_int] * 4
        for fn in (lib.mf_hip_store_probe, lib.mf_hip_store_probe_f32):
            fn.argtypes, fn.restype = [ctypes.c_int64, ctypes.c_int] + [ctypes.c_void_p] * 4, ctypes.c_int
        for fn in (lib.mf_hip_store_probe_in, lib.mf_hip_store_probe_in_f32):
            fn.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4
            fn.restype = ctypes.c_int
        lib.gmf_hip_get_tiles.argtypes = [ctypes.c_int] * 3
        for fn in (lib.mf_hip_hencky, lib.mf_hip_hencky_mixed):
            fn.argtypes, fn.restype = [ctypes.c_int64] + [ctypes.c_void_p] * 7, ctypes.c_int
        for fn in (lib.smf_hip_hencky_matrix, lib.smf_hip_hencky, lib.gmf_hip_hencky):
            fn.argtypes, fn.restype = [ctypes.c_int64] + [ctypes.c_void_p] * 6, ctypes.c_int
        lib.smf_hip_set_tile.argtypes = [ctypes.c_int] * 4
        lib.smf_hip_get_tile.argtypes = [ctypes.c_int] * 3
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


def _run(fn, kind, alpha, A, batch_shape, shapes, order, out_dtype, per_matrix=False, sort=False):
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
    extra = []
    if sort:                                           # (the sorted entry points' workspace, torch's allocator)
        nbytes = library().gmf_hip_sort_workspace(n)
        work = torch.empty(nbytes, dtype=torch.uint8, device=X.device)
        extra = [work.data_ptr(), nbytes]
    with torch.cuda.device(X.device):
        if per_matrix:
            r = fn(al.data_ptr() if n else None, n, X.data_ptr() if n else None, *ptrs, *extra, _stream_handle(torch))
        else:
            r = fn(KINDS[kind], float(alpha or 0.0), n, X.data_ptr() if n else None, *ptrs, *extra,
                   _stream_handle(torch))
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


def smf_packed(kind, v, alpha=None, order=0, precision="fp64"):
    """f(A) and its derivatives for symmetric matrices given as Voigt vectors v (..., 6), as the distinct values only:
    F (..., 6) Voigt, J (..., 21) = dF_ab/dA_cd for Voigt indices i = (a,b) <= k = (c,d), H (..., 126) = the Voigt
    second derivatives d2F_i/dv_p dv_q for p <= q, row i at 21 i (smf_hip_eval_packed).  expand_packed gives the
    Voigt or matrix layouts of smf."""
    _prepare(kind, alpha, order)
    mixed, dt = _precision(precision)
    lib = library()
    shape = tuple(v.shape)
    if shape[-1:] != (6,):
        raise ValueError(f"expected Voigt vectors (..., 6), got shape {shape}")
    return _run(lib.smf_hip_eval_packed_mixed if mixed else lib.smf_hip_eval_packed, kind, alpha, v, shape[:-1],
                ((6,), (21,), (126,)), order, dt)


def _tri6():
    import numpy as np
    T = np.zeros((6, 6), dtype=np.int64)
    for p in range(6):
        for q in range(6):
            a, b = min(p, q), max(p, q)
            T[p, q] = a * (11 - a) // 2 + b
    return T


def expand_packed(F, J=None, H=None, layout="voigt"):
    """the Voigt (F 6, J 6x6, H 6x6x6) or matrix (F 3x3, J 3^4, H 3^6) layout of packed results (NumPy arrays,
    any leading shape); the weights are exact (1, 2 / 1, 1/2), so the results are those of smf / smf_hip_eval*"""
    import numpy as np
    F = np.asarray(F)
    T = _tri6()
    lead = F.shape[:-1]
    if layout == "voigt":
        w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
        out = [F]
        if J is not None:
            out.append(np.asarray(J)[..., T] * w)
        if H is not None:
            out.append(np.asarray(H)[..., 21 * np.arange(6)[:, None, None] + T[None, :, :]])
        return tuple(out) if len(out) > 1 else out[0]
    if layout != "matrix":
        raise ValueError("layout must be 'voigt' or 'matrix'")
    vm = np.array([[0, 3, 4], [3, 1, 5], [4, 5, 2]])
    u = np.where(np.eye(3, dtype=bool), 1.0, 0.5)
    out = [F[..., vm]]
    if J is not None:
        out.append(np.asarray(J)[..., T[vm[:, :, None, None], vm[None, None, :, :]]])
    if H is not None:
        idx = 21 * vm[:, :, None, None, None, None] + T[vm[None, None, :, :, None, None], vm[None, None, None, None, :, :]]
        wt = u[:, :, None, None] * u[None, None, :, :]
        out.append(np.asarray(H)[..., idx] * wt)
    del lead
    return tuple(out) if len(out) > 1 else out[0]


def gmf(kind, A, alpha=None, order=0, precision="fp64", sort=False):
    """f(A) and its derivatives for general real 3x3 matrices A (..., 3, 3), principal branch; NaN results for a
    matrix with a real eigenvalue <= 0 (log, sqrt, pow).  order 0: F, 1: (F, J), 2: (F, J, H).
    precision "mixed": FP32 results (FP64 per-matrix spectral part, FP32 assembly of J and H).
    sort=True: through the workspace entry points (gmf_hip_eval_sorted*, a workspace of about 540 bytes per matrix
    from torch's allocator) that may evaluate the matrices in the order of their series lengths (the process-wide
    strategy per function, order and precision, gmf_hip_set_sort_strategy): gmf_hip_eval's results bitwise or, for a
    staged strategy, up to rounding (<= 1e-13 per matrix)."""
    _prepare(kind, alpha, order)
    mixed, dt = _precision(precision)
    lib = library()
    shape = tuple(A.shape)
    if shape[-2:] != (3, 3):
        raise ValueError(f"expected matrices (..., 3, 3), got shape {shape}")
    pa = kind == "pow" and _per_matrix(alpha)
    if sort:
        fn = ((lib.gmf_hip_eval_pow_sorted_mixed if mixed else lib.gmf_hip_eval_pow_sorted) if pa else
              (lib.gmf_hip_eval_sorted_mixed if mixed else lib.gmf_hip_eval_sorted))
    else:
        fn = ((lib.gmf_hip_eval_pow_mixed if mixed else lib.gmf_hip_eval_pow) if pa else
              (lib.gmf_hip_eval_mixed if mixed else lib.gmf_hip_eval))
    return _run(fn, kind, alpha, A, shape[:-2], ((3, 3), (3, 3, 3, 3), (3, 3, 3, 3, 3, 3)), order, dt, pa, sort)


def hencky(L, J, H=None):
    """value, gradient and tangent of the Hencky-type potential psi(X) = 1/2 |log X|^2 (reference theory.md:190):
    from L = log X (..., 3, 3), J = dL/dX (..., 3, 3, 3, 3) and optionally H = d2L/dX2 (..., 3^6), as
    gmf("log", X, order=2) or smf("log", X 3x3, order=2) return them, per matrix
        psi = 1/2 L:L (...),   g = J^T L (..., 3, 3) = dpsi/dX,   K = J^T J + L:H (..., 3, 3, 3, 3) = dg/dX
    (K is None without H).  float64 inputs run mf_hip_hencky, float32 ones (precision "mixed") mf_hip_hencky_mixed;
    sums in FP64 for both.  torch CUDA tensors stay on the device (torch's current stream); NumPy arrays are copied
    to the GPU and the results come back as NumPy arrays."""
    import torch
    to_numpy = not isinstance(L, torch.Tensor) or not L.is_cuda
    if isinstance(L, torch.Tensor):
        dt = L.dtype
    else:
        import numpy as np
        dt = torch.float32 if np.asarray(L).dtype == np.float32 else torch.float64
    if dt not in (torch.float64, torch.float32):
        raise ValueError(f"expected float64 or float32 inputs, got {dt}")
    ins = [torch.as_tensor(x, dtype=dt) for x in (L, J, H) if x is not None]
    ins = [(x if x.is_cuda else x.to("cuda")).contiguous() for x in ins]
    lead = tuple(ins[0].shape[:-2])
    for x, tail, name in zip(ins, ((3, 3), (3,) * 4, (3,) * 6), ("L", "J", "H")):
        if tuple(x.shape) != lead + tail:
            raise ValueError(f"{name} has shape {tuple(x.shape)}, expected {lead + tail}")
    n = math.prod(lead)
    dev = ins[0].device
    psi = torch.empty(lead, dtype=dt, device=dev)
    g = torch.empty(lead + (3, 3), dtype=dt, device=dev)
    K = torch.empty(lead + (3,) * 4, dtype=dt, device=dev) if H is not None else None
    lib = library()
    fn = lib.mf_hip_hencky_mixed if dt == torch.float32 else lib.mf_hip_hencky
    ptr = (lambda t: t.data_ptr() if t is not None and n else None)
    with torch.cuda.device(dev):
        r = fn(n, ptr(ins[0]), ptr(ins[1]), ptr(ins[2]) if H is not None else None, ptr(psi), ptr(g), ptr(K),
               _stream_handle(torch))
    if r != 0:
        raise ValueError(f"{fn.__name__} returned {r}")
    out = (psi, g, K)
    if to_numpy:
        out = tuple(o.cpu().numpy() if o is not None else None for o in out)
    return out


def smf_hencky(X):
    """log X, psi, g, K of the Hencky-type potential for symmetric X (..., 3, 3) in one fused kernel
    (smf_hip_hencky_matrix; FP64): the results of smf("log", X, order=2) followed by hencky, without J and H.
    Returns (L, psi, g, K); NaN results for a non-symmetric matrix.  Voigt vectors X (..., 6) = (X11, X22, X33, X12,
    X13, X23) run smf_hip_hencky: L (..., 6), g = dpsi/dv (..., 6), K = d2psi/dv2 (..., 6, 6)."""
    lib = library()
    if tuple(X.shape[-1:]) == (6,):
        return _fused_hencky(lib.smf_hip_hencky, X, voigt=True)
    return _fused_hencky(lib.smf_hip_hencky_matrix, X)


def gmf_hencky(X):
    """the same for general real X (..., 3, 3) (gmf_hip_hencky; FP64): gmf("log", X, order=2) followed by hencky in
    one kernel; NaN results for a matrix with a real eigenvalue <= 0"""
    return _fused_hencky(library().gmf_hip_hencky, X)


def _fused_hencky(fn, X, voigt=False):
    import torch
    to_numpy = not isinstance(X, torch.Tensor) or not X.is_cuda
    A = torch.as_tensor(X, dtype=torch.float64)
    A = (A if A.is_cuda else A.to("cuda")).contiguous()
    if not voigt and tuple(A.shape[-2:]) != (3, 3):
        raise ValueError(f"expected matrices (..., 3, 3) or Voigt vectors (..., 6), got shape {tuple(A.shape)}")
    lead = tuple(A.shape[:-1] if voigt else A.shape[:-2])
    n = math.prod(lead)
    shapes = ((6,), (), (6,), (6, 6)) if voigt else ((3, 3), (), (3, 3), (3, 3, 3, 3))
    outs = [torch.empty(lead + s, dtype=torch.float64, device=A.device) for s in shapes]
    with torch.cuda.device(A.device):
        r = fn(n, A.data_ptr() if n else None, *[o.data_ptr() if n else None for o in outs], _stream_handle(torch))
    if r != 0:
        raise ValueError(f"{fn.__name__} returned {r}")
    if to_numpy:
        outs = [o.cpu().numpy() for o in outs]
    return tuple(outs)


# ---- heterogeneous batches (include/matfun_hip.h, mf_hip_*)
class Group(ctypes.Structure):
    """one group of matrices with a common function: family 0 smf / 1 gmf, kind, order, count, alpha (pow)"""
    _fields_ = [("family", ctypes.c_int32), ("kind", ctypes.c_int32), ("order", ctypes.c_int32),
                ("count", ctypes.c_int32), ("alpha", ctypes.c_double)]


def _batch_lib():
    lib = library()
    if not getattr(lib, "_batch_ready", False):
        P = ctypes.c_void_p
        lib.mf_hip_plan_create.argtypes = [ctypes.POINTER(Group), ctypes.c_int64, ctypes.POINTER(P)]
        lib.mf_hip_plan_destroy.argtypes = [P]
        lib.mf_hip_plan_size.argtypes, lib.mf_hip_plan_size.restype = [P, ctypes.c_int], ctypes.c_int64
        for fn in (lib.mf_hip_batch_eval, lib.mf_hip_batch_eval_mixed):
            fn.argtypes = [P] * 7
        lib.mf_hip_plan_set_strategy.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.mf_hip_plan_tune.argtypes = [P] * 6 + [ctypes.c_int, P, ctypes.POINTER(ctypes.c_double)]
        lib.mf_hip_perturb.argtypes = [P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, P]
        lib.mf_hip_sequence_create.argtypes = [P] * 7 + [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                                         ctypes.POINTER(P)]
        lib.mf_hip_sequence_launch.argtypes = [P, P]
        lib.mf_hip_sequence_destroy.argtypes = [P]
        lib._batch_ready = True
    return lib


FAMILIES = {"smf": 0, "gmf": 1}


def _check_layout(t, name, dtype, count, exact):
    """t must be a contiguous tensor of `dtype` with `count` elements (exact) or at least `count`; the C ABI takes raw
    device pointers, so a tensor that is too short, strided or of another type would be read or written out of
    bounds by the kernels.  Returns the tensor's device (checked by the caller once every tensor passed this)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch CUDA tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: expected dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    if (t.numel() != count) if exact else (t.numel() < count):
        raise ValueError(f"{name}: {t.numel()} elements, expected {'' if exact else 'at least '}{count}")
    return name, t.device


def _check_batch_args(n, nj, nh, A=None, alpha=None, out=None, precision="fp64", names=("A",)):
    """checks every tensor a batch call passes to the C ABI before anything is launched: the inputs named in
    `names` (float64, n x 9 elements each), alpha (None or n float64), out = (F, J, H) in the precision's type with
    room for n x 9, nj x 81 and nh x 729 elements (J / H may be None when the batch has no row for them); then that
    all of them are on one GPU.  Raises ValueError."""
    import torch
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    ins = A if isinstance(A, (tuple, list)) else (A,)
    placed = []
    for t, name in zip(ins, names):
        if n == 0 and (t is None or (isinstance(t, torch.Tensor) and t.numel() == 0)):
            continue                                   # (an empty batch reads nothing)
        placed.append(_check_layout(t, name, torch.float64, 9 * n, exact=True))
    if alpha is not None:
        placed.append(_check_layout(alpha, "alpha", torch.float64, n, exact=True))
    if out is not None:
        dt = torch.float32 if precision == "mixed" else torch.float64
        for t, name, count in zip(out, ("F", "J", "H"), (9 * n, 81 * nj, 729 * nh)):
            if t is None or t.numel() == 0:
                if count > 0:
                    raise ValueError(f"{name}: the batch has {count} elements of it, got no tensor")
                continue
            placed.append(_check_layout(t, name, dt, count, exact=False))
    for name, dev in placed:
        if dev.type != "cuda":
            raise ValueError(f"{name}: expected a tensor on the GPU, got one on {dev}")
    if len({dev for _, dev in placed}) > 1:
        raise ValueError(f"tensors on different devices: {sorted(str(d) for _, d in placed)}")


class Batch:
    """A heterogeneous batch: groups (family, kind, order, count[, alpha]) of 3x3 matrices, evaluated by one kernel.

        b = matfun_amd.Batch([("smf", "log", 2, 27), ("gmf", "exp", 1, 8), ("smf", "pow", 1, 4, 0.5), ...])
        F, J, H = b.eval(A)                 A: (N, 3, 3) float64 CUDA tensor, groups one after another
        F, J, H = b.eval(A, alpha=al)       al: (N,) per-matrix exponents (read for pow groups)

    J holds a row (3, 3, 3, 3) per matrix of order >= 1 and H a row (3,)*6 per matrix of order 2, in group order.
    Device tensors in, device tensors out (torch's current stream)."""

    GROUP_DTYPE = [("family", "<i4"), ("kind", "<i4"), ("order", "<i4"), ("count", "<i4"), ("alpha", "<f8")]

    def __init__(self, groups):
        lib = _batch_lib()
        if hasattr(groups, "dtype"):                   # a structured array of GROUP_DTYPE (large batches)
            arr = groups
            ng = len(groups)
            ptr = ctypes.cast(arr.ctypes.data, ctypes.POINTER(Group)) if ng else (Group * 1)()
        else:
            ng = len(groups)
            arr = ptr = (Group * max(1, ng))()
            for i, g in enumerate(groups):
                fam, kind, order, count = g[:4]
                alpha = g[4] if len(g) > 4 else 0.0
                arr[i] = Group(FAMILIES.get(fam, fam) if isinstance(fam, str) else int(fam),
                               KINDS[kind] if isinstance(kind, str) else int(kind), int(order), int(count),
                               float(alpha))
        h = ctypes.c_void_p()
        r = lib.mf_hip_plan_create(ptr, ng, ctypes.byref(h))
        if r != 0:
            raise ValueError(f"mf_hip_plan_create returned {r}")
        self._h, self._lib = h, lib
        self.n, self.nj, self.nh, self.ntiles, self.nclasses = (int(lib.mf_hip_plan_size(h, w)) for w in range(5))
        self.groups = groups
        self._sequences = weakref.WeakSet()

    def close(self):
        """frees the plan; refused (RuntimeError) while a Sequence of this batch is open, since its graph uses the
        plan's device tables"""
        if self._h:
            if any(s._h for s in list(self._sequences)):
                raise RuntimeError("Batch.close: close the batch's open sequences first")
            if self._lib.mf_hip_plan_destroy(self._h) != 0:
                raise RuntimeError("mf_hip_plan_destroy refused: a sequence of the plan is still alive")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def outputs(self, precision="fp64"):
        import torch
        dt = torch.float32 if precision == "mixed" else torch.float64
        return (torch.empty((self.n, 3, 3), dtype=dt, device="cuda"),
                torch.empty((self.nj,) + (3,) * 4, dtype=dt, device="cuda"),
                torch.empty((self.nh,) + (3,) * 6, dtype=dt, device="cuda"))

    @staticmethod
    def _p(t):
        return None if t is None or t.numel() == 0 else t.data_ptr()

    def _live(self):
        if not self._h:
            raise RuntimeError("the batch is closed")

    def eval(self, A, alpha=None, out=None, precision="fp64"):
        import torch
        self._live()
        _check_batch_args(self.n, self.nj, self.nh, A, alpha, out, precision)
        F, J, H = out if out is not None else self.outputs(precision)
        fn = self._lib.mf_hip_batch_eval_mixed if precision == "mixed" else self._lib.mf_hip_batch_eval
        r = fn(self._h, self._p(A), self._p(alpha), self._p(F), self._p(J), self._p(H),
               torch.cuda.current_stream().cuda_stream)
        if r != 0:
            raise ValueError(f"mf_hip_batch_eval returned {r}")
        return F, J, H

    def set_strategy(self, mode, per_cu, order=1):
        """mf_hip_plan_set_strategy: mode 0 fused, 1 split, 2 concurrent split, 3 staged split (allocates the plan's
        workspace for the staged classes on first use)"""
        r = self._lib.mf_hip_plan_set_strategy(self._h, int(mode), int(per_cu), int(order))
        if r == -2:
            raise MemoryError("mf_hip_plan_set_strategy: the staged classes' workspace could not be allocated")
        if r != 0:
            raise ValueError(f"invalid strategy (mode {mode}, per_cu {per_cu}, order {order})")

    def tune(self, A, alpha=None, out=None, precision="fp64"):
        import torch
        self._live()
        _check_batch_args(self.n, self.nj, self.nh, A, alpha, out, precision)
        F, J, H = out if out is not None else self.outputs(precision)
        res = (ctypes.c_double * 4)()
        r = self._lib.mf_hip_plan_tune(self._h, self._p(A), self._p(alpha), self._p(F), self._p(J), self._p(H),
                                       1 if precision == "mixed" else 0, torch.cuda.current_stream().cuda_stream, res)
        if r != 0:
            raise ValueError(f"mf_hip_plan_tune returned {r}")
        return {"mode": int(res[0]), "per_cu": int(res[1]), "order": int(res[2]), "us": res[3]}

    def perturb(self, A, sigma, seed, iteration):
        import torch
        self._live()
        _check_batch_args(self.n, self.nj, self.nh, A)
        r = self._lib.mf_hip_perturb(self._h, self._p(A), float(sigma), int(seed), int(iteration),
                                     torch.cuda.current_stream().cuda_stream)
        if r != 0:
            raise ValueError(f"mf_hip_perturb returned {r}")

    def sequence(self, A0, A, out, iterations, sigma, seed, alpha=None, precision="fp64"):
        """a graph-captured Newton sequence (mf_hip_sequence_*): .launch() replays it on the current stream"""
        self._live()
        _check_batch_args(self.n, self.nj, self.nh, (A0, A), alpha, out, precision, names=("A0", "A"))
        s = Sequence(self, A0, A, out, iterations, sigma, seed, alpha, precision)
        self._sequences.add(s)
        return s


def group_array(family, kind, order, count, alpha=None):
    """groups as a structured array (Batch.GROUP_DTYPE) from equal-length integer arrays (and exponents)"""
    import numpy as np
    g = np.zeros(len(count), dtype=Batch.GROUP_DTYPE)
    g["family"], g["kind"], g["order"], g["count"] = family, kind, order, count
    if alpha is not None:
        g["alpha"] = alpha
    return g


class Sequence:
    def __init__(self, batch, A0, A, out, iterations, sigma, seed, alpha, precision):
        F, J, H = out
        h = ctypes.c_void_p()
        p = Batch._p
        r = batch._lib.mf_hip_sequence_create(batch._h, p(A0), p(A), p(alpha), p(F), p(J), p(H),
                                              1 if precision == "mixed" else 0, int(iterations), float(sigma),
                                              int(seed), ctypes.byref(h))
        if r != 0:
            raise ValueError(f"mf_hip_sequence_create returned {r}")
        self._h, self._lib, self._keep = h, batch._lib, (batch, A0, A, alpha, out)

    def launch(self):
        import torch
        r = self._lib.mf_hip_sequence_launch(self._h, torch.cuda.current_stream().cuda_stream)
        if r != 0:
            raise ValueError(f"mf_hip_sequence_launch returned {r}")

    def close(self):
        if self._h:
            self._lib.mf_hip_sequence_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
