"""NumPy front end of the CPU checkers (TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
cpu_baseline leg, never by the product).

    Checker("oracle")     this repository's scalar C restatement (oracle/liboracle.so, built by `make -C oracle`)
    Checker("ref")        the reference C library compiled from BodeTobias/matfun's own sources
                          (oracle/_ref/libmatfun_ref.so; None when it was not built)

Both expose smf(kind, alpha, v, order), smf_matrix(kind, alpha, A, order), gmf(kind, alpha, A, order) returning
(rc, F, J, H) with J / H None below the order, in the layouts of c/include/matfun/matfun.h.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"oracle": os.path.join(HERE, "liboracle.so"), "ref": os.path.join(HERE, "_ref", "libmatfun_ref.so")}
KINDS = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3}


def available(which):
    return os.path.exists(PATHS[which])


class Checker:
    def __init__(self, which="oracle"):
        path = PATHS[which]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.which = which
        self.lib = ctypes.CDLL(path)
        pre = "oracle_" if which == "oracle" else ""
        args = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 4
        self._smf = getattr(self.lib, pre + "smf_eval")
        self._smfm = getattr(self.lib, pre + "smf_eval_matrix")
        self._gmf = getattr(self.lib, pre + "gmf_eval")
        for f in (self._smf, self._smfm, self._gmf):
            f.argtypes, f.restype = args, ctypes.c_int
        if which == "oracle":
            aargs = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 4
            self._smfa, self._gmfa = self.lib.oracle_smf_eval_matrix_alpha, self.lib.oracle_gmf_eval_alpha
            self._smfva = self.lib.oracle_smf_eval_alpha
            for f in (self._smfa, self._gmfa, self._smfva):
                f.argtypes, f.restype = aargs, ctypes.c_int
            self.lib.oracle_set_threads.argtypes = [ctypes.c_int]
            self.lib.oracle_gmf_terms.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
            self.lib.oracle_gmf_terms.restype = ctypes.c_int
        else:
            self.lib.smf_simd_name.restype = ctypes.c_char_p

    def name(self):
        if self.which == "ref":
            return "reference C (" + self.lib.smf_simd_name().decode() + ")"
        return "oracle (scalar C restatement)"

    def set_threads(self, n):
        if self.which == "oracle":
            self.lib.oracle_set_threads(int(n))

    @staticmethod
    def _kind(kind):
        return KINDS[kind] if isinstance(kind, str) else int(kind)

    def _run(self, fn, kind, alpha, X, nin, nout, order):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, nin)
        n = X.shape[0]
        F = np.empty((n, nout))
        J = np.empty((n, nout * nout)) if order >= 1 else None
        H = np.empty((n, nout ** 3)) if order >= 2 else None
        p = lambda a: None if a is None else a.ctypes.data
        rc = fn(self._kind(kind), float(alpha), n, X.ctypes.data, F.ctypes.data, p(J), p(H))
        return rc, F, J, H

    def smf(self, kind, alpha, v, order=2):
        return self._run(self._smf, kind, alpha, v, 6, 6, order)

    def smf_matrix(self, kind, alpha, A, order=2):
        return self._run(self._smfm, kind, alpha, A, 9, 9, order)

    def gmf(self, kind, alpha, A, order=2):
        return self._run(self._gmf, kind, alpha, A, 9, 9, order)

    def _run_alpha(self, fn, alphas, X, order, nx=9):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, nx)
        alphas = np.ascontiguousarray(alphas, dtype=np.float64)
        n = X.shape[0]
        assert alphas.shape == (n,)
        F = np.empty((n, nx))
        J = np.empty((n, nx * nx)) if order >= 1 else None
        H = np.empty((n, nx ** 3)) if order >= 2 else None
        p = lambda a: None if a is None else a.ctypes.data
        rc = fn(2, alphas.ctypes.data, n, X.ctypes.data, F.ctypes.data, p(J), p(H))
        return rc, F, J, H

    def smf_alpha(self, alphas, v, order=2):
        return self._run_alpha(self._smfva, alphas, v, order, nx=6)

    def smf_matrix_alpha(self, alphas, A, order=2):
        return self._run_alpha(self._smfa, alphas, A, order)

    def gmf_alpha(self, alphas, A, order=2):
        return self._run_alpha(self._gmfa, alphas, A, order)

    def gmf_terms(self, kind, alpha, order, A):
        A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
        return self.lib.oracle_gmf_terms(self._kind(kind), float(alpha), int(order), A.ctypes.data)


def hencky(L, J, H=None):
    """The Newton step of the Hencky-type potential psi(X) = 1/2 |log X|^2 (reference theory.md:190: the Newton method
    for this potential with the tangent J^T J + log X : H), restated in NumPy from L = log X (n, 9), J = dL/dX (n, 81)
    and H = d2L/dX2 (n, 729) in the matrix layout: per matrix, summed over a,
        psi = 1/2 L_a L_a,   g_c = L_a J_ac,   K_ce = J_ac J_ae + L_a H_ace.
    Returns (psi (n,), g (n, 9), K (n, 81) or None).  Pinned in tests/test_hencky_oracle.py against what the
    mathematics fixes: g and K equal central differences of psi and g over X (through this oracle's gmf log), K is
    symmetric, and diagonal X gives the closed forms g = diag(log x / x), K_(cc)(cc) = (1 - log x_c) / x_c^2."""
    L = np.asarray(L, dtype=np.float64).reshape(-1, 9)
    n = L.shape[0]
    Jm = np.asarray(J, dtype=np.float64).reshape(n, 9, 9)
    psi = 0.5 * np.einsum("na,na->n", L, L)
    g = np.einsum("na,nac->nc", L, Jm)
    K = None
    if H is not None:
        Hm = np.asarray(H, dtype=np.float64).reshape(n, 9, 9, 9)
        K = (np.einsum("nac,nae->nce", Jm, Jm) + np.einsum("na,nace->nce", L, Hm)).reshape(n, 81)
    return psi, g, K


def hencky_voigt(F, J, H=None):
    """The same step in Voigt coordinates v = (X11, X22, X33, X12, X13, X23) from the Voigt-layout results of smf log
    (F (n, 6), J (n, 6, 6) = dF_i/dv_k, H (n, 6, 6, 6)): with m_i = 1 (i < 3) or 2 (the matrix entries of Voigt index i),
        psi = 1/2 m_i F_i^2,   g_k = dpsi/dv_k = m_i F_i J_ik,   K_kl = d2psi/dv_k dv_l = m_i (J_ik J_il + F_i H_ikl).
    Returns (psi (n,), g (n, 6), K (n, 36) or None).  Pinned in tests/test_hencky_oracle.py by central differences over
    v through this oracle's Voigt smf log."""
    F = np.asarray(F, dtype=np.float64).reshape(-1, 6)
    n = F.shape[0]
    m = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    Jv = np.asarray(J, dtype=np.float64).reshape(n, 6, 6)
    psi = 0.5 * np.einsum("i,ni,ni->n", m, F, F)
    g = np.einsum("i,ni,nik->nk", m, F, Jv)
    K = None
    if H is not None:
        Hv = np.asarray(H, dtype=np.float64).reshape(n, 6, 6, 6)
        K = (np.einsum("i,nik,nil->nkl", m, Jv, Jv) + np.einsum("i,ni,nikl->nkl", m, F, Hv)).reshape(n, 36)
    return psi, g, K

