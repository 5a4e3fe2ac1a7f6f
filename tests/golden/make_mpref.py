"""Writes tests/golden/mpref.npz: an independent high-precision reference for samples of the benchmark distributions
(none of the reference's or this repository's formulas involved: mpmath eigen-decompositions and central differences
in many digits).

    python tests/golden/make_mpref.py          (about a minute on 8 cores)

Symmetric (Voigt layout, as python/tests/test_mpref.py:33-63): 6 matrices of each class of the headline distribution
(matfun_amd.workloads.spd, seed 11: generic, two equal eigenvalues, three equal, relative gap 1e-8) x exp, log, sqrt,
pow 0.3; F from mpmath's symmetric eigensolver in 60 digits, J and H by central differences (h = 1e-20 s, 1e-15 s).

General (matrix layout): 5 matrices of each class of the general mix (matfun_amd.workloads.general, seed 12:
similarity with cond(V) <= 1e3, complex pair, simple shear -- defective --, 4 I + N(0, 1)) x exp, log, sqrt, pow 0.3,
in 120 digits: f(X) = V f(D) V^-1 from mpmath's eigensolver (principal branch) at X + d M and X - d M averaged
(d = 1e-45 s, M a fixed random direction; the average is f(X) + O(d^2), and the eigenvectors of a defective X stay
resolvable after the split), J and H by central differences (h = 1e-30 s, 1e-20 s), H over all pairs of
directions.  Matrices with a real eigenvalue <= 0 are skipped for log / sqrt / pow (outside the domain)."""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from matfun_amd import workloads as W   # noqa: E402
from tests.refchecks import mp_reference_voigt   # noqa: E402

FUNCS = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3)]


def smf_samples():
    """6 per class of the headline distribution, classes recovered from the eigenvalues"""
    A = W.spd(4000, seed=11)
    lam = np.linalg.eigvalsh(A.reshape(-1, 3, 3))
    rel = np.diff(lam, axis=1) / lam[:, 1:]
    tri = (rel < 1e-12).all(axis=1)
    dbl = (rel < 1e-12).any(axis=1) & ~tri
    gap = ((rel > 1e-9) & (rel < 1e-7)).any(axis=1)
    gen = ~tri & ~dbl & ~gap
    out, labels = [], []
    for name, sel in (("generic", gen), ("double", dbl), ("triple", tri), ("gap1e-8", gap)):
        idx = np.nonzero(sel)[0][:6]
        out.append(A[idx])
        labels += [name] * len(idx)
    return np.concatenate(out), labels


def gmf_samples():
    """5 per class of the general mix: the generator draws the class from a uniform u first (< 0.5 similarity,
    < 0.75 complex pair, < 0.9 shear, else 4 I + N)"""
    n = 400
    G = W.general(n, seed=12)
    u = np.random.default_rng(12).uniform(0.0, 1.0, n)
    out, labels = [], []
    for name, sel in (("similarity", u < 0.5), ("complex pair", (u >= 0.5) & (u < 0.75)),
                      ("shear", (u >= 0.75) & (u < 0.9)), ("4I+N", u >= 0.9)):
        idx = np.nonzero(sel)[0][:5]
        out.append(G[idx])
        labels += [name] * len(idx)
    return np.concatenate(out), labels


def gmf_reference(args):
    kind, alpha, a = args
    import mpmath as mp
    with mp.workdps(120):
        A0 = mp.matrix(3, 3)
        for k in range(9):
            A0[k // 3, k % 3] = mp.mpf(float(a[k]))
        s = max(abs(x) for x in A0)
        rng = np.random.default_rng(99)
        M = mp.matrix(rng.standard_normal((3, 3)).tolist())
        d = s * mp.mpf("1e-45")
        al = mp.mpf(0.5) if kind == "sqrt" else mp.mpf(alpha)

        def fe(e):
            if kind == "exp":
                return mp.exp(e)
            if kind == "log":
                return mp.log(e)
            return mp.exp(al * mp.log(e))

        def feig(X):
            E, V = mp.eig(X)
            return V * mp.diag([fe(e) for e in E]) * mp.inverse(V)

        def F(X):
            Y = (feig(X + d * M) + feig(X - d * M)) / 2
            return mp.matrix([[mp.re(Y[i, j]) for j in range(3)] for i in range(3)])

        def E(c):
            Z = mp.zeros(3, 3)
            Z[c // 3, c % 3] = 1
            return Z

        flat = lambda X: [float(X[k // 3, k % 3]) for k in range(9)]
        h1, h2 = s * mp.mpf("1e-30"), s * mp.mpf("1e-20")
        J = np.zeros((9, 9))
        for c in range(9):
            J[:, c] = flat((F(A0 + h1 * E(c)) - F(A0 - h1 * E(c))) / (2 * h1))
        H = np.zeros((9, 9, 9))
        for c in range(9):
            for e in range(c, 9):
                Ec, Ee = h2 * E(c), h2 * E(e)
                v = flat((F(A0 + Ec + Ee) - F(A0 + Ec - Ee) - F(A0 - Ec + Ee) + F(A0 - Ec - Ee)) / (4 * h2 * h2))
                H[:, c, e] = H[:, e, c] = v
        return np.array(flat(F(A0))), J.ravel(), H.ravel()


def smf_reference(args):
    kind, alpha, v = args
    F, J, H = mp_reference_voigt(kind, alpha, v, dps=60)
    return F, J.ravel(), H.ravel()


def in_domain(a):
    ev = np.linalg.eigvals(a.reshape(3, 3))
    return not np.any((np.abs(ev.imag) < 1e-12 * np.abs(ev).max()) & (ev.real <= 0))


def main():
    S, sl = smf_samples()
    V = W.voigt(S)
    G, gl = gmf_samples()
    sjobs = [(k, al, v) for k, al in FUNCS for v in V]
    skeys = [(k, al, i) for k, al in FUNCS for i in range(len(V))]
    gjobs, gkeys = [], []
    for k, al in FUNCS:
        for i, a in enumerate(G):
            if k == "exp" or in_domain(a):
                gjobs.append((k, al, a))
                gkeys.append((k, al, i))
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        sres = pool.map(smf_reference, sjobs)
        gres = pool.map(gmf_reference, gjobs)
    KIND = {"exp": 0, "log": 1, "pow": 2, "sqrt": 3}
    np.savez_compressed(
        os.path.join(HERE, "mpref.npz"),
        smf_kind=np.array([KIND[k] for k, _, _ in skeys]), smf_alpha=np.array([a for _, a, _ in skeys]),
        smf_v=np.array([V[i] for _, _, i in skeys]), smf_label=np.array([sl[i] for _, _, i in skeys]),
        smf_F=np.array([r[0] for r in sres]), smf_J=np.array([r[1] for r in sres]), smf_H=np.array([r[2] for r in sres]),
        gmf_kind=np.array([KIND[k] for k, _, _ in gkeys]), gmf_alpha=np.array([a for _, a, _ in gkeys]),
        gmf_A=np.array([G[i] for _, _, i in gkeys]), gmf_label=np.array([gl[i] for _, _, i in gkeys]),
        gmf_F=np.array([r[0] for r in gres]), gmf_J=np.array([r[1] for r in gres]), gmf_H=np.array([r[2] for r in gres]))
    print(f"wrote {len(sres)} symmetric and {len(gres)} general cases")


if __name__ == "__main__":
    main()
