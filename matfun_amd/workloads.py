"""Seeded synthetic inputs of the benchmark / parity workloads (host side, NumPy, FP64, row-major).

spd(n, seed)       symmetric positive definite A = Q diag(lam) Q^T: Q Haar-random orthogonal (QR of a Gaussian 3x3
                   with the signs of diag(R) moved into Q, Mezzadri 2007), lam_i = 10^U(-2, 2) independently; with
                   `special`, 15% of the matrices get lam_2 = lam_1 exactly, 5% lam_1 = lam_2 = lam_3, 10% a relative
                   gap lam_2 = lam_1 (1 + 1e-8); A is made exactly symmetric ((A + A^T) / 2).  `log10_range` moves the
                   eigenvalue range (e.g. (-2, log10(60)) keeps exp(A) inside the FP32 range of the mixed form).
general(n, seed)   general real matrices: 50% V diag(lam) V^-1 (lam ~ U[0.1, 10], V_ij ~ N(0,1) redrawn until
                   cond_2(V) <= 1e3); 25% a rotation-stretch Q blockdiag(s R2(theta), mu) Q^T (Q takes e3 to a random
                   axis; theta ~ U(-pi + 0.05, pi - 0.05), s, mu ~ U[0.5, 2]: a complex pair); 15% simple shear
                   I + gamma e1 e2^T (gamma ~ N(0,1), defective); 10% 4 I + N(0,1) entries.
voigt(A)           (A11, A22, A33, A12, A13, A23) of symmetric matrices.
"""
import numpy as np


def haar(rng, n):
    G = rng.standard_normal((n, 3, 3))
    Q, R = np.linalg.qr(G)
    d = np.sign(np.diagonal(R, axis1=1, axis2=2))
    d[d == 0] = 1.0
    return Q * d[:, None, :]


def spd(n, seed=0, special=True, log10_range=(-2.0, 2.0)):
    rng = np.random.default_rng(seed)
    Q = haar(rng, n)
    lam = 10.0 ** rng.uniform(log10_range[0], log10_range[1], (n, 3))
    if special:
        u = rng.uniform(0.0, 1.0, n)
        dbl, tri, gap = u < 0.15, (u >= 0.15) & (u < 0.20), (u >= 0.20) & (u < 0.30)
        lam[dbl, 1] = lam[dbl, 0]
        lam[tri, 1] = lam[tri, 0]
        lam[tri, 2] = lam[tri, 0]
        lam[gap, 1] = lam[gap, 0] * (1.0 + 1e-8)
    A = np.einsum("nij,nj,nkj->nik", Q, lam, Q)
    A = 0.5 * (A + np.swapaxes(A, 1, 2))
    return np.ascontiguousarray(A.reshape(n, 9))


def _rot_to_axis(rng, n):
    """rotations taking e3 to a random unit axis (Householder-free: random orthogonal then fix the third column)"""
    Q = haar(rng, n)
    return Q


def general(n, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, n)
    out = np.empty((n, 3, 3))
    # 50%: V diag(lam) V^-1 with cond(V) <= 1e3
    sel = np.nonzero(u < 0.50)[0]
    V = rng.standard_normal((sel.size, 3, 3))
    for _ in range(100):
        bad = np.linalg.cond(V) > 1e3
        if not bad.any():
            break
        V[bad] = rng.standard_normal((int(bad.sum()), 3, 3))
    lam = rng.uniform(0.1, 10.0, (sel.size, 3))
    out[sel] = np.einsum("nij,nj,njk->nik", V, lam, np.linalg.inv(V))
    # 25%: complex pair, Q blockdiag(s R2(theta), mu) Q^T
    sel = np.nonzero((u >= 0.50) & (u < 0.75))[0]
    th = rng.uniform(-np.pi + 0.05, np.pi - 0.05, sel.size)
    s = rng.uniform(0.5, 2.0, sel.size)
    mu = rng.uniform(0.5, 2.0, sel.size)
    B = np.zeros((sel.size, 3, 3))
    B[:, 0, 0] = s * np.cos(th)
    B[:, 0, 1] = -s * np.sin(th)
    B[:, 1, 0] = s * np.sin(th)
    B[:, 1, 1] = s * np.cos(th)
    B[:, 2, 2] = mu
    Q = _rot_to_axis(rng, sel.size)
    out[sel] = Q @ B @ np.swapaxes(Q, 1, 2)
    # 15%: simple shear I + gamma e1 e2^T
    sel = np.nonzero((u >= 0.75) & (u < 0.90))[0]
    out[sel] = np.eye(3)
    out[sel, 0, 1] = rng.standard_normal(sel.size)
    # 10%: 4 I + N(0, 1)
    sel = np.nonzero(u >= 0.90)[0]
    out[sel] = 4.0 * np.eye(3) + rng.standard_normal((sel.size, 3, 3))
    return np.ascontiguousarray(out.reshape(n, 9))


def voigt(A):
    A = np.asarray(A).reshape(-1, 9)
    return np.ascontiguousarray(A[:, [0, 4, 8, 1, 2, 5]])
