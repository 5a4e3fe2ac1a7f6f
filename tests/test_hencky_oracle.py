"""The oracle of the Hencky Newton step (oracle.pyoracle.hencky: psi = 1/2 |log X|^2, g = J^T L, K = J^T J + L:H,
reference theory.md:190) pinned by what the mathematics fixes, not by retyping its formula:
  - g and K are the derivatives of psi and g: central differences over every entry of X (general X through the
    oracle's gmf log; symmetric X through its smf log along the symmetric directions (E_cd + E_dc) / 2);
  - K is symmetric (a Hessian);
  - diagonal X = diag(x): g = diag(log x / x), K_(cc)(cc) = (1 - log x_c) / x_c^2, K_(cd)(cd) = (log[x_c, x_d])^2 for
    c != d (E_cd is nilpotent, so the second derivative of log along it vanishes).
A planted mistake (J in place of J^T) fails the difference check."""
import numpy as np
import pytest

from oracle.pyoracle import Checker, hencky


@pytest.fixture(scope="module")
def ck():
    return Checker("oracle")


def general_pos(n, seed):
    """general real matrices with eigenvalues of positive real part (log defined): 2 I + 0.6 N(0, 1)"""
    rng = np.random.default_rng(seed)
    return (2.0 * np.eye(3) + 0.6 * rng.standard_normal((n, 3, 3))).reshape(n, 9)


def spd_pos(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, 3, 3))
    return (np.einsum("nij,nkj->nik", M, M) + 0.5 * np.eye(3)).reshape(n, 9)


def psi_g(ck, fam, X, order):
    if fam == "gmf":
        rc, L, J, H = ck.gmf("log", 0.0, X, order)
    else:
        rc, L, J, H = ck.smf_matrix("log", 0.0, X, order)
    assert rc == 0
    return hencky(L, J, H)


def directions(sym):
    E = np.zeros((9, 9))
    for c in range(9):
        E[c, c] = 1.0
    if sym:
        E = 0.5 * (E + E.reshape(9, 3, 3).transpose(0, 2, 1).reshape(9, 9))
    return E


@pytest.mark.parametrize("fam", ["gmf", "smf"])
def test_gradient_and_tangent_are_derivatives(ck, fam):
    n = 24
    X = general_pos(n, 1) if fam == "gmf" else spd_pos(n, 2)
    psi, g, K = psi_g(ck, fam, X, 2)
    E = directions(fam == "smf")
    h = 1e-5
    for c in range(9):
        Xp, Xm = X + h * E[c], X - h * E[c]
        pp, gp, _ = psi_g(ck, fam, Xp, 1)
        pm, gm, _ = psi_g(ck, fam, Xm, 1)
        d_psi = (pp - pm) / (2 * h)
        d_g = (gp - gm) / (2 * h)
        # directional derivatives along E[c]: g . E[c] and K E[c]
        g_dir = g @ E[c]
        K_dir = np.einsum("nce,e->nc", K.reshape(n, 9, 9), E[c])
        assert np.max(np.abs(d_psi - g_dir) / (1 + np.abs(g_dir))) < 1e-8, (fam, c)
        err = np.linalg.norm(d_g - K_dir, axis=1) / np.maximum(np.linalg.norm(K_dir, axis=1), 1.0)
        assert err.max() < 1e-7, (fam, c, err.max())


def test_planted_transpose_fails(ck):
    n = 8
    X = general_pos(n, 3)
    rc, L, J, H = ck.gmf("log", 0.0, X, 2)
    g_wrong = np.einsum("nc,nac->na", L, J.reshape(n, 9, 9))          # J L instead of J^T L
    h = 1e-5
    E = directions(False)
    worst = 0.0
    for c in range(9):
        pp = psi_g(ck, "gmf", X + h * E[c], 1)[0]
        pm = psi_g(ck, "gmf", X - h * E[c], 1)[0]
        worst = max(worst, np.max(np.abs((pp - pm) / (2 * h) - g_wrong[:, c])))
    assert worst > 1e-3


@pytest.mark.parametrize("fam", ["gmf", "smf"])
def test_tangent_symmetric(ck, fam):
    n = 64
    X = general_pos(n, 4) if fam == "gmf" else spd_pos(n, 5)
    _, _, K = psi_g(ck, fam, X, 2)
    K = K.reshape(n, 9, 9)
    err = np.linalg.norm(K - K.transpose(0, 2, 1), axis=(1, 2)) / np.linalg.norm(K, axis=(1, 2))
    assert err.max() < 1e-13


@pytest.mark.parametrize("fam", ["gmf", "smf"])
def test_diagonal_closed_forms(ck, fam):
    x = np.array([[0.5, 2.0, 7.0], [1.0, 1.5, 0.25], [3.0, 3.5, 40.0]])
    n = len(x)
    X = np.zeros((n, 3, 3))
    for i in range(3):
        X[:, i, i] = x[:, i]
    psi, g, K = psi_g(ck, fam, X.reshape(n, 9), 2)
    lx = np.log(x)
    np.testing.assert_allclose(psi, 0.5 * (lx * lx).sum(axis=1), rtol=1e-14)
    g_ref = np.zeros((n, 3, 3))
    for i in range(3):
        g_ref[:, i, i] = lx[:, i] / x[:, i]
    np.testing.assert_allclose(g.reshape(n, 3, 3), g_ref, rtol=1e-13, atol=1e-15)
    K = K.reshape(n, 9, 9)
    for c in range(3):
        np.testing.assert_allclose(K[:, 4 * c, 4 * c], (1 - lx[:, c]) / x[:, c] ** 2, rtol=1e-12)
    if fam == "gmf":                       # (smf's derivatives act on the symmetric part: E_cd alone is not one)
        for c, d in ((0, 1), (0, 2), (1, 2), (2, 0)):
            dd = (lx[:, c] - lx[:, d]) / (x[:, c] - x[:, d])
            np.testing.assert_allclose(K[:, 3 * c + d, 3 * c + d], dd * dd, rtol=1e-12)


def test_voigt_gradient_and_tangent_are_derivatives(ck):
    """hencky_voigt: g and K are the derivatives of psi and g over the Voigt coordinates (central differences through
    the oracle's Voigt smf log), and they are the matrix-layout step summed over the entries of each Voigt index"""
    from oracle.pyoracle import hencky_voigt
    n = 24
    A = spd_pos(n, 6)
    v = A[:, [0, 4, 8, 1, 2, 5]]
    rc, F, J, H = ck.smf("log", 0.0, v, 2)
    psi, g, K = hencky_voigt(F, J, H)
    h = 1e-5
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        _, Fp, Jp, _ = ck.smf("log", 0.0, v + e, 1)
        _, Fm, Jm, _ = ck.smf("log", 0.0, v - e, 1)
        pp, gp, _ = hencky_voigt(Fp, Jp)
        pm, gm, _ = hencky_voigt(Fm, Jm)
        assert np.max(np.abs((pp - pm) / (2 * h) - g[:, k]) / (1 + np.abs(g[:, k]))) < 1e-8
        dg = (gp - gm) / (2 * h)
        Kk = K.reshape(n, 6, 6)[:, :, k]
        assert (np.linalg.norm(dg - Kk, axis=1) / np.maximum(np.linalg.norm(Kk, axis=1), 1.0)).max() < 1e-7
    # against the matrix-layout step: g_p = m_p g_c, K_pq = m_p m_q K_ce for entries c of p, e of q
    pm_, gm_, Km_ = psi_g(ck, "smf", A, 2)
    rep = [0, 4, 8, 1, 2, 5]
    mlt = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    np.testing.assert_allclose(psi, pm_, rtol=1e-13)
    np.testing.assert_allclose(g, gm_[:, rep] * mlt, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(K.reshape(n, 6, 6), Km_.reshape(n, 9, 9)[:, rep][:, :, rep] * np.outer(mlt, mlt),
                               rtol=1e-11, atol=1e-13)
