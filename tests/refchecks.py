"""The reference's own edge-case checks, restated once and run against any evaluator (no test functions here):
tests/test_oracle_refchecks.py runs them on the CPU oracle, tests/test_gpu_refchecks.py on the HIP library through its
C ABI.  Each check cites the reference test it restates and keeps that test's inputs, metric (max|x - ref| / max|ref|,
c/tests/test_c.c:18-25) and tolerance; where the reference compares with a formula, the formula is written here from
the mathematics (closed forms, identities, central differences), not taken from any implementation.

An evaluator has
    smf(kind, alpha, V, order)        Voigt vectors V (n, 6)     -> F (n, 6),  J (n, 36), H (n, 216)
    smf_matrix(kind, alpha, A, order) symmetric A (n, 9)         -> F (n, 9),  J (n, 81), H (n, 729) (and the rc)
    gmf(kind, alpha, A, order)        general A (n, 9)           -> F (n, 9),  J (n, 81), H (n, 729)
(J / H None below the order) and `gpu`: True when a non-symmetric smf matrix gives NaN results and rc 0 (the GPU
API, cuda/include/matfun/matfun_cuda.h:1-3) rather than rc -3 (the CPU API, c/src/matfun.cc:82)."""
import numpy as np

EPS = np.finfo(float).eps
VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def maxrel(a, b):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def voigt_to_matrix(v):
    v = np.asarray(v, dtype=np.float64).reshape(-1, 6)
    A = np.empty((len(v), 3, 3))
    for p, (i, j) in enumerate(VOIGT):
        A[:, i, j] = A[:, j, i] = v[:, p]
    return A.reshape(-1, 9)


KINDS5 = [("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 0.3), ("pow", 1.7)]
X6 = np.array([2.0, 1.5, 1.2, 0.3, -0.2, 0.1])            # test_c.c:54


# ---- symmetric matrices (c/tests/test_c.c)

def smf_diagonal(ev):
    """test_c.c:41-52: diagonal matrix (0.5, 2, 3): F is f of the diagonal, <= 4e-15"""
    xd = np.array([[0.5, 2.0, 3.0, 0.0, 0.0, 0.0]])
    fns = {"exp": np.exp, "log": np.log, "sqrt": np.sqrt}
    for kind, al in KINDS5:
        F = ev.smf(kind, al, xd, 0)[0]
        d = xd[0, :3]
        ref = np.zeros(6)
        ref[:3] = d ** al if kind == "pow" else fns[kind](d)
        assert maxrel(F, ref) <= 4e-15, (kind, al, maxrel(F, ref))


def smf_exp_log(ev):
    """test_c.c:54-58: exp(log A) = A, <= 1e-14"""
    L = ev.smf("log", 0.0, X6[None], 0)[0]
    E = ev.smf("exp", 0.0, L, 0)[0]
    assert maxrel(E, X6) <= 1e-14, maxrel(E, X6)


def smf_central_differences(ev):
    """test_c.c:60-84: J and H against central differences (h = 1e-5) of F and J, a random matrix and one with a
    double eigenvalue (2, 2, 5), <= 1e-8"""
    xs = np.array([[2.0, 1.5, 1.2, 0.3, -0.2, 0.1], [3.0, 3.0, 3.0, 1.0, 1.0, 1.0]])
    h = 1e-5
    for x in xs:
        for kind, al in KINDS5:
            _, J0, H0 = ev.smf(kind, al, x[None], 2)
            P = np.repeat(x[None], 12, axis=0)
            for p in range(6):
                P[2 * p, p] += h
                P[2 * p + 1, p] -= h
            Fp, Jp, _ = ev.smf(kind, al, P, 1)
            Jfd = np.empty((6, 6))
            Hfd = np.empty((6, 6, 6))
            for p in range(6):
                Jfd[:, p] = (Fp[2 * p] - Fp[2 * p + 1]) / (2 * h)
                Hfd[:, :, p] = (Jp[2 * p] - Jp[2 * p + 1]).reshape(6, 6) / (2 * h)
            assert maxrel(J0, Jfd) <= 1e-8, (x, kind, al, "J", maxrel(J0, Jfd))
            assert maxrel(H0, Hfd) <= 1e-8, (x, kind, al, "H", maxrel(H0, Hfd))


def smf_batch_equals_single(ev):
    """test_c.c:86-104: a batch of 37 equals the single calls, <= 1e-14"""
    t = 0.1 * np.arange(37)
    V = np.stack([2 + np.sin(t), 1.5 + np.cos(t), 1.2 + 0.5 * np.sin(2 * t), 0.3 * np.cos(3 * t), -0.2 * np.sin(t),
                  np.full(37, 0.1)], axis=1)
    Fb, Jb, Hb = ev.smf("log", 0.0, V, 2)
    for i in range(37):
        F1, J1, H1 = ev.smf("log", 0.0, V[i:i + 1], 2)
        e = max(maxrel(Fb[i], F1), maxrel(Jb[i], J1), maxrel(Hb[i], H1))
        assert e <= 1e-14, (i, e)


def smf_matrix_layout(ev):
    """test_c.c:106-138: the matrix layout is the Voigt results with the exact weights 1 (diagonal) and 1/2
    (off-diagonal) of dv / dA_cd (tolerance 0), and J, H have the index-pair symmetries (<= 1e-15)"""
    vi = [0, 3, 4, 3, 1, 5, 4, 5, 2]
    w = np.array([1.0 if c % 4 == 0 else 0.5 for c in range(9)])
    Fv, Jv, Hv = ev.smf("log", 0.0, X6[None], 2)
    Fm, Jm, Hm = ev.smf_matrix("log", 0.0, voigt_to_matrix(X6), 2)[:3]
    Jv, Hv = Jv.reshape(6, 6), Hv.reshape(6, 6, 6)
    assert np.array_equal(Fm[0], Fv[0][vi])
    assert np.array_equal(Jm[0].reshape(9, 9), Jv[np.ix_(vi, vi)] * w[None, :])
    assert np.array_equal(Hm[0].reshape(9, 9, 9), Hv[np.ix_(vi, vi, vi)] * w[None, :, None] * w[None, None, :])
    J = Jm[0].reshape(3, 3, 3, 3)
    H = Hm[0].reshape(3, 3, 3, 3, 3, 3)
    for T, perm in ((J, (1, 0, 2, 3)), (J, (0, 1, 3, 2)), (J, (2, 3, 0, 1)), (H, (1, 0, 2, 3, 4, 5)),
                    (H, (0, 1, 3, 2, 4, 5)), (H, (0, 1, 4, 5, 2, 3))):
        assert np.max(np.abs(T - np.transpose(T, perm))) <= 1e-15


def smf_nonsymmetric_in_batch(ev):
    """test_c.c:139-148: a batch of 37 copies with a round-off asymmetry (accepted) in matrix 5 and clear ones in 7
    and 30: NaN results there only (and rc -3 on the CPU API), the others equal the single call to 1e-12"""
    A0 = voigt_to_matrix(X6)[0]
    Ab = np.repeat(A0[None], 37, axis=0)
    Ab[5, 1] += 1e-12
    Ab[7, 1] += 0.1
    Ab[30, 5] -= 0.1
    F, _, _, rc = ev.smf_matrix("log", 0.0, Ab, 0)
    assert rc == (0 if ev.gpu else -3), rc
    F1 = ev.smf_matrix("log", 0.0, A0[None], 0)[0]
    for i in range(37):
        if i in (7, 30):
            assert np.isnan(F[i]).all(), i
        else:
            assert np.max(np.abs(F[i] - F1[0])) < 1e-12, i


def smf_scaling(ev):
    """test_c.c:152-174: extreme magnitudes s = 1e+-100, 1e+-250: sqrt(sA) = sqrt(s) sqrt(A), log(sA) = log(s) I +
    log(A), J_sqrt / sqrt(s), H_sqrt / s^1.5 (1e+-100 only), J_log / s -- <= 1e-14"""
    F0, J0, H0 = ev.smf("sqrt", 0.0, X6[None], 2)
    L0, JL0, _ = ev.smf("log", 0.0, X6[None], 1)
    for k, s in enumerate((1e100, 1e-100, 1e250, 1e-250)):
        xs = s * X6[None]
        F, J, H = ev.smf("sqrt", 0.0, xs, 2 if k < 2 else 1)
        L, JL, _ = ev.smf("log", 0.0, xs, 1)
        Lr = L0 + np.array([np.log(s)] * 3 + [0.0] * 3)
        err = max(maxrel(F, np.sqrt(s) * F0), maxrel(J, J0 / np.sqrt(s)), maxrel(JL, JL0 / s), maxrel(L, Lr))
        if k < 2:
            err = max(err, maxrel(H, H0 / (s * np.sqrt(s))))
        assert err <= 1e-14, (s, err)


# ---- general matrices (c/tests/test_c.c:176-258, python/tests/test_gmf.py)

def gmf_closed_forms(ev):
    """test_c.c:184-201 and test_gmf.py:55-76: N = 0.7 e1 e2^T (N^2 = 0): exp(N) = I + N, log(I + N) = N,
    sqrt(4 I + N) = 2 I + N / 4, (I + N)^2.5 = I + 2.5 N; rotation by 0.9 about z: exp of its generator, log and sqrt of
    the rotation -- <= 1e-15; the derivatives of exp at 0: J = identity (d_ac d_bd), H_abcdef = (d_ac d_de d_fb +
    d_ae d_fc d_db) / 2 (<= 1e-16)"""
    th = 0.9
    c, sn, ch, sh = np.cos(th), np.sin(th), np.cos(th / 2), np.sin(th / 2)
    I = np.eye(3).ravel()
    Nil = np.array([0, 0.7, 0, 0, 0, 0, 0, 0, 0.0])
    K = np.array([0, -th, 0, th, 0, 0, 0, 0, 0])
    R = np.array([c, -sn, 0, sn, c, 0, 0, 0, 1])
    Rh = np.array([ch, -sh, 0, sh, ch, 0, 0, 0, 1])
    for kind, al, A, ref in (("exp", 0.0, Nil, I + Nil), ("log", 0.0, I + Nil, Nil), ("sqrt", 0.0, 4 * I + Nil,
                                                                                         2 * I + Nil / 4),
                             ("pow", 2.5, I + Nil, I + 2.5 * Nil), ("exp", 0.0, K, R), ("log", 0.0, R, K),
                             ("sqrt", 0.0, R, Rh)):
        F = ev.gmf(kind, al, A[None], 0)[0]
        assert maxrel(F, ref) <= 1e-15, (kind, al, maxrel(F, ref))
    F, J, H = ev.gmf("exp", 0.0, np.zeros((1, 9)), 2)
    d = np.eye(3)
    assert np.array_equal(F[0], I)
    assert maxrel(J, np.einsum("ac,bd->abcd", d, d)) <= 1e-16
    Href = 0.5 * (np.einsum("ac,de,fb->abcdef", d, d, d) + np.einsum("ae,fc,db->abcdef", d, d, d))
    assert maxrel(H, Href) <= 1e-16


def gmf_central_differences(ev):
    """test_c.c:203-228: J and H against central differences (h = 1e-6), a random non-symmetric matrix and a nearly
    defective one; exp, log, sqrt, pow 1.7: J <= 1e-8, H <= 1e-7"""
    Ag = np.array([[1.6, 0.3, -0.2, -0.4, 1.1, 0.5, 0.2, -0.3, 1.4], [1.2, 0.8, 0.1, 0, 1.2, 0.5, 1e-9, 0, 1.2]])
    h = 1e-6
    for A in Ag:
        for kind, al in (("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 1.7)):
            _, J0, H0 = ev.gmf(kind, al, A[None], 2)
            P = np.repeat(A[None], 18, axis=0)
            for p in range(9):
                P[2 * p, p] += h
                P[2 * p + 1, p] -= h
            Fp, Jp, _ = ev.gmf(kind, al, P, 1)
            Jfd = np.empty((9, 9))
            Hfd = np.empty((81, 9))
            for p in range(9):
                Jfd[:, p] = (Fp[2 * p] - Fp[2 * p + 1]) / (2 * h)
                Hfd[:, p] = (Jp[2 * p] - Jp[2 * p + 1]) / (2 * h)
            assert maxrel(J0, Jfd) <= 1e-8, (A, kind, "J", maxrel(J0, Jfd))
            assert maxrel(H0, Hfd) <= 1e-7, (A, kind, "H", maxrel(H0, Hfd))


def gmf_batch_and_nan(ev):
    """test_c.c:230-255: a batch of 37 equals the single calls (<= 1e-14); a real eigenvalue < 0 (matrix 11) gives
    NaN for log in that matrix only"""
    t = 0.37 * np.arange(37)
    Ab = np.stack([1.5 + np.sin(t), 0.3 * np.cos(2 * t), np.full(37, 0.2), -0.4 * np.sin(t), np.full(37, 1.2),
                   0.5 * np.cos(t), 0.1 * (np.arange(37) % 3), np.full(37, -0.3), 1.4 + 0.2 * np.cos(3 * t)], axis=1)
    Ab[11] = [-1, 0, 0, 0, 2, 0, 0, 0, 3]
    Fg, Jg, Hg = ev.gmf("log", 0.0, Ab, 2)
    for i in range(37):
        F1, J1, H1 = ev.gmf("log", 0.0, Ab[i:i + 1], 2)
        if i == 11:
            assert np.isnan(Fg[i]).all() and np.isnan(F1).all()
            continue
        e = max(maxrel(Fg[i], F1), maxrel(Jg[i], J1), maxrel(Hg[i], H1))
        assert e <= 1e-14, (i, e)


def gmf_identities(ev):
    """test_gmf.py:79-98: exp(log A) = A and sqrt(A)^2 = A (<= 1e-14), the chain rule J_exp(log A) J_log(A) = I
    (<= 1e-14 absolute), J and H against central differences (1e-9 / 1e-8, h = 1e-6), H symmetric in (cd) <-> (ef)
    (<= 1e-15); A = I + 0.4 N(0, 1) (seed 1) + e2 e1^T"""
    rng = np.random.default_rng(1)
    A = (np.eye(3) + 0.4 * rng.standard_normal((3, 3)) + np.outer([0, 1, 0], [1, 0, 0])).reshape(1, 9)
    L, JL, _ = ev.gmf("log", 0.0, A, 1)
    E, JE, _ = ev.gmf("exp", 0.0, L, 1)
    S = ev.gmf("sqrt", 0.0, A, 0)[0].reshape(3, 3)
    assert maxrel(E, A) <= 1e-14 and maxrel(S @ S, A.reshape(3, 3)) <= 1e-14
    chain = JE.reshape(9, 9) @ JL.reshape(9, 9)
    assert np.max(np.abs(chain - np.eye(9))) <= 1e-14
    h = 1e-6
    for kind, al in (("exp", 0.0), ("log", 0.0), ("pow", 1.7)):
        _, J, H = ev.gmf(kind, al, A, 2)
        P = np.repeat(A, 18, axis=0)
        for c in range(9):
            P[2 * c, c] += h
            P[2 * c + 1, c] -= h
        Fp, Jp, _ = ev.gmf(kind, al, P, 1)
        for c in range(9):
            assert maxrel((Fp[2 * c] - Fp[2 * c + 1]) / (2 * h), J.reshape(9, 9)[:, c]) <= 1e-9, (kind, c)
            assert maxrel((Jp[2 * c] - Jp[2 * c + 1]).reshape(9, 9) / (2 * h), H.reshape(9, 9, 9)[:, :, c]) <= 1e-8
        H6 = H.reshape(3, 3, 3, 3, 3, 3)
        assert maxrel(H6, np.transpose(H6, (0, 1, 4, 5, 2, 3))) <= 1e-15


def gmf_symmetric_input_matches_smf(ev):
    """test_gmf.py:101-112: for symmetric A, gmf's derivatives projected on symmetric directions are smf's (matrix
    layout): F, J <= 1e-14, H <= 1e-13; A = B B^T + 0.5 I, B ~ N(0, 1) (seed 2)"""
    rng = np.random.default_rng(2)
    B = rng.standard_normal((5, 3, 3))
    A = (B @ np.swapaxes(B, -1, -2) + 0.5 * np.eye(3)).reshape(5, 9)
    for kind, al in (("exp", 0.0), ("log", 0.0), ("sqrt", 0.0), ("pow", 1.7)):
        F, J, H = (x.reshape((5,) + (3,) * k) for x, k in zip(ev.gmf(kind, al, A, 2), (2, 4, 6)))
        Fs, Js, Hs = (x.reshape((5,) + (3,) * k) for x, k in zip(ev.smf_matrix(kind, al, A, 2)[:3], (2, 4, 6)))
        Jp = 0.5 * (J + np.swapaxes(J, -1, -2))
        Hp = 0.5 * (H + np.swapaxes(H, -3, -4))
        Hp = 0.5 * (Hp + np.swapaxes(Hp, -1, -2))
        assert maxrel(F, Fs) <= 1e-14 and maxrel(Jp, Js) <= 1e-14 and maxrel(Hp, Hs) <= 1e-13, kind


def gmf_domain_and_batch(ev):
    """test_gmf.py:115-132: NaN for a real eigenvalue <= 0 (log, sqrt) only for that matrix (4); a complex pair with
    negative real part (-1 +- 2i, matrix 9) is in the domain; batch results equal single calls (F, J <= 1e-14,
    H <= 1e-13); exp(log A9) = A9; exp finite everywhere"""
    rng = np.random.default_rng(3)
    A = 1.5 * np.eye(3) + 0.3 * rng.standard_normal((13, 3, 3))
    A[4] = np.diag([-1.0, 2.0, 3.0])
    A[9] = np.array([[-1.0, 2.0, 0], [-2.0, -1.0, 0], [0, 0, 1.0]])
    A = A.reshape(13, 9)
    for kind in ("log", "sqrt"):
        F, J, H = ev.gmf(kind, 0.0, A, 2)
        assert np.isnan(F).any(axis=1).tolist() == [i == 4 for i in range(13)], kind
        for i in (0, 9, 12):
            Fi, Ji, Hi = ev.gmf(kind, 0.0, A[i:i + 1], 2)
            assert maxrel(Fi, F[i]) <= 1e-14 and maxrel(Ji, J[i]) <= 1e-14 and maxrel(Hi, H[i]) <= 1e-13
    L = ev.gmf("log", 0.0, A[9:10], 0)[0]
    assert maxrel(ev.gmf("exp", 0.0, L, 0)[0], A[9]) <= 1e-14
    assert np.isfinite(ev.gmf("exp", 0.0, A, 2)[2]).all()


def gmf_rotations_near_180(ev, deg):
    """test_gmf.py:152-168: rotations by theta close to 180 degrees about (1, 2, 3) / sqrt(14), in an orthonormal and a
    non-normal basis (I + 0.3 N(0, 1), seed 4): log R = theta K, sqrt R = rotation by theta / 2, R^0.25 = rotation by
    theta / 4 (closed forms), to 50 eps / (pi - theta)"""
    th = np.radians(deg)
    k = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    rot = lambda t: np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K
    X = np.eye(3) + 0.3 * np.random.default_rng(4).standard_normal((3, 3))
    tol = 50 * EPS / (np.pi - th)
    for Bm in (np.eye(3), X):
        Bi = np.linalg.inv(Bm)
        R = (Bm @ rot(th) @ Bi).reshape(1, 9)
        for kind, al, ref in (("log", 0.0, Bm @ (th * K) @ Bi), ("sqrt", 0.0, Bm @ rot(th / 2) @ Bi),
                              ("pow", 0.25, Bm @ rot(th / 4) @ Bi)):
            F = ev.gmf(kind, al, R, 0)[0]
            assert maxrel(F, ref) < tol, (deg, kind, maxrel(F, ref), tol)


# ---- independent high precision for powers with large exponents (python/tests/test_mpref.py:93-101)

def mp_reference_voigt(kind, alpha, v, dps=60):
    """F (6), J (6, 6), H (6, 6, 6) (Voigt layout) of the symmetric matrix with Voigt vector v in `dps` digits:
    F from mpmath's symmetric eigensolver, J and H by central differences of F in the same precision (none of the
    divided-difference formulas involved)"""
    import mpmath as mp
    with mp.workdps(dps):
        a = mp.mpf(alpha)
        f = {"exp": mp.exp, "log": mp.log, "sqrt": mp.sqrt}.get(kind, lambda x: x ** a)
        A0 = mp.matrix(3, 3)
        for p, (i, j) in enumerate(VOIGT):
            A0[i, j] = A0[j, i] = mp.mpf(float(v[p]))

        def F(A):
            E, Q = mp.eigsy(A)
            return Q * mp.diag([f(E[i]) for i in range(3)]) * Q.T

        def Ev(p):
            M = mp.zeros(3, 3)
            i, j = VOIGT[p]
            M[i, j] = M[j, i] = 1
            return M

        vec = lambda M: [float(M[i, j]) for i, j in VOIGT]
        s = max(abs(x) for x in A0)
        h1, h2 = s * mp.mpf("1e-20"), s * mp.mpf("1e-15")      # truncation O(h^2), round-off 10^-dps / h^2
        J = np.zeros((6, 6))
        for p in range(6):
            J[:, p] = vec((F(A0 + h1 * Ev(p)) - F(A0 - h1 * Ev(p))) / (2 * h1))
        H = np.zeros((6, 6, 6))
        for p in range(6):
            for q in range(p, 6):
                Ep, Eq = h2 * Ev(p), h2 * Ev(q)
                H[:, p, q] = H[:, q, p] = vec((F(A0 + Ep + Eq) - F(A0 + Ep - Eq) - F(A0 - Ep + Eq) + F(A0 - Ep - Eq))
                                              / (4 * h2 * h2))
        return np.array(vec(F(A0))), J, H


LARGE_ALPHA = [0.3, 1.7, 12.0, 13.5, 40.5, 60.5, 100.5, 300.5]
LAMBDAS = [(0.8, 1.0, 1.25), (1.0, 1.04, 1.06), (0.95, 1.0, 1.06)]


def large_alpha_inputs():
    Q = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))[0]
    out = []
    for lam in LAMBDAS:
        A = Q @ np.diag(lam) @ Q.T
        out.append(np.array([A[i, j] for i, j in VOIGT]))
    return out


def pow_large_alpha(ev, refs):
    """test_mpref.py:93-101: x^alpha with alpha up to 300.5 for eigenvalues in and near the cluster of the Taylor
    series, against the 60-digit reference `refs[(alpha, lambda set)]`, to max(1e-13, 10 alpha eps) (x^alpha has
    condition number alpha)"""
    for j, v in enumerate(large_alpha_inputs()):
        for al in LARGE_ALPHA:
            out = ev.smf("pow", al, v[None], 2)
            tol = max(1e-13, 10 * al * EPS)
            for name, x, r in zip("FJH", out, refs[(al, j)]):
                assert maxrel(x, r) <= tol, (al, LAMBDAS[j], name, maxrel(x, r), tol)


SMF_CHECKS = [smf_diagonal, smf_exp_log, smf_central_differences, smf_batch_equals_single, smf_matrix_layout,
              smf_nonsymmetric_in_batch, smf_scaling]
GMF_CHECKS = [gmf_closed_forms, gmf_central_differences, gmf_batch_and_nan, gmf_identities,
              gmf_symmetric_input_matches_smf, gmf_domain_and_batch]
ROTATION_DEGREES = [170.0, 179.0, 179.9, 179.999, 179.99999]
