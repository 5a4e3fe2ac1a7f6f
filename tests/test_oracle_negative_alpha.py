"""Per-matrix exponents alpha <= 0 (an extension: the reference API requires alpha > 0; configurations with alpha drawn
from U[-1, 1] need it).  The oracle's direct evaluation at alpha < 0 is pinned through the identity
A^alpha = (A^-1)^|alpha| evaluated by the oracle at the positive exponent |alpha| (which is bitwise the reference's
computation) and the chain rule:
    B = A^-1 (adjugate / det in long double, then rounded),  dB[E] = -B E B,  d2B[E1, E2] = B E1 B E2 B + B E2 B E1 B
    F(alpha; A) = F(|alpha|; B),  J(alpha; A)[E] = J(|alpha|; B)[dB[E]],
    H(alpha; A)[E1, E2] = H(|alpha|; B)[dB[E1], dB[E2]] + J(|alpha|; B)[d2B[E1, E2]]
with the symmetrised directions (E + E^T) / 2 for smf (its derivatives are those of A -> f((A + A^T)/2)).
Special cases: alpha = -1 gives A^-1; alpha = 0 gives I with J = H = 0.
Tolerance 1e-10 per matrix (normwise): the two sides are different floating-point computations of the same function
on matrices with eigenvalues in [0.1, 10] (condition <= 1e2)."""
import numpy as np
import pytest

from oracle.pyoracle import Checker
from matfun_amd import workloads as W
from tests.parity import normwise


def inv_ld(A):
    """inverses of 3x3 matrices by the adjugate in long double (correctly rounded to a few ulp)"""
    M = np.asarray(A, dtype=np.longdouble).reshape(-1, 3, 3)
    C = np.empty_like(M)
    for i in range(3):
        for j in range(3):
            r = [k for k in range(3) if k != i]
            c = [k for k in range(3) if k != j]
            C[:, j, i] = (-1) ** (i + j) * (M[:, r[0], c[0]] * M[:, r[1], c[1]] - M[:, r[0], c[1]] * M[:, r[1], c[0]])
    det = (M[:, 0, 0] * C[:, 0, 0] + M[:, 0, 1] * C[:, 1, 0] + M[:, 0, 2] * C[:, 2, 0])
    return (C / det[:, None, None]).astype(np.float64).reshape(-1, 9)


def directions(sym):
    E = np.zeros((9, 3, 3))
    for c in range(9):
        E[c, c // 3, c % 3] = 1.0
    return 0.5 * (E + np.swapaxes(E, 1, 2)) if sym else E


def chain(B, Fp, Jp, Hp, sym):
    """F, J, H at alpha from F, J, H at |alpha| evaluated at B = A^-1 (all in the 3x3 matrix layout)"""
    n = len(B)
    B = B.reshape(n, 3, 3)
    E = directions(sym)
    dB = -np.einsum("nik,ckl,nlj->ncij", B, E, B).reshape(n, 9, 9)            # [n, c, p]
    BEB = np.einsum("nik,ckl,nlj->ncij", B, E, B)                              # [n, c, 3, 3]
    # d2B[E_c, E_e] = B E_c B E_e B + B E_e B E_c B
    BEBE = np.einsum("ncik,ekl,nlj->nceij", BEB, E, B)                          # B E_c B E_e B
    d2B = (BEBE + np.swapaxes(BEBE, 1, 2)).reshape(n, 9, 9, 9)                  # [n, c, e, p]
    J = np.einsum("nmp,ncp->nmc", Jp.reshape(n, 9, 9), dB)
    H = (np.einsum("nmpq,ncp,neq->nmce", Hp.reshape(n, 9, 9, 9), dB, dB) +
         np.einsum("nmp,ncep->nmce", Jp.reshape(n, 9, 9), d2B))
    return Fp, J.reshape(n, 81), H.reshape(n, 729)


@pytest.fixture(scope="module")
def oracle():
    return Checker("oracle")


@pytest.mark.parametrize("alpha", [-1.0, -0.75, -0.3, -1e-3])
def test_smf_negative_alpha_by_inverse(oracle, alpha):
    A = W.spd(300, seed=41, log10_range=(-1.0, 1.0))
    n = len(A)
    _, F, J, H = oracle.smf_matrix_alpha(np.full(n, alpha), A, 2)
    B = inv_ld(A)
    B = 0.5 * (B + B.reshape(n, 3, 3).transpose(0, 2, 1).reshape(n, 9))
    _, Fp, Jp, Hp = oracle.smf_matrix_alpha(np.full(n, -alpha), B, 2)
    Fr, Jr, Hr = chain(B, Fp, Jp, Hp, sym=True)
    for name, x, y in (("F", F, Fr), ("J", J, Jr), ("H", H, Hr)):
        e = normwise(x, y)
        assert e.max() <= 1e-10, (alpha, name, e.max())


@pytest.mark.parametrize("alpha", [-1.0, -0.5, -0.3])
def test_gmf_negative_alpha_by_inverse(oracle, alpha):
    G = W.general(400, seed=42)
    # keep the well-conditioned matrices inside the domain (no real eigenvalue <= 0, moderate condition)
    ok = np.array([np.linalg.cond(g.reshape(3, 3)) < 50 for g in G])
    G = G[ok][:150]
    n = len(G)
    _, F, J, H = oracle.gmf_alpha(np.full(n, alpha), G, 2)
    bad = np.isnan(F).any(axis=1)
    B = inv_ld(G)
    _, Fp, Jp, Hp = oracle.gmf_alpha(np.full(n, -alpha), B, 2)
    Fr, Jr, Hr = chain(B, Fp, Jp, Hp, sym=False)
    for name, x, y in (("F", F, Fr), ("J", J, Jr), ("H", H, Hr)):
        e = normwise(x[~bad], y[~bad])
        assert e.max() <= 1e-9, (alpha, name, e.max())


def test_alpha_minus_one_is_the_inverse(oracle):
    A = W.spd(200, seed=43, log10_range=(-1.0, 1.0))
    _, F, _, _ = oracle.smf_matrix_alpha(np.full(len(A), -1.0), A, 0)
    assert normwise(F, inv_ld(A)).max() <= 1e-13
    G = W.general(200, seed=44)
    _, Fg, _, _ = oracle.gmf_alpha(np.full(len(G), -1.0), G, 0)
    ok = ~np.isnan(Fg).any(axis=1) & np.array([np.linalg.cond(g.reshape(3, 3)) < 1e3 for g in G])
    assert normwise(Fg[ok], inv_ld(G)[ok]).max() <= 1e-11


def test_alpha_zero_is_identity_with_zero_derivatives(oracle):
    A = W.spd(100, seed=45)
    _, F, J, H = oracle.smf_matrix_alpha(np.zeros(len(A)), A, 2)
    assert np.abs(F - np.eye(3).ravel()).max() <= 1e-14
    assert np.abs(J).max() == 0.0 and np.abs(H).max() == 0.0
    G = W.general(100, seed=46)
    _, Fg, Jg, Hg = oracle.gmf_alpha(np.zeros(len(G)), G, 2)
    ok = ~np.isnan(Fg).any(axis=1)
    assert np.abs(Fg[ok] - np.eye(3).ravel()).max() <= 1e-13
    # gmf assembles J and H from residue sums that cancel to zero only up to rounding (eigenvalue separations of the
    # random general matrices reach 1e-2, amplifying it by their inverse powers)
    assert np.abs(Jg[ok]).max() <= 1e-11 and np.abs(Hg[ok]).max() <= 1e-9


def voigt_to_matrix(Fv, Jv, Hv):
    """Voigt results -> matrix layout: entry (a, b) reads Voigt component k(a, b); derivative directions E_cd with
    c != d carry weight 1/2 (each Voigt off-diagonal component moves A_cd and A_dc)"""
    vm = np.array([0, 3, 4, 3, 1, 5, 4, 5, 2])
    w = np.array([1.0 if m % 4 == 0 else 0.5 for m in range(9)])
    n = len(Fv)
    F = Fv[:, vm]
    J = Jv.reshape(n, 6, 6)[:, vm][:, :, vm] * w[None, None, :]
    H = Hv.reshape(n, 6, 6, 6)[:, vm][:, :, vm][:, :, :, vm] * (w[:, None] * w[None, :])[None, None]
    return F, J.reshape(n, 81), H.reshape(n, 729)


def test_smf_voigt_alpha_equals_matrix_alpha(oracle):
    """the Voigt-layout per-matrix exponents give the matrix-layout results (pinned above) after the layout map"""
    A = W.spd(200, seed=47)
    al = np.random.default_rng(48).uniform(-1.0, 1.0, len(A))
    al[:3] = [0.0, -1.0, 1.0]
    v = A.reshape(-1, 3, 3)[:, [0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
    rc, Fv, Jv, Hv = oracle.smf_alpha(al, v, 2)
    assert rc == 0
    _, F, J, H = oracle.smf_matrix_alpha(al, A, 2)
    for x, y in zip(voigt_to_matrix(Fv, Jv, Hv), (F, J, H)):
        assert np.array_equal(x, y)
    # and for exponents > 0, the reference semantics: one uniform call per exponent
    for a in (0.3, 2.5):
        _, Fu, Ju, Hu = oracle.smf("pow", a, v, 2)
        _, Fa, Ja, Ha = oracle.smf_alpha(np.full(len(v), a), v, 2)
        assert np.array_equal(Fu, Fa) and np.array_equal(Ju, Ja) and np.array_equal(Hu, Ha)


def test_non_finite_exponent_rejected(oracle):
    A = W.spd(4, seed=49)
    for bad in (np.nan, np.inf):
        al = np.array([0.5, bad, 0.5, 0.5])
        assert oracle.smf_matrix_alpha(al, A, 1)[0] == -1
        assert oracle.gmf_alpha(al, A, 1)[0] == -1


def test_gmf_alpha_positive_equals_uniform_calls(oracle):
    """oracle_gmf_eval_alpha at alpha > 0 is the reference's computation: bitwise equal to the uniform-alpha entry
    point (which is itself bitwise the reference's scalar build, tests/test_oracle_fixtures.py), for exponents drawn
    per matrix among 0.3, 1.7, 60.5 (60.5: radius 3/alpha inside the Taylor cluster).  A per-matrix error that scaled
    alpha would pass the inverse identity above (A^(c alpha) = (A^-1)^(c |alpha|)) but not this."""
    G = W.general(600, seed=50)
    rng = np.random.default_rng(51)
    choices = np.array([0.3, 1.7, 60.5])
    al = choices[rng.integers(0, 3, len(G))]
    for order in (0, 1, 2):
        rc, F, J, H = oracle.gmf_alpha(al, G, order)
        assert rc == 0
        for a in choices:
            s = al == a
            _, Fu, Ju, Hu = oracle.gmf("pow", a, G[s], order)
            for x, y in ((F, Fu), (J, Ju), (H, Hu)):
                if y is not None:
                    assert np.array_equal(x[s], y, equal_nan=True), (a, order)


@pytest.mark.skipif(not __import__("oracle.pyoracle", fromlist=["available"]).available("ref"),
                    reason="reference C library not built (oracle/_ref)")
def test_gmf_alpha_positive_equals_reference_build():
    """... and against the compiled reference itself (scalar variant, fresh process for MATFUN_SIMD)"""
    import os
    import subprocess
    import sys
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r)
from oracle.pyoracle import Checker
from matfun_amd import workloads as W
o, r = Checker('oracle'), Checker('ref')
assert r.lib.smf_simd_name() == b'scalar'
G = W.general(800, seed=52)
for a in (0.3, 1.7, 60.5):
    for order in (0, 1, 2):
        x, y = o.gmf_alpha(np.full(len(G), a), G, order), r.gmf('pow', a, G, order)
        for p, q in zip(x[1:], y[1:]):
            if q is not None:
                assert np.array_equal(p, q, equal_nan=True), (a, order)
print('ok')
""" % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MATFUN_SIMD="scalar", OMP_NUM_THREADS="4")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "ok" in out.stdout, out.stderr[-2000:]


def inverse_closed_form(A, sym):
    """F = A^-1, J[E] = -A^-1 E A^-1, H[E1, E2] = A^-1 E1 A^-1 E2 A^-1 + A^-1 E2 A^-1 E1 A^-1 (the derivatives of
    the inverse; directions symmetrised for smf), in the matrix layout"""
    n = len(A)
    B = inv_ld(A).reshape(n, 3, 3)
    E = directions(sym)
    J = -np.einsum("nik,ckl,nlj->nijc", B, E, B)
    BEB = np.einsum("nik,ckl,nlj->ncij", B, E, B)
    BEBEB = np.einsum("ncik,ekl,nlj->nijce", BEB, E, B)
    H = BEBEB + np.swapaxes(BEBEB, 3, 4)
    return B.reshape(n, 9), J.reshape(n, 81), H.reshape(n, 729)


def test_alpha_minus_one_derivatives_are_those_of_the_inverse(oracle):
    """alpha = -1 at order 2 against the closed-form derivatives of the matrix inverse (no evaluation at a positive
    exponent involved): smf on SPD matrices (symmetric directions), gmf on well-conditioned general matrices"""
    A = W.spd(200, seed=53, log10_range=(-1.0, 1.0))
    _, F, J, H = oracle.smf_matrix_alpha(np.full(len(A), -1.0), A, 2)
    for x, y in zip((F, J, H), inverse_closed_form(A, sym=True)):
        assert normwise(x, y).max() <= 1e-13             # (measured 2.8e-15 / 5.4e-15 / 8.3e-15)
    G = W.general(400, seed=54)
    ok = np.array([np.linalg.cond(g.reshape(3, 3)) < 50 for g in G])
    G = G[ok][:150]
    _, F, J, H = oracle.gmf_alpha(np.full(len(G), -1.0), G, 2)
    good = ~np.isnan(F).any(axis=1)
    assert good.sum() > 100
    for x, y in zip((F, J, H), inverse_closed_form(G, sym=False)):
        assert normwise(x[good], y[good]).max() <= 1e-13  # (measured 1.2e-15 / 2.3e-15 / 1.1e-14)
