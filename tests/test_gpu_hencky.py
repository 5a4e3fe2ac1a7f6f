"""mf_hip_hencky / mf_hip_hencky_mixed -- value, gradient and tangent of psi(X) = 1/2 |log X|^2 from L = log X, J, H
(reference theory.md:190) -- on the GPU through the C ABI, against the oracle (oracle.pyoracle.hencky, pinned by
finite differences and closed forms in tests/test_hencky_oracle.py) on the same L, J, H, which come from this
library's own gmf / smf log at order 2.
Tolerances (per matrix, normwise): FP64 1e-13 (the kernel's FMA chain and NumPy's sums differ in order only);
FP32 form 1e-6 against the oracle on the same FP32 inputs (the kernel sums in FP64 and rounds once to FP32).
At 2^20 matrices: against a plain PyTorch FP64 reference on the device, and the symmetry of K."""
import numpy as np
import pytest

import matfun_amd
from oracle.pyoracle import hencky as hencky_oracle
from tests.gpu_util import call
from tests.parity import normwise

pytestmark = pytest.mark.gpu


def general_pos(n, seed):
    rng = np.random.default_rng(seed)
    return (2.0 * np.eye(3) + 0.6 * rng.standard_normal((n, 3, 3))).reshape(n, 9)


def spd_pos(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, 3, 3))
    return (np.einsum("nij,nkj->nik", M, M) + 0.5 * np.eye(3)).reshape(n, 9)


def run(L, J, H, want=(True, True, True), mixed=False):
    """mf_hip_hencky(_mixed) on host arrays L (n,9), J (n,81), H (n,729) or None; outputs as asked (psi, g, K),
    prefilled with 7 (never a result by accident); returns rc and the outputs (NumPy, float64)"""
    import torch
    lib = matfun_amd.library()
    dt = torch.float32 if mixed else torch.float64
    n = L.shape[0]
    dev = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)
    dL, dJ, dH = dev(L), dev(J), dev(H)
    outs = [torch.full(s, 7.0, dtype=dt, device="cuda") if w else None
            for s, w in zip(((n,), (n, 9), (n, 81)), want)]
    p = lambda t: None if t is None else t.data_ptr()
    fn = lib.mf_hip_hencky_mixed if mixed else lib.mf_hip_hencky
    rc = fn(n, p(dL), p(dJ), p(dH), *[p(o) for o in outs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, [None if o is None else o.cpu().double().numpy() for o in outs]


def inputs(fam, n, seed, mixed=False):
    X = general_pos(n, seed) if fam == "general" else spd_pos(n, seed)
    rc, L, J, H = call(fam, "log", 0.0, X, 2, mixed=mixed)
    assert rc == 0
    return L, J, H


@pytest.mark.parametrize("fam", ["general", "matrix"])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 257, 1000])
def test_fp64_matches_oracle(fam, n):
    L, J, H = inputs(fam, n, 10 + n)
    rc, (psi, g, K) = run(L, J, H)
    assert rc == 0
    rpsi, rg, rK = hencky_oracle(L, J, H)
    assert normwise(psi[:, None], rpsi[:, None]).max() <= 1e-13
    assert normwise(g, rg).max() <= 1e-13
    assert normwise(K, rK).max() <= 1e-13


@pytest.mark.parametrize("fam", ["general", "matrix"])
def test_mixed_matches_oracle(fam):
    n = 515
    L, J, H = inputs(fam, n, 3, mixed=True)               # FP32 values (returned as float64 copies)
    rc, (psi, g, K) = run(L, J, H, mixed=True)
    assert rc == 0
    rpsi, rg, rK = hencky_oracle(L, J, H)
    assert normwise(psi[:, None], rpsi[:, None]).max() <= 1e-6
    assert normwise(g, rg).max() <= 1e-6
    assert normwise(K, rK).max() <= 1e-6


def test_optional_outputs_equal_the_full_call():
    n = 100
    L, J, H = inputs("general", n, 5)
    rc, (psi, g, K) = run(L, J, H)
    assert rc == 0
    rc1, (p1, g1, k1) = run(L, None, None, want=(True, False, False))     # psi from L alone
    rc2, (p2, g2, k2) = run(L, J, None, want=(False, True, False))        # g without H
    rc3, (p3, g3, k3) = run(L, J, H, want=(False, False, True))
    assert rc1 == rc2 == rc3 == 0
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)      # (a matrix outside the log's domain: NaN)
    assert same(p1, psi) and same(g2, g) and same(k3, K)
    assert g1 is None and k1 is None and p2 is None and k2 is None and p3 is None and g3 is None


def test_nan_matrix_only_affects_itself():
    n = 70
    L, J, H = inputs("general", n, 6)
    L[33, 4] = np.nan
    rc, (psi, g, K) = run(L, J, H)
    assert rc == 0
    rpsi, rg, rK = hencky_oracle(L, J, H)
    assert np.isnan(psi[33]) and np.isnan(g[33]).all() and np.isnan(K[33]).all()
    keep = np.arange(n) != 33
    assert not np.isnan(K[keep]).any()
    assert normwise(K[keep], rK[keep]).max() <= 1e-13


def test_invalid_arguments():
    L, J, H = inputs("general", 4, 7)
    assert run(L, J, None, want=(False, False, True))[0] == -1              # K without H
    assert run(L, None, None, want=(False, True, False))[0] == -1           # g without J


def test_host_api_matches_c_abi():
    import torch
    n = 64
    L, J, H = inputs("general", n, 8)
    rc, (psi, g, K) = run(L, J, H)
    dL = torch.from_numpy(L.reshape(n, 3, 3)).cuda()
    dJ = torch.from_numpy(J.reshape(n, 3, 3, 3, 3)).cuda()
    dH = torch.from_numpy(H.reshape((n,) + (3,) * 6)).cuda()
    tpsi, tg, tK = matfun_amd.hencky(dL, dJ, dH)
    assert tg.shape == (n, 3, 3) and tK.shape == (n, 3, 3, 3, 3) and tpsi.shape == (n,)
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)
    assert same(tpsi.cpu().numpy(), psi) and same(tg.cpu().numpy().reshape(n, 9), g)
    assert same(tK.cpu().numpy().reshape(n, 81), K)
    npsi, ng, nK = matfun_amd.hencky(L.reshape(n, 3, 3), J.reshape(n, 3, 3, 3, 3))     # NumPy in, no H
    assert isinstance(ng, np.ndarray) and nK is None and same(ng.reshape(n, 9), g)


def test_large_batch_against_torch_fp64():
    import torch
    n = 1 << 20
    rng = torch.Generator(device="cuda").manual_seed(11)
    X = (2.0 * torch.eye(3, dtype=torch.float64, device="cuda") +
         0.6 * torch.randn((n, 3, 3), dtype=torch.float64, device="cuda", generator=rng)).contiguous()
    L, J, H = matfun_amd.gmf("log", X, order=2)
    psi, g, K = matfun_amd.hencky(L, J, H)
    torch.cuda.synchronize()
    Lf, Jm = L.reshape(n, 1, 9), J.reshape(n, 9, 9)
    g_ref = torch.bmm(Lf, Jm).reshape(n, 9)
    K_ref = torch.bmm(Jm.transpose(1, 2).contiguous(), Jm) + torch.bmm(Lf, H.reshape(n, 9, 81)).reshape(n, 9, 9)
    psi_ref = 0.5 * (Lf.reshape(n, 9) ** 2).sum(dim=1)
    ok = ~torch.isnan(K_ref.reshape(n, -1)).any(dim=1)
    assert ok.float().mean().item() > 0.99
    def rel(a, b):
        a, b = a.reshape(n, -1)[ok], b.reshape(n, -1)[ok]
        return ((a - b).norm(dim=1) / b.norm(dim=1)).max().item()
    assert rel(g, g_ref) <= 1e-13 and rel(K, K_ref) <= 1e-13 and rel(psi[:, None], psi_ref[:, None]) <= 1e-13
    Km = K.reshape(n, 9, 9)
    assert rel(Km, Km.transpose(1, 2).contiguous()) <= 1e-12


# ---- the fused form for symmetric X: smf_hip_hencky_matrix against smf_hip_eval_matrix(log) + mf_hip_hencky
def fused(X, want=(True, True, True, True)):
    import torch
    lib = matfun_amd.library()
    n = X.shape[0]
    dX = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64)).cuda()
    outs = [torch.full(s, 7.0, dtype=torch.float64, device="cuda") if w else None
            for s, w in zip(((n, 9), (n,), (n, 9), (n, 81)), want)]
    p = lambda t: None if t is None else t.data_ptr()
    rc = lib.smf_hip_hencky_matrix(n, p(dX), *[p(o) for o in outs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, [None if o is None else o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000])
def test_fused_symmetric_equals_composed(n):
    X = spd_pos(n, 20 + n)
    X[n // 2] = np.array([[2.0, 0.5, 0.0], [0.5, 2.0, 0.0], [0.0, 0.0, 2.0]]).ravel()     # a repeated eigenvalue
    if n > 2:
        X[1] = np.diag([0.3, 0.3, 0.3]).ravel()                                           # a multiple of I
    rc, (L, psi, g, K) = fused(X)
    assert rc == 0
    rc2, L2, J2, H2 = call("matrix", "log", 0.0, X, 2)
    assert rc2 == 0
    assert np.array_equal(L, L2)                           # the same eigen-decomposition and F
    rpsi, rg, rK = hencky_oracle(L2, J2, H2)
    assert normwise(psi[:, None], rpsi[:, None]).max() <= 1e-13
    assert normwise(g, rg).max() <= 1e-13
    assert normwise(K, rK).max() <= 1e-13
    Km = K.reshape(n, 9, 9)
    assert np.array_equal(Km, Km.transpose(0, 2, 1))       # assembled from its distinct values


def test_fused_nan_and_optional_outputs():
    n = 130
    X = spd_pos(n, 31)
    X[70, 1] += 1e-3                                        # not symmetric: NaN for that matrix only
    rc, (L, psi, g, K) = fused(X)
    assert rc == 0
    assert np.isnan(K[70]).all() and np.isnan(psi[70]) and np.isnan(L[70]).all()
    keep = np.arange(n) != 70
    assert not np.isnan(K[keep]).any()
    rc1, (L1, p1, g1, K1) = fused(X, want=(False, False, False, True))
    rc2, (L2, p2, g2, K2) = fused(X, want=(False, True, True, False))
    assert rc1 == rc2 == 0 and L1 is None and L2 is None
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)
    assert same(K1, K) and same(p2, psi) and same(g2, g)
    lib = matfun_amd.library()
    assert lib.smf_hip_hencky_matrix(-1, None, None, None, None, None, None) == -1
    assert lib.smf_hip_hencky_matrix(5, None, None, None, None, None, None) == 0     # nothing asked for
    assert lib.smf_hip_hencky_matrix(0, None, None, None, None, None, None) == 0


def test_fused_large_batch_against_composed():
    import torch
    n = 1 << 20
    from matfun_amd import workloads as W
    X = torch.from_numpy(W.spd(n, seed=5)).cuda()
    L, psi, g, K = matfun_amd.smf_hencky(X.reshape(n, 3, 3))
    L2, J2, H2 = matfun_amd.smf("log", X.reshape(n, 3, 3), order=2)
    psi2, g2, K2 = matfun_amd.hencky(L2, J2, H2)
    torch.cuda.synchronize()
    assert torch.equal(L, L2)

    def rel(a, b):
        a, b = a.reshape(n, -1), b.reshape(n, -1)
        return ((a - b).norm(dim=1) / b.norm(dim=1).clamp_min(1e-300)).max().item()
    assert rel(g, g2) <= 1e-13 and rel(K, K2) <= 1e-13 and rel(psi[:, None], psi2[:, None]) <= 1e-13


# ---- the fused form for general X: gmf_hip_hencky against gmf_hip_eval(log) + mf_hip_hencky
def fused_g(X, want=(True, True, True, True)):
    import torch
    lib = matfun_amd.library()
    n = X.shape[0]
    dX = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64)).cuda()
    outs = [torch.full(s, 7.0, dtype=torch.float64, device="cuda") if w else None
            for s, w in zip(((n, 9), (n,), (n, 9), (n, 81)), want)]
    p = lambda t: None if t is None else t.data_ptr()
    rc = lib.gmf_hip_hencky(n, p(dX), *[p(o) for o in outs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, [None if o is None else o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 1000])
def test_fused_general_equals_composed(n):
    X = general_pos(n, 40 + n)
    X[n // 2] = np.array([[2.0, 1.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0]]).ravel()     # a Jordan block
    if n > 2:
        X[1] = np.array([[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 0.5]]).ravel()    # a complex pair
    rc, (L, psi, g, K) = fused_g(X)
    assert rc == 0
    rc2, L2, J2, H2 = call("general", "log", 0.0, X, 2)
    assert rc2 == 0
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)
    assert same(L, L2)                                     # the same phase 1 and F
    rc3, (p3, g3, K3) = run(L2, J2, H2)                    # the composed GPU path
    rpsi, rg, rK = hencky_oracle(L2, J2, H2)
    for x, r3, r in ((psi[:, None], p3[:, None], rpsi[:, None]), (g, g3, rg), (K, K3, rK)):
        assert normwise(x, r3).max() <= 1e-13 and normwise(x, r).max() <= 1e-13


def test_fused_general_nan_and_optional_outputs():
    n = 130
    X = general_pos(n, 41)
    X[70] = np.diag([1.0, -2.0, 3.0]).ravel()             # a negative real eigenvalue: NaN for that matrix only
    rc, (L, psi, g, K) = fused_g(X)
    assert rc == 0
    assert np.isnan(K[70]).all() and np.isnan(psi[70]) and np.isnan(g[70]).all()
    rc1, (L1, p1, g1, K1) = fused_g(X, want=(False, False, False, True))
    rc2, (L2, p2, g2, K2) = fused_g(X, want=(True, True, True, False))
    assert rc1 == rc2 == 0 and L1 is None and K2 is None
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)
    assert same(K1, K) and same(p2, psi) and same(g2, g) and same(L2, L)
    lib = matfun_amd.library()
    assert lib.gmf_hip_hencky(-1, None, None, None, None, None, None) == -1
    assert lib.gmf_hip_hencky(0, None, None, None, None, None, None) == 0


def test_fused_general_large_batch_against_composed():
    import torch
    n = 1 << 20
    rng = torch.Generator(device="cuda").manual_seed(12)
    X = (2.0 * torch.eye(3, dtype=torch.float64, device="cuda") +
         0.6 * torch.randn((n, 3, 3), dtype=torch.float64, device="cuda", generator=rng)).contiguous()
    L, psi, g, K = matfun_amd.gmf_hencky(X)
    L2, J2, H2 = matfun_amd.gmf("log", X, order=2)
    psi2, g2, K2 = matfun_amd.hencky(L2, J2, H2)
    torch.cuda.synchronize()
    ok = ~torch.isnan(K2.reshape(n, -1)).any(dim=1)
    assert torch.equal(torch.isnan(K.reshape(n, -1)).any(dim=1), ~ok)
    assert torch.equal(L[ok], L2[ok])

    def rel(a, b):
        a, b = a.reshape(n, -1)[ok], b.reshape(n, -1)[ok]
        return ((a - b).norm(dim=1) / b.norm(dim=1).clamp_min(1e-300)).max().item()
    assert rel(g, g2) <= 1e-13 and rel(K, K2) <= 1e-13 and rel(psi[:, None], psi2[:, None]) <= 1e-13


def test_newton_on_the_gradient_converges_quadratically():
    """Newton's method for g(X) = T (g = dpsi/dX of psi = 1/2 |log X|^2, T = g(X*)) with the fused step's tangent K:
    quadratic convergence (reference theory.md:190), i.e. K is the derivative of g as the GPU computes both; with
    the Gauss-Newton tangent J^T J (H dropped) the residual only shrinks by a constant factor per step.  X* = 2 I +
    0.1 N(0, 1) (tangent condition <= 20), start X* + 1e-3 N(0, 1); measured on the 99th percentile of the per-matrix
    relative residual over 4096 matrices (the oracle on the CPU: 2.5e-3, 3.7e-6, 8.4e-11, 7e-16; Gauss-Newton 2.5e-3,
    1.7e-3, 1.2e-3, ...).  (Wider draws, 2 I + 0.3 N(0, 1), include tangents of condition 1e4 whose Newton basin the
    start misses.)"""
    import torch
    n = 4096
    gen = torch.Generator(device="cuda").manual_seed(13)
    eye = torch.eye(3, dtype=torch.float64, device="cuda")
    Xs = (2.0 * eye + 0.1 * torch.randn((n, 3, 3), dtype=torch.float64, device="cuda", generator=gen)).contiguous()
    X0 = (Xs + 1e-3 * torch.randn((n, 3, 3), dtype=torch.float64, device="cuda", generator=gen)).contiguous()
    T = matfun_amd.gmf_hencky(Xs)[2].reshape(n, 9)
    ok = ~torch.isnan(T).any(dim=1)
    assert ok.all()
    T = T[ok]

    def solve(gauss_newton, iters):
        X, res = X0[ok].contiguous(), []
        m = X.shape[0]
        for _ in range(iters):
            if gauss_newton:
                L, J = matfun_amd.gmf("log", X, order=1)
                Jm = J.reshape(m, 9, 9)
                K = torch.bmm(Jm.transpose(1, 2).contiguous(), Jm)
                g = torch.bmm(L.reshape(m, 1, 9), Jm).reshape(m, 9)
            else:
                _, _, g, K = matfun_amd.gmf_hencky(X)
                K, g = K.reshape(m, 9, 9), g.reshape(m, 9)
            r = g - T
            rel = torch.nan_to_num(r.norm(dim=1) / T.norm(dim=1), nan=float("inf"))
            res.append(torch.quantile(rel, 0.99).item())
            dx = torch.linalg.solve_ex(K, r.unsqueeze(-1))[0].squeeze(-1)
            X = (X.reshape(m, 9) - dx).reshape(m, 3, 3).contiguous()
        return res

    res = solve(False, 6)
    assert res[-1] <= 1e-12, res
    for a, b in zip(res, res[1:]):                 # quadratic while the error is above the rounding level
        if a > 1e-9:
            assert b <= 50.0 * a * a + 1e-13, res
    gn = solve(True, 6)
    assert not (gn[-1] <= 1e-9), (gn, res)


@pytest.mark.parametrize("n", [1, 63, 65, 1000])
def test_fused_voigt_equals_matrix_form_and_oracle(n):
    """smf_hip_hencky (Voigt coordinates): log X bitwise the Voigt smf log, psi bitwise the matrix form, g / K exactly
    m_p / m_p m_q times the matrix form's entries (weights 1, 2: exact), and <= 1e-13 against the oracle's Voigt step
    (oracle.pyoracle.hencky_voigt, pinned by differences over v) on the Voigt log of the same matrices"""
    import torch
    from oracle.pyoracle import hencky_voigt
    A = spd_pos(n, 50 + n)
    rep = [0, 4, 8, 1, 2, 5]
    v = np.ascontiguousarray(A[:, rep])
    L, psi, g, K = matfun_amd.smf_hencky(torch.from_numpy(v).cuda())
    L, psi, g, K = (t.cpu().numpy() for t in (L, psi, g, K))
    assert L.shape == (n, 6) and g.shape == (n, 6) and K.shape == (n, 6, 6)
    rc, (Lm, pm, gm, Km) = fused(A)
    assert rc == 0
    mlt = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    assert np.array_equal(psi, pm)
    assert np.array_equal(g, gm[:, rep] * mlt)
    assert np.array_equal(K, Km.reshape(n, 9, 9)[:, rep][:, :, rep] * np.outer(mlt, mlt))
    rc2, F2, J2, H2 = call("voigt", "log", 0.0, v, 2)
    assert rc2 == 0 and np.array_equal(L, F2)
    rpsi, rg, rK = hencky_voigt(F2, J2, H2)
    assert normwise(psi[:, None], rpsi[:, None]).max() <= 1e-13
    assert normwise(g, rg).max() <= 1e-13
    assert normwise(K.reshape(n, 36), rK).max() <= 1e-13


@pytest.mark.parametrize("which", ["smf", "gmf"])
def test_fused_step_replays_in_a_graph(which):
    """the fused entry points allocate nothing and record no memset: captured in a graph, every replay (with eager
    work on the outputs between replays) gives the eager results bitwise"""
    import torch
    lib = matfun_amd.library()
    n = 777
    X = spd_pos(n, 60) if which == "smf" else general_pos(n, 61)
    fn = lib.smf_hip_hencky_matrix if which == "smf" else lib.gmf_hip_hencky
    dX = torch.from_numpy(X).cuda()
    outs = [torch.empty(s, dtype=torch.float64, device="cuda") for s in ((n, 9), (n,), (n, 9), (n, 81))]
    assert fn(n, dX.data_ptr(), *[o.data_ptr() for o in outs], torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph):
            rc = fn(n, dX.data_ptr(), *[o.data_ptr() for o in outs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    assert rc == 0
    for _ in range(3):
        for o in outs:
            o.fill_(7.0)                           # (eager work between the replays)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for o, r in zip(outs, ref):
            assert torch.equal(torch.nan_to_num(o, nan=-1.0), torch.nan_to_num(r, nan=-1.0))


def test_voigt_newton_converges_quadratically():
    """Newton's method in Voigt coordinates for g(v) = T on symmetric positive definite matrices with smf_hip_hencky's
    6 x 6 tangent (the form an FE code assembles): quadratic convergence to the rounding level (the oracle on the
    CPU, same draws: 1.1e-3, 4.2e-6, 1.4e-10, 1.8e-15 at the 99th percentile)"""
    import torch
    n = 4096
    gen = torch.Generator(device="cuda").manual_seed(14)
    M = 0.1 * torch.randn((n, 3, 3), dtype=torch.float64, device="cuda", generator=gen)
    Xs = 2.0 * torch.eye(3, dtype=torch.float64, device="cuda") + 0.5 * (M + M.transpose(1, 2))
    vs = Xs.reshape(n, 9)[:, [0, 4, 8, 1, 2, 5]].contiguous()
    T = matfun_amd.smf_hencky(vs)[2]
    v = (vs + 1e-3 * torch.randn((n, 6), dtype=torch.float64, device="cuda", generator=gen)).contiguous()
    res = []
    for _ in range(5):
        _, _, g, K = matfun_amd.smf_hencky(v)
        r = g - T
        res.append(torch.quantile(r.norm(dim=1) / T.norm(dim=1), 0.99).item())
        v = (v - torch.linalg.solve_ex(K, r.unsqueeze(-1))[0].squeeze(-1)).contiguous()
    assert res[-1] <= 1e-13, res
    for a, b in zip(res, res[1:]):
        if a > 1e-9:
            assert b <= 50.0 * a * a + 1e-13, res
