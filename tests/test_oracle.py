"""Pins of the FP64 oracle (CPU only).

Each function of oracle/fgls_oracle.py is checked against something that does
not restate its formula:
  * the reference fixture test_cases/test_XL/B_exp.out (B, 20 traits);
  * the reference's own fgls_chol compiled from its sources
    (oracle/_ref/libfgls_ref.so, when built) -- B and V;
  * closed forms with an explicit inverse of M (brute force, small n);
  * invariances: sigma scale, SNP permutation and scaling, intercept shift;
  * planted mistakes (in the oracle's inputs/conventions) that the pins catch.
"""
import os

import numpy as np
import pytest

from oracle import fgls_oracle as orc
from oracle import ref_runner
from tests.fgls_data import make_problem, write_problem_dir
from tests.golden.test_xl_fixture import load_fixture, synth_xr, write_dir


@pytest.fixture(scope="module")
def txl():
    return synth_xr()


def test_oracle_reproduces_reference_fixture(txl):
    prm, Phi, H, Sig, XL, XR, Y, Bexp = txl
    B, V = orc.fgls_chol(Phi, H, Sig, XL, XR, Y, wXR=1, x_b=prm["x_b"])
    assert orc.max_rel_err(B, Bexp.ravel()) < 1e-12


def test_fixture_pins_conventions(txl):
    """Plausible convention mistakes fail the B_exp.out pin: h used as h^2,
    row-major X_L, a perturbed M."""
    prm, Phi, H, Sig, XL, XR, Y, Bexp = txl
    t = prm["t"]
    sub = slice(0, 200)
    ref = Bexp[:2, sub].ravel()

    def run(Phi_, H_, XL_, Y_):
        B, _ = orc.fgls_chol(Phi_, H_[:2], Sig[:2], XL_, XR[:, sub], Y_[:, :2], wXR=1)
        return orc.max_rel_err(B, ref)

    assert run(Phi, H, XL, Y) < 1e-12
    assert run(Phi, H ** 2, XL, Y) > 1e-6                                   # h^2 instead of h
    n = prm["n"]
    XLrow = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "test_XL", "XL.in")).reshape(n, 3)
    assert run(Phi, H, XLrow, Y) > 1e-3                                      # row-major X_L
    Phi_t = Phi + 0.01 * np.eye(n)
    assert run(Phi_t, H, XL, Y) > 1e-8                                       # perturbed M


def brute_force(Phi, h, sigma, XL, x, y):
    """b = (X'M^-1X)^-1 X'M^-1 y and V = res_sigma (X'M^-1X)^-1 with an
    explicit inverse of M; res_sigma from the covariate-only GLS fit."""
    n = Phi.shape[0]
    M = sigma * (h * Phi + (1 - h) * np.eye(n))
    Mi = np.linalg.inv(M)
    X = np.column_stack([XL, x])
    S = X.T @ Mi @ X
    b = np.linalg.solve(S, X.T @ Mi @ y)
    SL = XL.T @ Mi @ XL
    bL = np.linalg.solve(SL, XL.T @ Mi @ y)
    r = y - XL @ bL
    res = (r @ Mi @ r) / (n - XL.shape[1])
    V = res * np.linalg.inv(S)
    return b, V


@pytest.mark.parametrize("n,wXL", [(40, 3), (64, 1), (50, 5)])
def test_oracle_matches_brute_force(n, wXL):
    prob = make_problem(n=n, wXL=wXL, m=6, seed=100 + n, h2=0.55)
    sigma = 1.7
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, sigma))
    tr = orc.trait_precompute(L, prob.XL, prob.y)
    b, V = orc.solve_snps(prob.XR, tr, 1)
    for i in range(prob.m):
        bb, VV = brute_force(prob.Phi, prob.h2, sigma, prob.XL, prob.XR[:, i], prob.y)
        assert np.allclose(b[i], bb, rtol=1e-10, atol=0)
        low = np.tril(VV, -1)
        assert np.allclose(np.tril(V[i], -1), low, rtol=1e-9, atol=1e-15)
        assert np.allclose(np.diag(V[i]), np.sqrt(np.diag(VV)), rtol=1e-10)
        assert np.all(np.triu(V[i], 1) == 0.0)


def test_oracle_invariances():
    prob = make_problem(n=120, wXL=3, m=30, seed=4, h2=0.35)
    L1 = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    L2 = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 4.2))
    t1 = orc.trait_precompute(L1, prob.XL, prob.y)
    t2 = orc.trait_precompute(L2, prob.XL, prob.y)
    b1, V1 = orc.solve_snps(prob.XR, t1, 1)
    b2, V2 = orc.solve_snps(prob.XR, t2, 1)
    assert orc.max_rel_err(b2, b1) < 1e-12                 # sigma scale
    assert orc.max_rel_err(V2, V1) < 1e-12
    perm = np.random.default_rng(0).permutation(prob.m)
    b3, V3 = orc.solve_snps(prob.XR[:, perm], t1, 1)      # SNP permutation
    assert np.array_equal(b3, b1[perm])
    b4, V4 = orc.solve_snps(prob.XR * -3.0, t1, 1)        # SNP scaling
    assert orc.max_rel_err(b4[:, 3] * -3.0, b1[:, 3]) < 1e-12
    assert orc.max_rel_err(V4[:, 3, 3] * 3.0, V1[:, 3, 3]) < 1e-12
    b5, _ = orc.solve_snps(prob.XR + 0.75, t1, 1)         # intercept shift (column 0 is 1)
    assert orc.max_rel_err(b5[:, 3], b1[:, 3]) < 1e-10


@pytest.mark.skipif(not ref_runner.available(), reason="reference sources were not built (oracle/Makefile)")
@pytest.mark.parametrize("n,wXL,m,t,x_b,wXR", [(300, 3, 700, 2, 128, 1), (150, 2, 90, 1, 40, 2)])
def test_oracle_matches_reference_build(tmp_path, n, wXL, m, t, x_b, wXR):
    """B and V of the oracle against the reference's own fgls_chol."""
    prob = make_problem(n=n, wXL=wXL, m=m, seed=n * t, t=t, wXR=wXR)
    H = np.linspace(0.25, 0.65, t)
    Sig = np.linspace(0.9, 1.3, t)
    d = str(tmp_path)
    write_problem_dir(prob, d, H=H, Sig=Sig)
    p = wXL + wXR
    ref_runner.fgls_chol_dir(d, n, p, m, t, wXL, x_b, threads=2)
    B = np.fromfile(os.path.join(d, "B.out"))
    V = np.fromfile(os.path.join(d, "V.out"))
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=wXR, x_b=x_b)
    assert B.size == Bo.size and V.size == Vo.size
    e = orc.parity_errors(Bo, Vo, B, V, p)
    assert e["B_scaled"] < 1e-12 and e["V_scaled"] < 1e-12, e
    assert orc.parity_errors(B, V, B, V, p)["V_upper_zero"]


@pytest.mark.skipif(not ref_runner.available(), reason="reference sources were not built (oracle/Makefile)")
def test_reference_build_reproduces_fixture(tmp_path):
    d = str(tmp_path / "txl")
    prm = write_dir(d)
    ref_runner.fgls_chol_dir(d, prm["n"], prm["p"], prm["m"], prm["t"], prm["wXL"], prm["x_b"], threads=2)
    B = np.fromfile(os.path.join(d, "B.out"))
    assert orc.max_rel_err(B, np.fromfile(os.path.join(d, "B_exp.out"))) < 1e-12


def test_eigen_oracle_reproduces_reference_fixture(txl):
    """The eigenbasis restatement (fgls_eigen.c) against test_XL/B_exp.out."""
    prm, Phi, H, Sig, XL, XR, Y, Bexp = txl
    sub = slice(0, 400)
    B, _ = orc.fgls_eigen(Phi, H[:4], Sig[:4], XL, XR[:, sub], Y[:, :4], wXR=1)
    assert orc.max_rel_err(B, Bexp[:4, sub].ravel()) < 1e-10


@pytest.mark.parametrize("n,wXL,m,t", [(200, 3, 300, 3), (97, 0, 50, 2), (150, 6, 80, 1)])
def test_eigen_oracle_matches_cholesky_oracle(n, wXL, m, t):
    """M^-1 = Z diag(1/(a W + b)) Z' = L^-T L^-1: both restatements give the same
    b and V (two different factorisations, so agreement is not by construction)."""
    prob = make_problem(n=n, wXL=wXL, m=m, seed=n + t, t=t)
    H = np.linspace(0.15, 0.85, t)
    Sig = np.linspace(0.5, 2.0, t)
    Be, Ve = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y)
    Bc, Vc = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y)
    e = orc.parity_errors(Be, Ve, Bc, Vc, wXL + 1)
    assert e["B_scaled"] < 1e-11 and e["V_scaled"] < 1e-11 and e["V_upper_zero"], e


def test_eigen_oracle_catches_planted_mistakes():
    """Plausible mistakes in the eigenbasis weights fail the Cholesky pin:
    a and b swapped, Winv without the square root, Z in place of Z'."""
    prob = make_problem(n=120, wXL=3, m=40, seed=4, t=1, h2=0.3)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, 0.3, 1.2))
    bc, Vc = orc.solve_snps(prob.XR, orc.trait_precompute(L, prob.XL, prob.y), 1)
    Z, W = orc.eigen_decompose(prob.Phi)

    def err(Zm, Wm, h, sigma, sqrt=True):
        tr = orc.eigen_trait(Zm, Wm, prob.XL, prob.y, h, sigma)
        if not sqrt:
            tr.Winv = tr.Winv ** 2
        b, V = orc.eigen_solve_snps(Zm, tr, prob.XR, 1)
        return orc.parity_errors(b, V.transpose(0, 2, 1).reshape(len(b), -1), bc,
                                 Vc.transpose(0, 2, 1).reshape(len(bc), -1), 4)["B_scaled"]

    assert err(Z, W, 0.3, 1.2) < 1e-11
    assert err(Z, W, 0.7, 1.2) > 1e-4                 # a and b swapped (h used as 1 - h)
    assert err(Z, W, 0.3, 1.2, sqrt=False) > 1e-4     # Winv^2 in place of Winv
    assert err(Z.T.copy(), W, 0.3, 1.2) > 1e-4        # rows of Z in place of columns


@pytest.mark.skipif(not ref_runner.available(), reason="reference sources were not built (oracle/Makefile)")
@pytest.mark.parametrize("wXR", [1, 2])
def test_eigen_oracle_matches_reference_build(tmp_path, wXR):
    """B and V of the eigenbasis restatement against the reference's own
    fgls_eigen (3 traits with their own h2 and sigma2; one and two columns
    per SNP).  The reference's
    asynchronous B/V writes (fgls_eigen.c:574-600, :646-650) now and then
    leave a wrong V behind (about one run in six here, with one compute
    thread too) and, more rarely, unwritten B entries: B and V must both
    match on one of up to five runs."""
    n, wXL, m, t, x_b = 300, 3, 500, 3, 128
    prob = make_problem(n=n, wXL=wXL, m=m, seed=3, t=t, wXR=wXR)
    H = np.array([0.4, 0.7, 0.2])
    Sig = np.array([1.0, 2.0, 0.5])
    d = str(tmp_path)
    write_problem_dir(prob, d, H=H, Sig=Sig)
    Bo, Vo = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=wXR, x_b=x_b)
    errors = []
    for attempt in range(5):
        for f in ("B.out", "V.out"):
            if os.path.exists(os.path.join(d, f)):
                os.remove(os.path.join(d, f))
        ref_runner.fgls_eigen_dir(d, n, wXL + wXR, m, t, wXL, x_b, threads=1)
        B = np.fromfile(os.path.join(d, "B.out"))
        V = np.fromfile(os.path.join(d, "V.out"))
        with np.errstate(divide="ignore", invalid="ignore"):
            e = orc.parity_errors(Bo, Vo, B, V, wXL + wXR)
        errors.append(e)
        if e["B_scaled"] < 1e-12 and e["V_scaled"] < 1e-12 and e["V_upper_zero"]:
            break
    else:
        raise AssertionError(errors)


def test_fixture_files_match_driver_stdin():
    prm, Phi, H, Sig, XL, Y, Bexp = load_fixture()
    assert (prm["n"], prm["p"], prm["m"], prm["t"], prm["wXL"]) == (500, 4, 2000, 20, 3)
    assert np.allclose(Phi, Phi.T)
    assert Bexp.shape == (20, 2000, 4)


def test_max_rel_err_semantics():
    """driver.c:159-177: 0/0 skipped, x/0 reported."""
    assert orc.max_rel_err([0.0, 1.0], [0.0, 1.0]) == 0.0
    assert orc.max_rel_err([1e-3, 1.0], [0.0, 1.0]) == np.inf
    v = np.zeros(4)
    v[2] = 5.0          # flat index 2 = (row 0, col 1): strict upper of a 2x2 block
    assert orc.max_rel_err(v, np.zeros(4), p=2, lower_only=True) == 0.0


def test_parity_metric_catches_planted_errors():
    """The scaled metric flags a wrong standard error, a wrong covariance, a
    transposed V block and a non-zero upper triangle; the element-wise metric
    alone is blind to absolute errors on tiny entries."""
    prob = make_problem(n=80, wXL=3, m=20, seed=8)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    tr = orc.trait_precompute(L, prob.XL, prob.y)
    b, V = orc.solve_snps(prob.XR, tr, 1)
    Vf = V.transpose(0, 2, 1).reshape(prob.m, -1)
    assert orc.parity_errors(b, Vf, b, Vf, 4)["V_scaled"] == 0.0
    bad = Vf.copy(); bad[3, 5] *= 1 + 1e-6                     # a standard error (c=1, r=1)
    assert orc.parity_errors(b, bad, b, Vf, 4)["V_scaled"] > 1e-7
    bad = Vf.copy(); bad[3, 1] += 1e-6 * Vf[3, 0]               # a covariance (c=0, r=1)
    assert orc.parity_errors(b, bad, b, Vf, 4)["V_scaled"] > 1e-9
    bad = V.reshape(prob.m, -1).copy()                          # row-major instead of column-major
    assert orc.parity_errors(b, bad, b, Vf, 4)["V_scaled"] > 1e-3
    assert not orc.parity_errors(b, bad, b, Vf, 4)["V_upper_zero"]
    bb = b.copy(); bb[7, 3] += 1e-6 * Vf[7, 15]                # SNP coefficient off by 1e-6 se
    assert orc.parity_errors(bb, Vf, b, Vf, 4)["B_scaled"] > 1e-7
    bb = b.copy(); bb[7, 0] *= 1 + 1e-6                         # the intercept of one SNP
    e = orc.parity_errors(bb, Vf, b, Vf, 4)
    assert 1e-7 < e["B_norm"] < 1e-6 and e["V_norm"] == 0.0
    bad = Vf.copy(); bad[2, 0] *= 1 + 1e-6                      # V_00 of one SNP
    assert orc.parity_errors(b, bad, b, Vf, 4)["V_norm"] > 1e-8


# --- adjoint of l = sum_i w_i b_i[p-1]^2 (no reference counterpart) ---------

def _adj_problem(n=48, wXL=3, m=40, seed=7, h2=0.45):
    pr = make_problem(n=n, wXL=wXL, m=m, seed=seed, h2=h2)
    w = np.random.default_rng(seed + 1).uniform(0.0, 1.0, m)
    return pr, w


@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_adjoint_oracle_matches_torch_autograd(sigma):
    from tests.adjoint_ref import dense_grads
    pr, w = _adj_problem()
    loss, gy, gh = orc.snp_loss_adjoint(pr.Phi, pr.h2, sigma, pr.XL, pr.XR, pr.y, w)
    loss_r, gy_r, gh_r = dense_grads(pr.y, pr.h2, pr.Phi, pr.XL, pr.XR, w, sigma)
    assert abs(loss - loss_r) <= 1e-10 * abs(loss_r)
    assert np.max(np.abs(gy - gy_r)) <= 1e-9 * np.max(np.abs(gy_r))
    assert abs(gh - gh_r) <= 1e-9 * abs(gh_r)


def test_adjoint_oracle_matches_central_differences():
    """Step 1e-6 central differences of the oracle's forward loss."""
    pr, w = _adj_problem(n=40, m=30, seed=11)
    sigma = 1.0

    def loss_of(y, h):
        L = orc.cholesky_lower(orc.build_M(pr.Phi, h, sigma))
        tr = orc.trait_precompute(L, pr.XL, y)
        b, _ = orc.solve_snps(pr.XR, tr, 1)
        return float(w @ (b[:, -1] ** 2))

    loss, gy, gh = orc.snp_loss_adjoint(pr.Phi, pr.h2, sigma, pr.XL, pr.XR, pr.y, w)
    eps = 1e-6
    fd_h = (loss_of(pr.y, pr.h2 + eps) - loss_of(pr.y, pr.h2 - eps)) / (2 * eps)
    assert abs(fd_h - gh) <= 1e-6 * abs(gh)
    d = np.random.default_rng(3).standard_normal(pr.n)
    fd_y = (loss_of(pr.y + eps * d, pr.h2) - loss_of(pr.y - eps * d, pr.h2)) / (2 * eps)
    assert abs(fd_y - gy @ d) <= 1e-6 * abs(gy @ d)


def test_adjoint_oracle_catches_planted_mistakes():
    """The tolerances above separate the slips a restatement could make: a
    dropped factor 2 of the square or a flipped sign of dM/dh."""
    from tests.adjoint_ref import dense_grads
    pr, w = _adj_problem(seed=5)
    _, gy_r, gh_r = dense_grads(pr.y, pr.h2, pr.Phi, pr.XL, pr.XR, w)
    _, gy, gh = orc.snp_loss_adjoint(pr.Phi, pr.h2, 1.0, pr.XL, pr.XR, pr.y, w)
    tol_y = 1e-9 * np.max(np.abs(gy_r))
    assert np.max(np.abs(gy - gy_r)) <= tol_y and abs(gh - gh_r) <= 1e-9 * abs(gh_r)
    assert np.max(np.abs(0.5 * gy - gy_r)) > 1e6 * tol_y
    assert abs(-gh - gh_r) > 1e6 * 1e-9 * abs(gh_r)


def test_error_stats_by_hand():
    """error_stats on four hand-checked entries (p = 1): relative errors of b
    are 0, 0.2/2.2, 0.9 and 0 (0/0); only the first two have |b| >= 1e-3 se;
    var = se^2 gives 0, 3, 0, 0."""
    B, Bref = np.array([1.0, 2.0, 1e-6, 0.0]), np.array([1.0, 2.2, 1e-5, 0.0])
    V, Vref = np.array([1.0, 2.0, 1.0, 3.0]), np.array([1.0, 1.0, 1.0, 3.0])
    s = orc.error_stats(B, V, Bref, Vref, 1)
    assert s["entries"] == 4
    assert s["b"]["max"] == pytest.approx(0.9) and s["b"]["median"] == pytest.approx(0.1 / 2.2)
    assert s["b"]["max_if_b_ge_1e-3_se"] == pytest.approx(0.2 / 2.2)
    assert s["b"]["share_b_ge_1e-3_se"] == 0.5
    assert s["var"]["max"] == pytest.approx(3.0) and s["var"]["median"] == 0.0
    # p > 1: only the diagonal of V enters var, in the on-disk order c*p + r
    V2 = np.array([[4.0, 0.5, 0.0, 9.0]])
    s2 = orc.error_stats(np.ones((1, 2)), V2 * [[1.0, 1.0, 1.0, 2.0]], np.ones((1, 2)), V2, 2)
    assert s2["var"]["max"] == pytest.approx(3.0) and s2["b"]["max"] == 0.0
