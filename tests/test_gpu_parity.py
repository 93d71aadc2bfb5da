"""GPU parity of the HIP path (through the C-ABI) against the FP64 oracle.

Bar: FP64 path within 1e-10 of the oracle -- coefficients relative to
max(|b|, se), standard errors element-wise, covariances relative to
max(|V_rc|, se_r se_c) -- and the strict upper triangle of V exactly 0
(oracle.fgls_oracle.parity_errors, which also reports the reference driver's
plain element-wise metric, driver.c:159-177).
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem

pytestmark = pytest.mark.gpu

TOL = 1e-10


def ceil4(x):
    return (x + 3) // 4 * 4


def setup(pkg, prob, wXR=1, max_block=4096, h2=None, sigma=1.0):
    import torch
    h2 = prob.h2 if h2 is None else h2
    L = orc.cholesky_lower(orc.build_M(prob.Phi, h2, sigma))
    tr = orc.trait_precompute(L, prob.XL, prob.y)
    ctx = pkg.DeviceContext(prob.n, prob.wXL, wXR, max_block)
    ctx.set_trait(torch.from_numpy(np.ascontiguousarray(L.T)).cuda(),
                  pkg.colmajor(prob.XL) if prob.wXL else None,
                  torch.from_numpy(prob.y).cuda())
    return tr, ctx


def v_flat(V):
    """oracle V (count x p x p, [r][c]) -> on-disk order rows (count x p*p)."""
    return V.transpose(0, 2, 1).reshape(V.shape[0], -1)


def check_bv(B, V, b_ref, V_ref, p, tol=TOL):
    """Parity bar: b relative to max(|b|, se), V's standard errors
    element-wise, covariances relative to max(|V_rc|, se_r se_c); the strict
    upper triangle of V exactly 0 (oracle.parity_errors)."""
    e = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), b_ref, v_flat(V_ref), p)
    assert e["V_upper_zero"], "strict upper triangle of V must be exactly 0"
    assert e["B_scaled"] < tol and e["V_scaled"] < tol, e
    return e


@pytest.mark.parametrize("n,wXL", [(64, 3), (500, 3), (1000, 3), (333, 2), (1031, 5)])
def test_trait_precompute(pkg, gpu, n, wXL):
    prob = make_problem(n=n, wXL=wXL, m=4, seed=n)
    tr, ctx = setup(pkg, prob)
    Bt, Vtl, res = ctx.trait()
    assert orc.max_rel_err(Bt, tr.B_t) < 1e-11
    assert orc.max_rel_err(np.tril(Vtl), tr.V_tl) < 1e-11
    assert np.all(np.triu(Vtl, 1) == 0.0)
    assert abs(res - tr.res_sigma) / tr.res_sigma < 1e-11
    ctx.close()


@pytest.mark.parametrize("n,k", [(500, 200), (1000, 64), (333, 70), (1031, 129)])
def test_trsm_matches_lapack(pkg, gpu, n, k):
    """fgls_hip_trsm = dtrsm(L, L, N, N) in place (fgls_chol_gpu.c:550)."""
    prob = make_problem(n=n, wXL=3, m=k, seed=2 * n + k)
    tr, ctx = setup(pkg, prob)
    X = pkg.colmajor(prob.XR, ld=ceil4(n))
    ctx.trsm(X)
    Z = X.cpu().numpy()[:, :n].T
    Zr = sla.solve_triangular(tr.L, prob.XR, lower=True)
    err = np.abs(Z - Zr).max() / np.abs(Zr).max()
    assert err < 1e-13, err
    # padding rows of the device buffer are untouched (zero)
    assert np.all(X.cpu().numpy()[:, n:] == 0.0)
    ctx.close()


@pytest.mark.parametrize("n,k", [(500, 200), (1000, 300), (333, 70), (1031, 129), (40, 5)])
def test_trmm_matches_lapack(pkg, gpu, n, k):
    """Z = L^-1 X through the device inverse equals dtrsm's result."""
    prob = make_problem(n=n, wXL=3, m=k, seed=3 * n + k)
    tr, ctx = setup(pkg, prob)
    X = pkg.colmajor(prob.XR, ld=ceil4(n))
    Z = ctx.trmm(X).cpu().numpy()
    Zr = sla.solve_triangular(tr.L, prob.XR, lower=True)
    err = np.abs(Z[:, :n].T - Zr).max() / np.abs(Zr).max()
    assert err < 1e-13, err
    assert np.all(Z[:, n:] == 0.0)
    ctx.close()


@pytest.mark.parametrize("n,m,max_block", [(1000, 2000, 2048), (500, 129, 512), (333, 1, 64),
                                           (1031, 300, 1024), (128, 256, 256)])
def test_solve_block_matches_oracle(pkg, gpu, n, m, max_block):
    prob = make_problem(n=n, wXL=3, m=m, seed=n + m)
    tr, ctx = setup(pkg, prob, max_block=max_block)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    B, V = ctx.solve_block(pkg.colmajor(prob.XR, ld=ceil4(n)), snp_base=5)
    ctx.check()
    check_bv(B, V, b_ref, V_ref, 4)
    ctx.close()


@pytest.mark.parametrize("wXL", [0, 1, 5, 7])
def test_solve_block_other_p(pkg, gpu, wXL):
    prob = make_problem(n=700, wXL=wXL, m=300, seed=11 + wXL, dosage="imputed", h2=0.3)
    tr, ctx = setup(pkg, prob, max_block=512)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    B, V = ctx.solve_block(pkg.colmajor(prob.XR))
    ctx.check()
    check_bv(B, V, b_ref, V_ref, wXL + 1)
    ctx.close()


@pytest.mark.parametrize("wXL,wXR", [(3, 2), (2, 3)])
def test_generic_wxr(pkg, gpu, wXL, wXR):
    """wXR > 1: TRMM to HBM then the wavefront per-SNP kernel."""
    m = 150
    prob = make_problem(n=600, wXL=wXL, m=m, seed=21 + wXR, wXR=wXR)
    tr, ctx = setup(pkg, prob, wXR=wXR, max_block=256)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, wXR)
    B, V = ctx.solve_block(pkg.colmajor(prob.XR), count=m)
    ctx.check()
    check_bv(B, V, b_ref, V_ref, wXL + wXR)
    ctx.close()


def test_trsm_then_snp_stage(pkg, gpu):
    """The reference's own structure: in-place TRSM of the block
    (fgls_chol_gpu.c:550), then the per-SNP stage (:619-687) on the result."""
    prob = make_problem(n=900, wXL=3, m=333, seed=5)
    tr, ctx = setup(pkg, prob, max_block=512)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    X = pkg.colmajor(prob.XR)
    ctx.trsm(X)
    B, V = ctx.snp_stage(X)
    ctx.check()
    check_bv(B, V, b_ref, V_ref, 4)
    ctx.close()


def test_reference_fixture_test_xl(pkg, gpu):
    """test_cases/test_XL/B_exp.out (t=20) through the device path, trait by trait."""
    import torch
    from tests.golden.test_xl_fixture import synth_xr
    prm, Phi, H, Sig, XL, XR, Y, Bexp = synth_xr()
    n, p, m, t, wXL = prm["n"], prm["p"], prm["m"], prm["t"], prm["wXL"]
    ctx = pkg.DeviceContext(n, wXL, 1, max_block=m)
    Xd = pkg.colmajor(XR)
    worst = 0.0
    for j in range(t):
        L = orc.cholesky_lower(orc.build_M(Phi, H[j], Sig[j]))
        ctx.set_trait(torch.from_numpy(np.ascontiguousarray(L.T)).cuda(), pkg.colmajor(XL),
                      torch.from_numpy(np.ascontiguousarray(Y[:, j])).cuda())
        B, V = ctx.solve_block(Xd)
        ctx.check()
        worst = max(worst, orc.max_rel_err(B.cpu().numpy(), Bexp[j]))
    ctx.close()
    assert worst < 1e-10, worst


def test_singular_snp_reports_index(pkg, gpu):
    """An all-zero SNP column makes S_i singular (zero last pivot): dpotrf(V)
    fails (fgls_chol.c:420-426); the device records the first such SNP."""
    prob = make_problem(n=300, wXL=3, m=100, seed=9)
    prob.XR[:, 37] = 0.0
    prob.XR[:, 80] = 0.0
    tr, ctx = setup(pkg, prob, max_block=128)
    B, V = ctx.solve_block(pkg.colmajor(prob.XR), snp_base=1000)
    with pytest.raises(pkg.FglsError) as ei:
        ctx.check()
    assert ei.value.code == pkg.FGLS_ERR_NOTSPD
    assert ei.value.bad_snp == 1037
    assert ei.value.info == 4
    ctx.check()  # cleared
    ctx.close()


def test_sigma_invariance_and_scaling(pkg, gpu):
    """b and V do not depend on sigma; scaling a SNP column by c scales its
    coefficient by 1/c and its standard error by 1/|c| (closed form).  The two
    sigmas give different factors, so b and V agree to FP64 rounding of the
    whole solve (element-wise relative, the driver's metric), not bitwise."""
    prob = make_problem(n=400, wXL=3, m=64, seed=13)
    _, ctx1 = setup(pkg, prob, sigma=1.0, max_block=64)
    _, ctx2 = setup(pkg, prob, sigma=3.7, max_block=64)
    X = pkg.colmajor(prob.XR)
    B1, V1 = ctx1.solve_block(X)
    B2, V2 = ctx2.solve_block(X)
    assert orc.max_rel_err(B2.cpu().numpy(), B1.cpu().numpy()) < 1e-11
    assert orc.max_rel_err(V2.cpu().numpy(), V1.cpu().numpy(), p=4, lower_only=True) < 1e-11
    B3, V3 = ctx1.solve_block(pkg.colmajor(prob.XR * -2.5))
    assert orc.max_rel_err(B3.cpu().numpy()[:, 3] * -2.5, B1.cpu().numpy()[:, 3]) < 1e-12
    assert orc.max_rel_err(V3.cpu().numpy()[:, 15] * 2.5, V1.cpu().numpy()[:, 15]) < 1e-12
    ctx1.close()
    ctx2.close()
