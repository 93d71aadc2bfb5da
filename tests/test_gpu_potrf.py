"""Device Cholesky of M = sigma2 (h2 Phi + (1 - h2) I) (fgls_hip_set_trait_phi),
the replacement of the host M construction and dpotrf (fgls_chol_gpu.c:406-430).
Checked against LAPACK's factor (scipy dpotrf through the oracle) and, end to
end, against the oracle's b and V."""
import numpy as np
import pytest
import torch

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem, make_phi

pytestmark = pytest.mark.gpu


def _colmajor(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()


@pytest.mark.parametrize("n,h2,sigma2", [(64, 0.5, 1.0), (100, 0.3, 2.0), (257, 0.8, 0.7),
                                          (1000, 0.5, 1.0), (2049, 0.45, 1.3)])
def test_device_factor_matches_lapack(gpu, pkg, n, h2, sigma2):
    print(f"device potrf n={n} h2={h2} sigma2={sigma2}", flush=True)
    rng = np.random.default_rng(n)
    Phi = make_phi(rng, n)
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, 2))])
    y = rng.standard_normal(n)
    with pkg.DeviceContext(n, 3, 1, max_block=64) as dc:
        dc.set_trait_phi(_colmajor(Phi), h2, sigma2, _colmajor(XL), torch.from_numpy(y).cuda())
        L = np.tril(dc.factor().cpu().numpy().T)
    Lref = orc.cholesky_lower(orc.build_M(Phi, h2, sigma2))
    err = np.linalg.norm(L - Lref) / np.linalg.norm(Lref)
    assert err < 1e-13, err
    M = orc.build_M(Phi, h2, sigma2)
    assert np.linalg.norm(L @ L.T - M) / np.linalg.norm(M) < 1e-14


def test_set_trait_phi_solve_matches_oracle(gpu, pkg):
    """b and V through the device factor vs the oracle (scaled 1e-10)."""
    n, m = 777, 900
    prob = make_problem(n=n, wXL=3, m=m, seed=5, h2=0.55)
    ldx = (n + 3) // 4 * 4
    X = torch.zeros((m, ldx), dtype=torch.float64)
    X[:, :n] = torch.from_numpy(np.ascontiguousarray(prob.XR.T))
    with pkg.DeviceContext(n, 3, 1, max_block=m) as dc:
        dc.set_trait_phi(_colmajor(prob.Phi), prob.h2, 1.4, _colmajor(prob.XL),
                         torch.from_numpy(prob.y.copy()).cuda())
        B, V = dc.solve_block(X.cuda(), m)
        dc.check()
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.4))
    tr = orc.trait_precompute(L, prob.XL, prob.y)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    e = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), b_ref,
                          V_ref.transpose(0, 2, 1).reshape(m, -1), 4)
    assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10 and e["V_upper_zero"], e


@pytest.mark.parametrize("bad_col", [0, 100, 130])
def test_device_factor_reports_first_bad_pivot(gpu, pkg, bad_col):
    """A non-SPD M: the error carries LAPACK's info (order of the first
    non-positive leading minor), like the host dpotrf (fgls_chol_gpu.c:416-430)."""
    n = 200
    Phi = np.eye(n)
    Phi[bad_col, bad_col] = -5.0
    h2 = 0.9
    M = orc.build_M(Phi, h2, 1.0)
    import scipy.linalg as sla
    _, info = sla.lapack.dpotrf(M, lower=1)
    assert info == bad_col + 1
    XL = np.ones((n, 1))
    with pkg.DeviceContext(n, 1, 1, max_block=16) as dc:
        with pytest.raises(pkg.FglsError) as ei:
            dc.set_trait_phi(_colmajor(Phi), h2, 1.0, _colmajor(XL), torch.ones(n, dtype=torch.float64).cuda())
        assert ei.value.code == pkg.FGLS_ERR_NOTSPD
        assert f"dpotrf(M) failed (info: {info})" in ei.value.msg
        # the context stays usable with an SPD trait
        dc.set_trait_phi(_colmajor(np.eye(n)), 0.5, 1.0, _colmajor(XL), torch.ones(n, dtype=torch.float64).cuda())


def _inverse_through_trmm(dc, n):
    """L^-1 as the device product L^-1 I (fgls_hip_trmm), returned (n, n)."""
    ldx = (n + 3) // 4 * 4
    eye = torch.zeros((n, ldx), dtype=torch.float64)
    eye[:, :n] = torch.eye(n, dtype=torch.float64)
    Z = dc.trmm(eye.cuda(), n)
    return Z[:, :n].cpu().numpy().T   # Z[c, r] = L^-1(r, c)


@pytest.mark.parametrize("n", [5, 63, 64, 100, 257, 700, 1000, 2049, 8400])
def test_device_inverse_of_factor(gpu, pkg, n):
    """L^-1 formed with the device factor (recursive inversion): lower
    triangular (zeros above the diagonal, exactly) and equal to the inverse of
    LAPACK's factor; also L^-1 M L^-T = I.  n = 8400 has top-level products
    with more than 4,096 64 x 64 output blocks (the tiled product; smaller
    ones split K)."""
    print(f"device factor + inverse n={n}", flush=True)
    rng = np.random.default_rng(n + 11)
    Phi = make_phi(rng, n)
    h2, sigma2 = 0.6, 1.7
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, 2))])
    with pkg.DeviceContext(n, 3, 1, max_block=64) as dc:
        dc.set_trait_phi(_colmajor(Phi), h2, sigma2, _colmajor(XL),
                         torch.from_numpy(rng.standard_normal(n)).cuda())
        T = _inverse_through_trmm(dc, n)
    M = orc.build_M(Phi, h2, sigma2)
    Tref = np.linalg.inv(orc.cholesky_lower(M))
    assert np.all(np.triu(T, 1) == 0.0)
    assert np.linalg.norm(T - Tref) / np.linalg.norm(Tref) < 1e-13
    assert np.abs(T @ M @ T.T - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("n", [64, 129, 1000, 1500, 8400])
def test_inverse_of_given_factor(gpu, pkg, n):
    """fgls_hip_set_trait with the caller's factor: L^-1 by the same recursion
    without the factorisation, whatever the factor's strict upper triangle
    holds."""
    print(f"inverse of a given factor n={n}", flush=True)
    rng = np.random.default_rng(n + 3)
    M = orc.build_M(make_phi(rng, n), 0.4, 1.0)
    L = orc.cholesky_lower(M)
    Lg = L + np.triu(rng.standard_normal((n, n)), 1)     # unspecified upper part
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, 2))])
    with pkg.DeviceContext(n, 3, 1, max_block=64) as dc:
        dc.set_trait(_colmajor(Lg), _colmajor(XL), torch.from_numpy(rng.standard_normal(n)).cuda())
        T = _inverse_through_trmm(dc, n)
    Tref = np.linalg.inv(L)
    assert np.all(np.triu(T, 1) == 0.0)
    assert np.linalg.norm(T - Tref) / np.linalg.norm(Tref) < 1e-13


@pytest.mark.parametrize("n", [200, 1100])
def test_device_factor_reads_lower_triangle_of_phi(gpu, pkg, n):
    """set_trait_phi reads only the lower triangle of Phi (NaN above it)."""
    rng = np.random.default_rng(n + 5)
    Phi = make_phi(rng, n)
    Phl = Phi.copy()
    Phl[np.triu_indices(n, 1)] = np.nan
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, 2))])
    y = rng.standard_normal(n)
    with pkg.DeviceContext(n, 3, 1, max_block=64) as dc:
        dc.set_trait_phi(_colmajor(Phl), 0.45, 1.2, _colmajor(XL), torch.from_numpy(y).cuda())
        L = np.tril(dc.factor().cpu().numpy().T)
        T = _inverse_through_trmm(dc, n)
    Lref = orc.cholesky_lower(orc.build_M(Phi, 0.45, 1.2))
    assert np.all(np.isfinite(L)) and np.all(np.isfinite(T))
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) < 1e-13
    assert np.linalg.norm(T - np.linalg.inv(Lref)) / np.linalg.norm(np.linalg.inv(Lref)) < 1e-13
