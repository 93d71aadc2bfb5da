"""GPU tests of the two digit forms of L^-1 in the mixed mode.

Six digits: row i of L^-1 scaled by its largest entry, six balanced radix-254
int8 digits per entry.  Five digits: row i scaled by its largest entry
outside its diagonal 64-column block, five digits everywhere plus a top digit
on that block; its per-entry error bound is 254 off_i / max_i (the "loss")
times the six-digit bound, and the context uses it (FGLS_SLICES_AUTO) when the
loss over all rows is at most 16.

Both forms are checked against the FP64 oracle at the mixed path's bar on
every product that reads the slices: 8-bit and 2-bit genotype tiles, dosage
digit planes, several columns per SNP (the stored L^-1 X_R) and the mixed
adjoint; the loss the device reports is checked against the one computed from
the oracle's L^-1.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem
from tests.test_gpu_parity import setup, v_flat

pytestmark = pytest.mark.gpu

# the mixed path's bar (as in test_gpu_mixed)
TOL_SCALED = 2e-11
TOL_VAR = 1e-12
MAX_LOSS = 16.0


def loss_of(L):
    """254 max_i (largest |L^-1_ik| outside row i's diagonal 64-block) / (largest |L^-1_ik|)."""
    n = L.shape[0]
    Li = np.abs(sla.solve_triangular(L, np.eye(n), lower=True))
    mx = Li.max(axis=1)
    k = np.arange(n)[None, :]
    d0 = (np.arange(n) // 64 * 64)[:, None]
    off = np.where((k >= d0) & (k < d0 + 64), 0.0, Li).max(axis=1)
    return float(254.0 * np.max(off / mx))


def check(B, V, b_ref, V_ref, p):
    e = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), b_ref, v_flat(V_ref), p)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, e
    assert e["var_rel"] < TOL_VAR, e
    return e


def related_problem(n, m, seed):
    """Pairs of closely related samples (Phi_{i, i+n/2} = 0.5) at h2 = 0.8: the
    off-block entries of L^-1 are a large fraction of the row maximum."""
    prob = make_problem(n=n, wXL=3, m=m, seed=seed, h2=0.8)
    Phi = np.eye(n)
    h = n // 2
    idx = np.arange(h)
    Phi[idx, idx + h] = Phi[idx + h, idx] = 0.5
    prob.Phi = Phi
    return prob


@pytest.mark.parametrize("n,m,wXL,h2", [(1000, 700, 3, 0.4), (333, 257, 3, 0.6), (64, 40, 1, 0.5),
                                        (96, 50, 2, 0.5), (2500, 300, 5, 0.3), (4113, 200, 3, 0.5)])
def test_both_forms_match_oracle(pkg, gpu, n, m, wXL, h2):
    print(f"slice forms n={n} m={m} wXL={wXL}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=11 * n + m, h2=h2)
    tr, ctx = setup(pkg, prob, max_block=max(m, 64))
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, h2, 1.0))
    loss_ref = loss_of(L)
    ns, loss = ctx.slice_info()
    assert abs(loss - loss_ref) <= 1e-8 * max(loss_ref, 1.0), (loss, loss_ref)
    assert ns == (5 if loss_ref <= MAX_LOSS else 6), (ns, loss_ref)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    B, V = ctx.solve_block_i8(X8)
    ctx.check()
    check(B, V, b_ref, V_ref, wXL + 1)
    # 2-bit tiles read the same slices: bitwise the same b and V
    B2, V2 = ctx.solve_block_u2t(pkg.tile_u2(X8, n, m), m)
    ctx.check()
    assert np.array_equal(B2.cpu().numpy(), B.cpu().numpy()) and np.array_equal(V2.cpu().numpy(), V.cpu().numpy())
    # six digits on the same context
    ctx.set_slice_form(pkg.FGLS_SLICES_SIX)
    assert ctx.slice_info() == (6, -1.0)
    B6, V6 = ctx.solve_block_i8(X8)
    ctx.check()
    check(B6, V6, b_ref, V_ref, wXL + 1)
    B6u, _ = ctx.solve_block_u2t(pkg.tile_u2(X8, n, m), m)
    assert np.array_equal(B6u.cpu().numpy(), B6.cpu().numpy())
    # and back: the automatic form again gives the first result bitwise
    ctx.set_slice_form(pkg.FGLS_SLICES_AUTO)
    assert ctx.slice_info()[0] == ns
    B5, _ = ctx.solve_block_i8(X8)
    assert np.array_equal(B5.cpu().numpy(), B.cpu().numpy())
    ctx.close()


def test_related_samples_keep_six_digits(pkg, gpu):
    """A matrix whose off-block entries of L^-1 are large: the automatic form
    falls back to six digits (loss > 16) and still meets the bar."""
    prob = related_problem(600, 200, seed=5)
    tr, ctx = setup(pkg, prob, max_block=256)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    loss_ref = loss_of(L)
    assert loss_ref > MAX_LOSS
    ns, loss = ctx.slice_info()
    assert ns == 6 and abs(loss - loss_ref) <= 1e-8 * loss_ref, (ns, loss, loss_ref)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    B, V = ctx.solve_block_i8(ctx.to_i8(pkg.colmajor(prob.XR)))
    ctx.check()
    check(B, V, b_ref, V_ref, 4)
    ctx.close()


@pytest.mark.parametrize("n,m", [(1000, 300), (777, 129)])
def test_five_digit_dosages_match_oracle(pkg, gpu, n, m):
    """Dosage digit planes (x = (128 hi + lo) / 1000) on the five-digit form."""
    print(f"five-digit dosages n={n} m={m}", flush=True)
    prob = make_problem(n=n, wXL=3, m=m, seed=n + 3 * m, h2=0.45, dosage="imputed")
    tr, ctx = setup(pkg, prob, max_block=m)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    assert ctx.slice_info()[0] == 5
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    D8 = ctx.to_dosage_i8(pkg.colmajor(prob.XR), 1000.0)
    B, V = ctx.solve_block_dosage(D8, 1000.0)
    ctx.check()
    check(B, V, b_ref, V_ref, 4)
    ctx.set_slice_form(pkg.FGLS_SLICES_SIX)
    B6, V6 = ctx.solve_block_dosage(D8, 1000.0)
    ctx.check()
    check(B6, V6, b_ref, V_ref, 4)
    ctx.close()


@pytest.mark.parametrize("n,m,wXR", [(1000, 90, 2), (650, 64, 3)])
def test_five_digit_several_columns_per_snp(pkg, gpu, n, m, wXR):
    """wXR > 1: the int8 product stores L^-1 X_R from the five-digit form."""
    print(f"five-digit wXR={wXR} n={n} m={m}", flush=True)
    prob = make_problem(n=n, wXL=3, m=m, seed=n + wXR, wXR=wXR)
    tr, ctx = setup(pkg, prob, wXR=wXR, max_block=m)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    assert ctx.slice_info()[0] == 5
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, wXR)
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    B, V = ctx.solve_block_i8(X8)
    ctx.check()
    check(B, V, b_ref, V_ref, 3 + wXR)
    B2, V2 = ctx.solve_block_u2t(pkg.tile_u2(X8, n, m * wXR), m)
    assert np.array_equal(B2.cpu().numpy(), B.cpu().numpy())
    ctx.close()


def test_five_digit_mixed_adjoint_matches_oracle(pkg, gpu):
    """The mixed adjoint's z = L^-1 x on the five-digit form (H = G L^-1
    keeps six digits) against the oracle's adjoint."""
    import torch
    from cugwas_proto_amd.adjoint import snp_loss_grad
    n, m = 1000, 600
    pr = make_problem(n=n, wXL=3, m=m, seed=21, h2=0.5)
    w = np.random.default_rng(22).uniform(0.0, 1.0, m)
    XR = torch.from_numpy(np.ascontiguousarray(pr.XR.T)).cuda()
    XL = torch.from_numpy(np.ascontiguousarray(pr.XL.T)).cuda()
    y = torch.from_numpy(pr.y.copy()).cuda()
    wd = torch.from_numpy(w).cuda()
    L = orc.cholesky_lower(orc.build_M(pr.Phi, pr.h2, 1.0))
    assert loss_of(L) <= MAX_LOSS   # the forward's slices are the five-digit form
    loss, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, 256, precision="mixed")
    loss_o, gy_o, gh_o = orc.snp_loss_adjoint(pr.Phi, pr.h2, 1.0, pr.XL, pr.XR, pr.y, w)
    gy = gy.cpu().numpy()
    assert abs(loss - loss_o) <= 1e-9 * abs(loss_o)
    assert np.max(np.abs(gy - gy_o)) <= 1e-9 * np.max(np.abs(gy_o))
    assert abs(gh - gh_o) <= 1e-9 * abs(gh_o)
