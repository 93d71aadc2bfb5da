"""GPU parity for wide models: p = 12 and p = 16 (the largest p the per-SNP
stage is built for), which a scan with sex, age and 10 principal components
reaches.  Every p from 1 to 16 has its own instantiation of the finalize,
the per-SNP wavefront kernel, the FP64 epilogues and the int8 epilogue; the
other GPU tests stop at p = 8.  Each path is checked against the FP64 oracle:
FP64 (fused product and reductions), mixed (int8 genotypes), dosages (two
int8 digit planes), the eigen variant in both precisions (two traits), and
wXR > 1 at p = 16 through the wavefront per-SNP kernel.
"""
import numpy as np
import pytest
import torch

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem
from tests.test_gpu_mixed import TOL_SCALED, TOL_VAR
from tests.test_gpu_parity import check_bv, setup, v_flat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wXL", [11, 15])
def test_wide_p_fp64_mixed_dosage(pkg, gpu, wXL):
    n, m = 900, 333
    p = wXL + 1
    print(f"wide p={p}: fp64, mixed, dosages", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=40 + wXL, h2=0.45)
    tr, ctx = setup(pkg, prob, max_block=512)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    XR = pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4)
    B, V = ctx.solve_block(XR, snp_base=7)
    ctx.check()
    check_bv(B, V, b_ref, V_ref, p)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    B8, V8 = ctx.solve_block_i8(ctx.to_i8(XR), snp_base=7)
    ctx.check()
    e = orc.parity_errors(B8.cpu().numpy(), V8.cpu().numpy(), b_ref, v_flat(V_ref), p)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED and e["var_rel"] < TOL_VAR, e
    # dosages on the same context
    D = np.round(np.random.default_rng(wXL).uniform(0.0, 2.0, size=(n, m)), 3)
    bd, Vd = orc.solve_snps(D, tr, 1)
    D8 = ctx.to_dosage_i8(pkg.colmajor(D), 1000.0)
    BD, VD = ctx.solve_block_dosage(D8, 1000.0)
    ctx.check()
    e = orc.parity_errors(BD.cpu().numpy(), VD.cpu().numpy(), bd, v_flat(Vd), p)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED and e["var_rel"] < TOL_VAR, e
    ctx.close()


@pytest.mark.parametrize("wXL", [11, 15])
def test_wide_p_eigen(pkg, gpu, wXL):
    """The eigen variant (one product Z' X_R, two traits) at p = 12 and 16,
    FP64 and mixed, against the oracle's Cholesky form of each trait."""
    n, m = 700, 257
    p = wXL + 1
    print(f"wide p={p}: eigen fp64 and mixed", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=50 + wXL, h2=0.4, t=2)
    H = [0.4, 0.7]
    Z, W = pkg.eigen_decompose(prob.Phi, threads=4)
    refs = []
    for k, h in enumerate(H):
        trk = orc.trait_precompute(orc.cholesky_lower(orc.build_M(prob.Phi, h, 1.0)), prob.XL, prob.Y[:, k])
        refs.append(orc.solve_snps(prob.XR, trk, 1))
    for mode in (pkg.FGLS_MODE_FP64, pkg.FGLS_MODE_MIXED_I8):
        with pkg.EigenContext(n, wXL, max_traits=2, max_block=512) as eg:
            eg.set_mode(mode)
            eg.set_basis(torch.from_numpy(np.ascontiguousarray(Z.T)).cuda(), torch.from_numpy(W).cuda(),
                         pkg.colmajor(prob.XL))
            eg.set_traits(H, [1.0, 1.0], pkg.colmajor(prob.Y))
            XR = pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4)
            if mode == pkg.FGLS_MODE_FP64:
                BE, VE = eg.solve_block(XR)
            else:
                BE, VE = eg.solve_block_i8(pkg.to_i8(XR, n))
            eg.check()
            for k in range(2):
                e = orc.parity_errors(BE[k].cpu().numpy(), VE[k].cpu().numpy(), refs[k][0], v_flat(refs[k][1]), p)
                tol = 1e-10 if mode == pkg.FGLS_MODE_FP64 else TOL_SCALED
                assert e["V_upper_zero"] and e["B_scaled"] < tol and e["V_scaled"] < tol, (mode, k, e)


def test_wide_p_generic_wxr(pkg, gpu):
    """wXR = 3 at p = 16 (wXL = 13): the TRMM to HBM, then the wavefront
    per-SNP kernel; and the in-place TRSM followed by the same kernel."""
    print("wide p=16 with wXR=3", flush=True)
    wXL, wXR, m = 13, 3, 120
    prob = make_problem(n=650, wXL=wXL, m=m, seed=61, wXR=wXR)
    tr, ctx = setup(pkg, prob, wXR=wXR, max_block=256)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, wXR)
    ld = (prob.n + 3) // 4 * 4
    B, V = ctx.solve_block(pkg.colmajor(prob.XR, ld=ld), count=m)
    ctx.check()
    check_bv(B, V, b_ref, V_ref, wXL + wXR)
    X = pkg.colmajor(prob.XR, ld=ld)
    ctx.trsm(X)
    B2, V2 = ctx.snp_stage(X, count=m)
    ctx.check()
    check_bv(B2, V2, b_ref, V_ref, wXL + wXR)
    ctx.close()


def test_p_above_16_is_an_argument_error(pkg, gpu):
    with pytest.raises(pkg.FglsError) as ei:
        pkg.DeviceContext(100, 16, 1, max_block=64)
    assert ei.value.code == pkg.FGLS_ERR_ARG and "p<=16" in ei.value.msg
    with pytest.raises(pkg.FglsError) as ei:
        pkg.EigenContext(100, 16, max_traits=1, max_block=64)
    assert ei.value.code == pkg.FGLS_ERR_ARG
