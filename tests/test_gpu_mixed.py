"""GPU tests of the mixed-precision path (int8 genotypes, 6 int8 digit slices
of L^-1 in balanced radix 254, exact int32 MFMA accumulation): its b and V
against the FP64 oracle within the stated bound, and against the device's own
FP64 path."""
import numpy as np
import pytest

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem
from tests.test_gpu_parity import setup, v_flat

pytestmark = pytest.mark.gpu

# Stated bound of the mixed path: b relative to max(|b|, se), standard errors
# element-wise and covariances relative to max(|V_rc|, se_r se_c), <= 2e-11;
# the variances var(b_ik) element-wise <= 1e-12 (the representation of L^-1
# keeps 3.7e-15 of each row's maximum; the measured errors at n = 1e4 are
# 7e-13 and 3e-14).  The plain element-wise error of b and of the covariances
# is reported, not gated: it is unbounded for entries near zero (at any
# precision).
TOL_SCALED = 2e-11
TOL_VAR = 1e-12


@pytest.mark.parametrize("n,m,wXL,h2", [(1000, 2000, 3, 0.4), (333, 257, 3, 0.6), (2500, 700, 5, 0.3),
                                        (64, 10, 1, 0.5), (4096, 300, 3, 0.5)])
def test_mixed_matches_oracle(pkg, gpu, n, m, wXL, h2):
    print(f"mixed vs oracle n={n} m={m} wXL={wXL}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=7 * n + m, h2=h2)
    tr, ctx = setup(pkg, prob, max_block=max(m, 64))
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    B, V = ctx.solve_block_i8(X8, snp_base=3)
    ctx.check()
    e = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), b_ref, v_flat(V_ref), wXL + 1)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, e
    assert e["var_rel"] < TOL_VAR, e
    assert e["B_norm"] < TOL_VAR and e["V_norm"] < TOL_VAR, e
    # the FP64 path of the same context agrees too (mode switch keeps the trait)
    ctx.set_mode(pkg.FGLS_MODE_FP64)
    B64, V64 = ctx.solve_block(pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4))
    e64 = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), B64.cpu().numpy(), V64.cpu().numpy(), wXL + 1)
    assert e64["B_scaled"] < TOL_SCALED and e64["V_scaled"] < TOL_SCALED, e64
    ctx.close()


def test_mixed_mode_set_before_trait_and_ragged_blocks(pkg, gpu):
    import torch
    prob = make_problem(n=777, wXL=3, m=300, seed=5)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    tr = orc.trait_precompute(L, prob.XL, prob.y)
    ctx = pkg.DeviceContext(prob.n, 3, 1, max_block=128)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    ctx.set_trait(torch.from_numpy(np.ascontiguousarray(L.T)).cuda(), pkg.colmajor(prob.XL),
                  torch.from_numpy(prob.y).cuda())
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    Bs, Vs = [], []
    for i0 in range(0, prob.m, 128):
        B, V = ctx.solve_block_i8(X8[i0:i0 + 128], snp_base=i0)
        Bs.append(B.cpu().numpy())
        Vs.append(V.cpu().numpy())
    ctx.check()
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    e = orc.parity_errors(np.concatenate(Bs), np.concatenate(Vs), b_ref, v_flat(V_ref), 4)
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, e
    ctx.close()


def test_to_i8_rejects_non_integer_genotypes(pkg, gpu):
    prob = make_problem(n=200, wXL=3, m=10, seed=1)
    _, ctx = setup(pkg, prob, max_block=16)
    X = prob.XR.copy()
    X[5, 3] = 0.5
    X[7, 4] = 300.0
    with pytest.raises(pkg.FglsError) as ei:
        ctx.to_i8(pkg.colmajor(X))
    assert "2 genotype values" in ei.value.msg
    ctx.close()


def test_mixed_singular_snp_reports_index(pkg, gpu):
    prob = make_problem(n=300, wXL=3, m=100, seed=9)
    prob.XR[:, 41] = 0.0
    tr, ctx = setup(pkg, prob, max_block=128)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    ctx.solve_block_i8(ctx.to_i8(pkg.colmajor(prob.XR)), snp_base=10)
    with pytest.raises(pkg.FglsError) as ei:
        ctx.check()
    assert ei.value.bad_snp == 51 and ei.value.info == 4
    ctx.close()


@pytest.mark.parametrize("n,m,wXL,block", [(1000, 700, 3, 700), (333, 300, 7, 128), (2500, 257, 5, 257)])
def test_mixed_dosages_match_oracle(pkg, gpu, n, m, wXL, block):
    """Imputed dosages (3 decimals in [0, 2]) on the int8 path as two digit
    planes: same bound as genotypes against the FP64 oracle."""
    print(f"mixed dosages n={n} m={m} wXL={wXL}", flush=True)
    import torch
    prob = make_problem(n=n, wXL=wXL, m=m, seed=3 * n + m, h2=0.45, dosage="imputed")
    tr, ctx = setup(pkg, prob, max_block=block)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, 1)
    XR = pkg.colmajor(prob.XR)
    Bs, Vs = [], []
    for i0 in range(0, m, block):
        cnt = min(block, m - i0)
        X8 = ctx.to_dosage_i8(XR[i0:i0 + cnt].contiguous(), 1000.0)
        B, V = ctx.solve_block_dosage(X8, 1000.0, snp_base=i0)
        Bs.append(B.cpu().numpy())
        Vs.append(V.cpu().numpy())
    ctx.check()
    e = orc.parity_errors(np.concatenate(Bs), np.concatenate(Vs), b_ref, v_flat(V_ref), wXL + 1)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, e
    assert e["var_rel"] < TOL_VAR, e
    # values that are not multiples of 1/1000 are rejected
    bad = XR[:4].clone()
    bad[0, 5] += 0.0004
    with pytest.raises(pkg.FglsError):
        ctx.to_dosage_i8(bad, 1000.0)
    # ... also when the offset is tiny (1e-8 / scale): nothing is rounded
    bad = XR[:4].clone()
    bad[1, 7] += 1e-11
    with pytest.raises(pkg.FglsError) as ei:
        ctx.to_dosage_i8(bad, 1000.0)
    assert "1 dosage values" in ei.value.msg
    ctx.to_dosage_i8(XR[:4].clone(), 1000.0)      # the exact values pass
    ctx.close()


def test_tiled_genotype_entry_points(pkg, gpu):
    """tile_i8 + the *_i8t entry points give bitwise the results of the
    column-major entry points (Cholesky and eigen contexts, genotypes and
    dosage digit planes, a count that is not a multiple of 16)."""
    import torch
    prob = make_problem(n=650, wXL=3, m=203, seed=9, t=1)
    _, ctx = setup(pkg, prob, max_block=256)
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    B1, V1 = ctx.solve_block_i8(X8)
    X8T = pkg.tile_i8(X8, prob.n)
    assert X8T.numel() == pkg.lib().fgls_hip_x8t_bytes(prob.n, prob.m, 1)
    B2, V2 = ctx.solve_block_i8t(X8T, prob.m)
    assert torch.equal(B1, B2) and torch.equal(V1, V2)
    D = np.round(np.random.default_rng(1).uniform(0, 2, size=(prob.n, prob.m)), 3)
    D8 = ctx.to_dosage_i8(pkg.colmajor(D), 1000.0)
    B3, V3 = ctx.solve_block_dosage(D8, 1000.0)
    B4, V4 = ctx.solve_block_i8t(pkg.tile_i8(D8, prob.n, prob.m, planes=2), prob.m, planes=2, scale=1000.0)
    assert torch.equal(B3, B4) and torch.equal(V3, V4)
    ctx.check()
    Z, W = pkg.eigen_decompose(prob.Phi, threads=4)
    e = pkg.EigenContext(prob.n, 3, max_traits=2, max_block=256)
    e.set_mode(pkg.FGLS_MODE_MIXED_I8)
    e.set_basis(torch.from_numpy(np.ascontiguousarray(Z.T)).cuda(), torch.from_numpy(W).cuda(),
                pkg.colmajor(prob.XL))
    e.set_traits([0.4, 0.7], [1.0, 1.0], pkg.colmajor(np.column_stack([prob.y, prob.y[::-1]])))
    E1 = e.solve_block_i8(X8)
    E2 = e.solve_block_i8t(X8T, prob.m)
    assert torch.equal(E1[0], E2[0]) and torch.equal(E1[1], E2[1])
    e.check()
    e.close()
    ctx.close()


@pytest.mark.parametrize("n,m", [(650, 203), (1001, 77), (64, 16), (7, 5), (4096, 300)])
def test_u2_tiles_equal_i8_tiles(pkg, gpu, n, m):
    """2-bit tiles (tile_u2 from int8 columns, tile_u2_packed from the 2-bit
    file layout) give bitwise the results of the 8-bit tiles in both
    variants: the product unpacks them to the same int8 operands.  Padding
    rows of the int8 input (rows >= n) may hold anything."""
    import torch
    print(f"u2 tiles n={n} m={m}", flush=True)
    prob = make_problem(n=n, wXL=min(3, n - 1), m=m, seed=n + 17)
    _, ctx = setup(pkg, prob, max_block=max(m, 64))
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    X8 = ctx.to_i8(pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4))
    B1, V1 = ctx.solve_block_i8t(pkg.tile_i8(X8, n), m)
    X8g = X8.clone()
    X8g[:, n:] = 100                                         # garbage padding rows
    T2 = pkg.tile_u2(X8g, n)
    assert T2.numel() == pkg.x2t_bytes(n, m) == ((m + 15) // 16) * 16 * ((n + 63) // 64) * 16
    B2, V2 = ctx.solve_block_u2t(T2, m)
    P = torch.from_numpy(pkg.pack_u2(prob.XR)).cuda()
    B3, V3 = ctx.solve_block_u2t(pkg.tile_u2_packed(P, n), m)
    ctx.check()
    assert torch.equal(B1, B2) and torch.equal(V1, V2)
    assert torch.equal(B1, B3) and torch.equal(V1, V3)
    assert torch.equal(T2, pkg.tile_u2_packed(P, n))
    # the same through a block_reduce / block_finalize pair
    ctx.block_reduce_u2t(T2, m)
    B4 = torch.empty_like(B1)
    V4 = torch.empty_like(V1)
    ctx.block_finalize(m, B4, V4)
    assert torch.equal(B1, B4) and torch.equal(V1, V4)
    # eigen variant, two traits
    Z, W = pkg.eigen_decompose(prob.Phi, threads=4)
    with pkg.EigenContext(n, prob.wXL, max_traits=2, max_block=max(m, 64)) as e:
        e.set_mode(pkg.FGLS_MODE_MIXED_I8)
        e.set_basis(torch.from_numpy(np.ascontiguousarray(Z.T)).cuda(), torch.from_numpy(W).cuda(),
                    pkg.colmajor(prob.XL))
        e.set_traits([0.4, 0.7], [1.0, 1.0], pkg.colmajor(np.column_stack([prob.y, prob.y[::-1]])))
        E1 = e.solve_block_i8t(pkg.tile_i8(X8, n), m)
        E2 = e.solve_block_u2t(T2, m)
        e.check()
        assert torch.equal(E1[0], E2[0]) and torch.equal(E1[1], E2[1])
    ctx.close()


def test_u2_tiles_reject_values_outside_0_3(pkg, gpu):
    import torch
    n, m = 300, 40
    X8 = torch.randint(0, 3, (m, 320), dtype=torch.int8, device="cuda")
    X8[:, n:] = -5                                           # padding: ignored
    pkg.tile_u2(X8, n)
    X8[3, 17] = 3                                            # 3 fits in two bits: allowed
    pkg.tile_u2(X8, n)
    X8[5, 299] = 4
    X8[6, 0] = -1
    with pytest.raises(pkg.FglsError) as ei:
        pkg.tile_u2(X8, n)
    assert "2 genotype values" in ei.value.msg
    P = torch.from_numpy(pkg.pack_u2(np.zeros((n, m)))).cuda()
    P[7, 74] |= 0xC0                                         # individual 299: code 3
    with pytest.raises(pkg.FglsError) as ei:
        pkg.tile_u2_packed(P, n)
    assert "1 2-bit codes" in ei.value.msg


@pytest.mark.parametrize("n,m,wXL,wXR", [(650, 120, 3, 2), (1001, 77, 13, 3), (64, 20, 1, 2), (333, 257, 2, 4)])
def test_mixed_several_columns_per_snp(pkg, gpu, n, m, wXL, wXR):
    """wXR > 1 in the mixed mode: the int8 product stores L^-1 X_R, the FP64
    per-SNP stage follows; against the oracle within the mixed bound, the
    FP64 path of the same context within the same bound, and the column-major,
    8-bit tiled and 2-bit tiled inputs bitwise equal to one another."""
    import torch
    print(f"mixed wXR={wXR} n={n} m={m} wXL={wXL}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=3 * n + wXR, wXR=wXR)
    tr, ctx = setup(pkg, prob, wXR=wXR, max_block=max(m, 64))
    ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, wXR)
    p = wXL + wXR
    X8 = ctx.to_i8(pkg.colmajor(prob.XR))
    assert X8.shape[0] == m * wXR
    B, V = ctx.solve_block_i8(X8, snp_base=5)
    ctx.check()
    e = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), b_ref, v_flat(V_ref), p)
    assert e["V_upper_zero"]
    assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, e
    assert e["var_rel"] < TOL_VAR, e
    B2, V2 = ctx.solve_block_i8t(pkg.tile_i8(X8, n), m)
    assert torch.equal(B, B2) and torch.equal(V, V2)
    T2 = pkg.tile_u2(X8, n)
    B3, V3 = ctx.solve_block_u2t(T2, m)
    assert torch.equal(B, B3) and torch.equal(V, V3)
    P = torch.from_numpy(pkg.pack_u2(prob.XR)).cuda()
    assert torch.equal(T2, pkg.tile_u2_packed(P, n))
    # two ragged blocks through reduce / finalize
    h = (m + 1) // 2
    Bs, Vs = [], []
    for i0, c in ((0, h), (h, m - h)):
        ctx.block_reduce_i8(X8[i0 * wXR:(i0 + c) * wXR], c)
        Bb = torch.empty((c, p), dtype=torch.float64, device="cuda")
        Vb = torch.empty((c, p * p), dtype=torch.float64, device="cuda")
        ctx.block_finalize(c, Bb, Vb, snp_base=i0)
        Bs.append(Bb)
        Vs.append(Vb)
    assert torch.equal(torch.cat(Bs), B) and torch.equal(torch.cat(Vs), V)
    # dosage digits need wXR = 1
    D8 = ctx.to_dosage_i8(pkg.colmajor(prob.XR[:, :wXR * 4]), 1000.0)
    with pytest.raises(pkg.FglsError) as ei:
        ctx.solve_block_dosage(D8, 1000.0, count=4)
    assert ei.value.code == pkg.FGLS_ERR_ARG and "wXR = 1" in ei.value.msg
    ctx.set_mode(pkg.FGLS_MODE_FP64)
    B64, V64 = ctx.solve_block(pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4), count=m)
    e64 = orc.parity_errors(B.cpu().numpy(), V.cpu().numpy(), B64.cpu().numpy(), V64.cpu().numpy(), p)
    assert e64["B_scaled"] < TOL_SCALED and e64["V_scaled"] < TOL_SCALED, e64
    ctx.close()
