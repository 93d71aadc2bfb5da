"""GPU tests of the eigenbasis variant (the reference's fgls_eigen, src/fgls_eigen.c):
the device context (one product Z' X_R per block for every trait, FP64 and the
int8 mixed path) against the oracle's restatement of fgls_eigen and against
the Cholesky oracle (the same b and V up to rounding), and the file-level
entry point fgls_eigen / driver.x variant 'e'."""
import os
import subprocess

import numpy as np
import pytest

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem, write_problem_dir
from tests.test_gpu_mixed import TOL_SCALED, TOL_VAR
from tests.test_gpu_parity import v_flat

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cugwas-proto_amd", "lib", "driver.x")

TOL_FP64 = 1e-10


def eig_setup(pkg, prob, H, Sig, max_block, mixed=False):
    import torch
    Z, W = pkg.eigen_decompose(prob.Phi, threads=4)
    e = pkg.EigenContext(prob.n, prob.wXL, max_traits=len(H), max_block=max_block)
    if mixed:
        e.set_mode(pkg.FGLS_MODE_MIXED_I8)
    e.set_basis(torch.from_numpy(np.ascontiguousarray(Z.T)).cuda(), torch.from_numpy(W).cuda(),
                pkg.colmajor(prob.XL) if prob.wXL else None)
    e.set_traits(H, Sig, pkg.colmajor(prob.Y[:, :len(H)]))
    return e


def oracle_traits(prob, H, Sig):
    Z, W = orc.eigen_decompose(prob.Phi)
    out = []
    for j in range(len(H)):
        tr = orc.eigen_trait(Z, W, prob.XL, prob.Y[:, j], float(H[j]), float(Sig[j]))
        b, V = orc.eigen_solve_snps(Z, tr, prob.XR, 1)
        out.append((tr, b, V))
    return out


@pytest.mark.parametrize("n,m,wXL,t", [(1000, 2000, 3, 1), (333, 257, 3, 3), (2500, 700, 5, 2),
                                       (64, 10, 1, 1), (300, 129, 0, 2)])
def test_eigen_fp64_matches_oracle(pkg, gpu, n, m, wXL, t):
    print(f"eigen fp64 n={n} m={m} wXL={wXL} t={t}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=11 * n + m, h2=0.4, t=t)
    H = np.linspace(0.2, 0.8, t)
    Sig = np.linspace(0.7, 1.9, t)
    e = eig_setup(pkg, prob, H, Sig, max_block=max(m, 64))
    B, V = e.solve_block(pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4), snp_base=5)
    e.check()
    p = wXL + 1
    for j, (tr, b_ref, V_ref) in enumerate(oracle_traits(prob, H, Sig)):
        Bt, Vtl, res = e.trait(j)
        assert abs(res - tr.res_sigma) <= 1e-12 * tr.res_sigma
        assert np.allclose(Bt, tr.B_t, rtol=1e-11, atol=1e-11 * np.abs(tr.B_t).max(initial=1.0))
        err = orc.parity_errors(B[j].cpu().numpy(), V[j].cpu().numpy(), b_ref, v_flat(V_ref), p)
        assert err["V_upper_zero"]
        assert err["B_scaled"] < TOL_FP64 and err["V_scaled"] < TOL_FP64, (j, err)
        # the Cholesky form of the same trait gives the same b and V
        L = orc.cholesky_lower(orc.build_M(prob.Phi, float(H[j]), float(Sig[j])))
        bc, Vc = orc.solve_snps(prob.XR, orc.trait_precompute(L, prob.XL, prob.Y[:, j]), 1)
        errc = orc.parity_errors(B[j].cpu().numpy(), V[j].cpu().numpy(), bc, v_flat(Vc), p)
        assert errc["B_scaled"] < TOL_FP64 and errc["V_scaled"] < TOL_FP64, (j, errc)
    e.close()


@pytest.mark.parametrize("n,m,wXL,t", [(1000, 2000, 3, 1), (777, 300, 3, 3), (2500, 700, 7, 2)])
def test_eigen_mixed_matches_oracle(pkg, gpu, n, m, wXL, t):
    print(f"eigen mixed n={n} m={m} wXL={wXL} t={t}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=13 * n + m, h2=0.5, t=t)
    H = np.linspace(0.3, 0.7, t)
    Sig = np.ones(t)
    e = eig_setup(pkg, prob, H, Sig, max_block=max(m, 64), mixed=True)
    dc = pkg.DeviceContext(prob.n, wXL, 1, max_block=16)   # for the int8 conversion only
    X8 = dc.to_i8(pkg.colmajor(prob.XR))
    B, V = e.solve_block_i8(X8)
    e.check()
    for j, (tr, b_ref, V_ref) in enumerate(oracle_traits(prob, H, Sig)):
        err = orc.parity_errors(B[j].cpu().numpy(), V[j].cpu().numpy(), b_ref, v_flat(V_ref), wXL + 1)
        assert err["B_scaled"] < TOL_SCALED and err["V_scaled"] < TOL_SCALED, (j, err)
        assert err["var_rel"] < TOL_VAR, (j, err)
    dc.close()
    e.close()


def test_eigen_mixed_dosages_and_ragged_blocks(pkg, gpu):
    import torch
    prob = make_problem(n=900, wXL=3, m=300, seed=21, h2=0.3, t=2, dosage="imputed")
    H, Sig = np.array([0.3, 0.6]), np.array([1.0, 1.3])
    e = eig_setup(pkg, prob, H, Sig, max_block=128, mixed=True)
    dc = pkg.DeviceContext(prob.n, 3, 1, max_block=16)
    Bs, Vs = [], []
    for i0 in range(0, prob.m, 128):
        X = pkg.colmajor(prob.XR[:, i0:i0 + 128])
        D8 = dc.to_dosage_i8(X, 1000.0)
        B, V = e.solve_block_dosage(D8, 1000.0, snp_base=i0)
        Bs.append(B.cpu().numpy())
        Vs.append(V.cpu().numpy())
    e.check()
    B = np.concatenate(Bs, axis=1)
    V = np.concatenate(Vs, axis=1)
    for j, (tr, b_ref, V_ref) in enumerate(oracle_traits(prob, H, Sig)):
        err = orc.parity_errors(B[j], V[j], b_ref, v_flat(V_ref), 4)
        assert err["B_scaled"] < TOL_SCALED and err["V_scaled"] < TOL_SCALED, (j, err)
        assert err["var_rel"] < TOL_VAR, (j, err)
    dc.close()
    e.close()


def test_eigen_rejects_non_positive_M_and_bad_strides(pkg, gpu):
    import torch
    prob = make_problem(n=200, wXL=3, m=20, seed=2, t=2)
    e = eig_setup(pkg, prob, np.array([0.4, 0.5]), np.ones(2), max_block=32)
    # h2 = 1.5: a_k W + b_k = 1.5 W - 0.5 < 0 for the small eigenvalues of Phi
    with pytest.raises(pkg.FglsError) as ei:
        e.set_traits([0.4, 1.5], [1.0, 1.0], pkg.colmajor(prob.Y))
    assert ei.value.code == pkg.FGLS_ERR_NOTSPD and "not positive definite" in ei.value.msg
    with pytest.raises(pkg.FglsError) as ei:
        e.solve_block(pkg.colmajor(prob.XR))    # the failed set_traits left no traits set
    assert ei.value.code == pkg.FGLS_ERR_STATE
    e.set_traits([0.4, 0.5], [1.0, 1.0], pkg.colmajor(prob.Y))
    B = torch.empty((2, 20 * 4), dtype=torch.float64, device="cuda")
    V = torch.empty((2, 20 * 16), dtype=torch.float64, device="cuda")
    lib = pkg.lib()
    X = pkg.colmajor(prob.XR)
    rc = lib.fgls_hip_eig_solve_block(e._h, X.data_ptr(), X.shape[1], 20, B.data_ptr(), 40,
                                      V.data_ptr(), 320, 0, None)
    assert rc == pkg.FGLS_ERR_ARG and "strides" in pkg.last_error()
    e.close()


def test_device_eigendecomposition(pkg, gpu):
    """fgls_hip_eigen_decompose against the host dsyevr: eigenvalues, the
    reconstruction of Phi and orthogonality of Z; and the eigen context built
    from it solves as from the host decomposition."""
    import torch
    prob = make_problem(n=700, wXL=3, m=200, seed=31, t=1)
    Phi_d = pkg.colmajor(prob.Phi)
    Z, W = pkg.eigen_decompose_device(Phi_d)
    Zc = Z.cpu().numpy().T            # column-major rows -> numpy columns
    Wc = W.cpu().numpy()
    Zh, Wh = pkg.eigen_decompose(prob.Phi, threads=4)
    assert np.abs(Wc - Wh).max() < 1e-12 * np.abs(Wh).max()
    assert np.abs(Zc @ np.diag(Wc) @ Zc.T - prob.Phi).max() < 1e-12
    assert np.abs(Zc.T @ Zc - np.eye(prob.n)).max() < 1e-12
    e = pkg.EigenContext(prob.n, 3, max_traits=1, max_block=256)
    e.set_basis(Z, W, pkg.colmajor(prob.XL))
    e.set_traits([0.4], [1.0], pkg.colmajor(prob.Y))
    B, V = e.solve_block(pkg.colmajor(prob.XR))
    e.check()
    (tr, b_ref, V_ref), = oracle_traits(prob, [0.4], [1.0])
    err = orc.parity_errors(B[0].cpu().numpy(), V[0].cpu().numpy(), b_ref, v_flat(V_ref), 4)
    assert err["B_scaled"] < TOL_FP64 and err["V_scaled"] < TOL_FP64, err
    e.close()


def write_dir(prob, d, H, Sig):
    write_problem_dir(prob, d, H=H, Sig=Sig)


@pytest.mark.parametrize("opts", [None, {"precision": "mixed"}, {"traits_per_pass": 2, "block": 100},
                                  {"precision": "mixed", "xr_format": "i8"}, {"eig": "host"}])
def test_fgls_eigen_files_match_oracle(pkg, gpu, tmp_path, opts):
    n, wXL, m, t = 500, 3, 777, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=99, t=t)
    d = str(tmp_path)
    H = np.array([0.25, 0.5, 0.75])
    Sig = np.array([0.9, 1.0, 1.7])
    write_dir(prob, d, H, Sig)
    if opts and opts.get("xr_format") == "i8":
        prob.XR.T.astype(np.int8).tofile(os.path.join(d, "XR.i8"))
        xr = os.path.join(d, "XR.i8")
    else:
        xr = os.path.join(d, "XR.in")
    p = wXL + 1
    f = lambda s: os.path.join(d, s)
    pkg.fgls_eigen(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"), xr,
                   f("Y.in"), f("B.out"), f("V.out"), options=opts)
    B = np.fromfile(f("B.out"), dtype="<f8")
    V = np.fromfile(f("V.out"), dtype="<f8")
    assert B.size == t * m * p and V.size == t * m * p * p
    Bo, Vo = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, 1, 256)
    err = orc.parity_errors(B, V, Bo, Vo, p)
    tol = TOL_SCALED if opts and opts.get("precision") == "mixed" else TOL_FP64
    assert err["V_upper_zero"] and err["B_scaled"] < tol and err["V_scaled"] < tol, err
    assert not any(s.endswith(".tmp") for s in os.listdir(d))   # no Z'X temporaries


def test_driver_variant_e(pkg, gpu, tmp_path):
    """driver.x e <dir> 1 (driver.c:70-77) on the test_XL fixture: B against
    B_exp.out, V against the eigen oracle."""
    from tests.golden.test_xl_fixture import write_dir as write_fixture
    d = str(tmp_path / "test_XL")
    prm = write_fixture(d)
    stdin = open(os.path.join(d, "input.in")).read()
    res = subprocess.run([DRIVER, "e", d, "1", "--check"], input=stdin, capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stderr
    assert res.stdout.strip().splitlines()[-1].strip().isdigit()
    n, p, m, t = prm["n"], prm["p"], prm["m"], prm["t"]
    B = np.fromfile(os.path.join(d, "B.out"), dtype="<f8")
    Bexp = np.fromfile(os.path.join(d, "B_exp.out"), dtype="<f8")
    assert B.size == t * m * p
    assert orc.max_rel_err(B, Bexp) < 1e-10
    Phi, H, Sig, XL, XR, Y = orc.read_inputs(d, n, p, m, t, prm["wXL"])
    Bo, Vo = orc.fgls_eigen(Phi, H, Sig, XL, XR, Y, wXR=1, x_b=prm["x_b"])
    V = np.fromfile(os.path.join(d, "V.out"), dtype="<f8")
    err = orc.parity_errors(B, V, Bo, Vo, p)
    assert err["B_scaled"] < TOL_FP64 and err["V_scaled"] < TOL_FP64 and err["V_upper_zero"], err


@pytest.mark.parametrize("mixed", [False, True])
def test_eigen_singular_snp_reports_trait_and_index(pkg, gpu, tmp_path, mixed):
    """An all-zero SNP column makes S_i singular for every trait (zero last
    pivot, fgls_eigen.c:540-547): each trait records its first failing SNP;
    check() reports trait 0 and clears; the file entry point fails with the
    reference's message."""
    prob = make_problem(n=300, wXL=3, m=100, seed=19, t=2)
    prob.XR[:, 37] = 0.0
    prob.XR[:, 80] = 0.0
    H, Sig = np.array([0.3, 0.6]), np.array([1.0, 1.0])
    e = eig_setup(pkg, prob, H, Sig, max_block=128, mixed=mixed)
    if mixed:
        dc = pkg.DeviceContext(prob.n, 3, 1, max_block=16)
        B, V = e.solve_block_i8(dc.to_i8(pkg.colmajor(prob.XR)), snp_base=1000)
        dc.close()
    else:
        B, V = e.solve_block(pkg.colmajor(prob.XR), snp_base=1000)
    with pytest.raises(pkg.FglsError) as ei:
        e.check()
    assert ei.value.code == pkg.FGLS_ERR_NOTSPD
    assert (ei.value.trait, ei.value.bad_snp, ei.value.info) == (0, 1037, 4)
    e.check()   # cleared
    assert np.isnan(B[1, 37].cpu().numpy()).all() and np.isfinite(B[1, 36].cpu().numpy()).all()
    # an empty block is a no-op
    e.solve_block(pkg.colmajor(prob.XR[:, :1]), count=0)
    e.close()
    d = str(tmp_path)
    write_problem_dir(prob, d, H=H, Sig=Sig)
    f = lambda s: os.path.join(d, s)
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_eigen(prob.n, 4, prob.m, 2, 3, 1, 50, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                       f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"),
                       options={"precision": "mixed" if mixed else "fp64"})
    assert ei.value.code == pkg.FGLS_ERR_NOTSPD and "i: 37" in ei.value.msg


@pytest.mark.parametrize("n,m,wXL,wXR,t", [(700, 301, 3, 2, 2), (257, 130, 1, 3, 3), (400, 97, 5, 11, 1)])
def test_eigen_wide_matches_oracle(pkg, gpu, n, m, wXL, wXR, t):
    """Several columns per SNP (fgls_hip_eig_solve_block_wide): against the
    oracle's fgls_eigen restatement (pinned at wXR = 2 by the reference's own
    fgls_eigen) and the Cholesky form of each trait; p = 16 in the last case."""
    print(f"eigen wide n={n} m={m} wXL={wXL} wXR={wXR} t={t}", flush=True)
    prob = make_problem(n=n, wXL=wXL, m=m, seed=7 * n + m, h2=0.4, t=t, wXR=wXR)
    H = np.linspace(0.25, 0.75, t)
    Sig = np.linspace(0.8, 1.6, t)
    e = eig_setup(pkg, prob, H, Sig, max_block=max(m, 64))
    XR = pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4)
    B, V = e.solve_block_wide(XR, wXR, snp_base=3)
    e.check()
    p = wXL + wXR
    assert B.shape == (t, m, p) and V.shape == (t, m, p * p)
    Bo, Vo = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y[:, :t], wXR=wXR)
    err = orc.parity_errors(B.cpu().numpy().ravel(), V.cpu().numpy().ravel(), Bo, Vo, p)
    assert err["V_upper_zero"] and err["B_scaled"] < TOL_FP64 and err["V_scaled"] < TOL_FP64, err
    for j in range(t):
        L = orc.cholesky_lower(orc.build_M(prob.Phi, float(H[j]), float(Sig[j])))
        bc, Vc = orc.solve_snps(prob.XR, orc.trait_precompute(L, prob.XL, prob.Y[:, j]), wXR)
        errc = orc.parity_errors(B[j].cpu().numpy(), V[j].cpu().numpy(), bc, v_flat(Vc), p)
        assert errc["B_scaled"] < TOL_FP64 and errc["V_scaled"] < TOL_FP64, (j, errc)
    e.close()


def test_eigen_wide_files_and_errors(pkg, gpu, tmp_path):
    """fgls_eigen with wXR = 2 through files (reader, one pass of 3 traits,
    several blocks) against the oracle; the mixed mode and p > 16 are errors."""
    n, wXL, wXR, m, t = 500, 3, 2, 401, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=21, t=t, wXR=wXR)
    d = str(tmp_path)
    H = np.array([0.3, 0.5, 0.7])
    Sig = np.array([1.0, 1.2, 0.8])
    write_dir(prob, d, H, Sig)
    p = wXL + wXR
    f = lambda s: os.path.join(d, s)
    pkg.fgls_eigen(n, p, m, t, wXL, wXR, 128, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"),
                   f("XR.in"), f("Y.in"), f("B.out"), f("V.out"), options={"block": 128})
    B = np.fromfile(f("B.out"), dtype="<f8")
    V = np.fromfile(f("V.out"), dtype="<f8")
    assert B.size == t * m * p and V.size == t * m * p * p
    Bo, Vo = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR, 128)
    err = orc.parity_errors(B, V, Bo, Vo, p)
    assert err["V_upper_zero"] and err["B_scaled"] < TOL_FP64 and err["V_scaled"] < TOL_FP64, err
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_eigen(n, p, m, t, wXL, wXR, 128, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"),
                       f("XR.in"), f("Y.in"), f("B.out"), f("V.out"), options={"precision": "mixed"})
    assert ei.value.code == pkg.FGLS_ERR_ARG
    e = eig_setup(pkg, prob, H, Sig, max_block=64)
    XR = pkg.colmajor(prob.XR[:, :64 * wXR], ld=(n + 3) // 4 * 4)
    with pytest.raises(pkg.FglsError) as ei:
        e.solve_block_wide(XR, 14)          # p = 3 + 14 > 16
    assert ei.value.code == pkg.FGLS_ERR_ARG
    e.set_mode(pkg.FGLS_MODE_MIXED_I8)
    with pytest.raises(pkg.FglsError) as ei:
        e.solve_block_wide(XR, wXR)
    assert ei.value.code == pkg.FGLS_ERR_STATE
    e.close()
