"""CPU-side checks of the C-ABI library (no compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(fgls_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_the_boundary():
    names = declared_functions()
    for required in ("fgls_chol_gpu", "fgls_chol_gpu_ex", "fgls_hip_ctx_create", "fgls_hip_set_trait",
                     "fgls_hip_solve_block", "fgls_hip_trsm", "fgls_hip_snp_stage", "fgls_hip_check",
                     "fgls_eigen", "fgls_eigen_gpu_ex", "fgls_hip_eig_create", "fgls_hip_eig_solve_block"):
        assert required in names


def test_host_eigendecomposition(pkg):
    """fgls_eigen_decompose (dsyevr, fgls_eigen.c:780-806) against numpy: the
    reconstruction, orthogonality and ascending order; only the lower triangle
    of Phi is read."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((90, 180))
    Phi = A @ A.T / 180
    Z, W = pkg.eigen_decompose(np.tril(Phi) + np.triu(np.full_like(Phi, 7.0), 1), threads=2)
    assert np.abs(Z @ np.diag(W) @ Z.T - Phi).max() < 1e-13
    assert np.abs(Z.T @ Z - np.eye(90)).max() < 1e-13
    assert np.allclose(W, np.linalg.eigvalsh(Phi), rtol=0, atol=1e-13)
    assert np.all(np.diff(W) >= 0)


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_cleanly(pkg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert pkg.device_count() == 0
    with pytest.raises(pkg.FglsError) as ei:
        pkg.DeviceContext(100, 3, 1, 64)
    assert ei.value.code == pkg.FGLS_ERR_ARG
    assert "device 0 not available" in ei.value.msg


def test_argument_errors(pkg):
    lib = pkg.lib()
    h = ctypes.c_void_p()
    assert lib.fgls_hip_ctx_create(ctypes.byref(h), 0, 3, 1, 64, 0) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_ctx_create(ctypes.byref(h), 100, 3, 0, 64, 0) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_ctx_create(ctypes.byref(h), 100, 15, 2, 64, 0) == pkg.FGLS_ERR_ARG  # p > 16
    assert lib.fgls_hip_solve_block(None, None, 0, 1, None, None, 0, None) == pkg.FGLS_ERR_ARG
    assert b"NULL context" in lib.fgls_hip_last_error()


def test_host_solver_input_errors(pkg, tmp_path):
    """File errors surface before any device work (defined errors, not crashes)."""
    d = str(tmp_path)
    f = lambda s: os.path.join(d, s)
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(10, 4, 5, 1, 3, 1, 5, 1, 1, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"))
    assert ei.value.code == pkg.FGLS_ERR_IO and "Phi.in" in ei.value.msg
    np.zeros(50).tofile(f("Phi.in"))                      # short: 50 of 100 doubles
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(10, 4, 5, 1, 3, 1, 5, 1, 1, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"))
    assert ei.value.code == pkg.FGLS_ERR_IO and "read 50 elements - 100 expected" in ei.value.msg
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(10, 5, 5, 1, 3, 1, 5, 1, 1, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"))
    assert ei.value.code == pkg.FGLS_ERR_ARG                # p != wXL + wXR


def test_driver_usage_and_exit_codes(pkg, tmp_path):
    drv = os.path.join(ROOT, "cugwas-proto_amd", "lib", "driver.x")
    assert os.path.exists(drv)
    r = subprocess.run([drv], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([drv, "e", str(tmp_path), "1"], input="10 4 5 1 3 5 1 1", capture_output=True, text=True)
    assert r.returncode == 1 and "[Error]" in r.stderr and "Phi.in" in r.stderr   # variant e: fgls_eigen
    r = subprocess.run([drv, "g", str(tmp_path), "1"], input="10 4 5 1 3 5 1 1", capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("[Error]: ") and "Phi.in" in r.stderr


def test_new_entry_points_reject_bad_arguments(pkg):
    """NULL contexts and pointers are defined errors (no device needed)."""
    lib = pkg.lib()
    assert lib.fgls_hip_set_trait_phi(None, None, 0, 0.5, 1.0, None, 0, None, None) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_get_factor(None, None, 0, None) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_adjoint_begin(None, None, 0, 1.0, None) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_adjoint_block(None, None, 0, 1, None, 0, None) == pkg.FGLS_ERR_ARG
    assert lib.fgls_hip_adjoint_end(None, None, None, None, None) == pkg.FGLS_ERR_ARG
    assert b"NULL context" in lib.fgls_hip_last_error()


def test_host_factor_matches_lapack(pkg):
    """fgls_factor_M (host M construction + dpotrf, fgls_chol_gpu.c:406-430)
    against scipy's Cholesky of the oracle's M, and its info on a non-SPD M."""
    from oracle import fgls_oracle as orc
    from tests.fgls_data import make_phi
    rng = np.random.default_rng(3)
    Phi = make_phi(rng, 150)
    L = pkg.factor_M(Phi, 0.35, 1.7, threads=2)
    Lref = orc.cholesky_lower(orc.build_M(Phi, 0.35, 1.7))
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) < 1e-14
    Phi_bad = np.eye(50)
    Phi_bad[20, 20] = -5.0
    with pytest.raises(pkg.FglsError) as ei:
        pkg.factor_M(Phi_bad, 0.9, 1.0)
    assert ei.value.code == pkg.FGLS_ERR_NOTSPD and "info: 21" in ei.value.msg


def test_pack_u2_layout(pkg):
    """The 2-bit layout of xr_format=u2: individual r in bits 2(r%4)..2(r%4)+1
    of byte r/4, first individual in the low bits, columns padded to whole
    bytes; anything but 0, 1, 2 is rejected."""
    X = np.array([[0, 2], [1, 2], [2, 0], [1, 1], [2, 0]])          # n = 5, m = 2
    P = pkg.pack_u2(X)
    assert P.dtype == np.uint8 and P.shape == (2, 2)
    assert P[0, 0] == (0 | 1 << 2 | 2 << 4 | 1 << 6) and P[0, 1] == 2
    assert P[1, 0] == (2 | 2 << 2 | 0 << 4 | 1 << 6) and P[1, 1] == 0
    for bad in (np.array([[3]]), np.array([[-1]]), np.array([[0.5]])):
        with pytest.raises(pkg.FglsError):
            pkg.pack_u2(bad)


def test_pack_i16_layout(pkg):
    """The 16-bit dosage layout of xr_format=i16: SNP-major int16 v = rint(1000 x);
    values that are not multiples of 1/1000, or out of range, are rejected."""
    X = np.array([[0.0, 1.999], [0.5, 0.001], [2.0, 1.234]])
    P = pkg.pack_i16(X, 1000.0)
    assert P.dtype == np.int16 and P.shape == (2, 3)
    assert P.tolist() == [[0, 500, 2000], [1999, 1, 1234]]
    for bad in (np.array([[0.0005]]), np.array([[17.0]]), np.array([[1.234 + 1e-11]])):
        with pytest.raises(pkg.FglsError):
            pkg.pack_i16(bad, 1000.0)
    # every 3-decimal value in [0, 2] passes (x * 1000 is within 2 eps of v)
    allv = np.arange(2001) / 1000.0
    assert np.array_equal(pkg.pack_i16(allv, 1000.0)[0], np.arange(2001))
