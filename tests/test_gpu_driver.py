"""The drop-in boundary on the GPU: the C host solver fgls_chol_gpu and the
reference CLI contract (driver.x), files in and files out."""
import os
import subprocess

import numpy as np
import pytest

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem, write_problem_dir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cugwas-proto_amd", "lib", "driver.x")


def run_driver(d, stdin, variant="g", extra=(), env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([DRIVER, variant, d, "1", *extra], input=stdin, capture_output=True,
                          text=True, env=e, timeout=300)


def read_out(d, name, count):
    return np.fromfile(os.path.join(d, name), dtype="<f8", count=count)


def test_driver_reference_fixture(gpu, tmp_path):
    """test_cases/test_XL end to end through the CLI (t=20 traits, x_b=2000)."""
    from tests.golden.test_xl_fixture import write_dir
    d = str(tmp_path / "test_XL")
    prm = write_dir(d)
    stdin = open(os.path.join(d, "input.in")).read()
    res = run_driver(d, stdin, extra=("--check",))
    assert res.returncode == 0, res.stderr
    lines = res.stdout.strip().splitlines()
    assert lines[-1].strip().isdigit()          # mean ms per repetition is the last line
    n, p, m, t = prm["n"], prm["p"], prm["m"], prm["t"]
    B = read_out(d, "B.out", t * m * p)
    Bexp = read_out(d, "B_exp.out", t * m * p)
    assert orc.max_rel_err(B, Bexp) < 1e-10
    # V against the oracle on the same synthetic X_R
    Phi, H, Sig, XL, XR, Y = orc.read_inputs(d, n, p, m, t, prm["wXL"])
    Bo, Vo = orc.fgls_chol(Phi, H, Sig, XL, XR, Y, wXR=1, x_b=prm["x_b"])
    V = read_out(d, "V.out", t * m * p * p)
    e = orc.parity_errors(B, V, Bo, Vo, p)
    assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10 and e["V_upper_zero"], e


@pytest.mark.parametrize("n,wXL,m,t,x_b,wXR", [(1000, 3, 10000, 1, 256, 1),
                                                 (333, 2, 517, 3, 100, 1),
                                                 (200, 3, 64, 2, 7, 2)])
def test_fgls_chol_gpu_matches_oracle(pkg, gpu, tmp_path, n, wXL, m, t, x_b, wXR):
    prob = make_problem(n=n, wXL=wXL, m=m, seed=n + m + t, t=t, wXR=wXR)
    d = str(tmp_path)
    H = np.linspace(0.3, 0.7, t)
    Sig = np.linspace(0.8, 1.5, t)
    write_problem_dir(prob, d, H=H, Sig=Sig)
    p = wXL + wXR
    f = lambda s: os.path.join(d, s)
    pkg.fgls_chol_gpu(n, p, m, t, wXL, wXR, x_b, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                      f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"))
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=wXR)
    B = read_out(d, "B.out", t * m * p)
    V = read_out(d, "V.out", t * m * p * p)
    assert B.size == Bo.size and V.size == Vo.size
    e = orc.parity_errors(B, V, Bo, Vo, p)
    assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10 and e["V_upper_zero"], e
    # block size, the serialised debug mode and the number of traits sharing
    # one read of X_R do not change a single bit
    for opts in ({"block": x_b}, {"serial": 1, "block": 3 * x_b}, {"traits_per_pass": 1},
                 {"traits_per_pass": 2, "block": 2 * x_b}):
        pkg.fgls_chol_gpu(n, p, m, t, wXL, wXR, x_b, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B2.out"), f("V2.out"), options=opts)
        assert np.array_equal(read_out(d, "B2.out", t * m * p), B)
        assert np.array_equal(read_out(d, "V2.out", t * m * p * p), V)


def test_driver_errors(gpu, tmp_path):
    """Defined errors instead of the reference's crashes (common.c:11-17)."""
    prob = make_problem(n=128, wXL=3, m=40, seed=3)
    d = str(tmp_path)
    write_problem_dir(prob, d)
    stdin = "128 4 40 1 3 16 1 1"
    ok = run_driver(d, stdin)
    assert ok.returncode == 0, ok.stderr
    # truncated X_R: the last block cannot be read completely
    xr = os.path.join(d, "XR.in")
    data = open(xr, "rb").read()
    open(xr, "wb").write(data[: len(data) - 8 * 128 * 3])
    res = run_driver(d, stdin)
    assert res.returncode == 1 and "[Error]:" in res.stderr and "not read completely" in res.stderr
    open(xr, "wb").write(data)
    # missing file
    os.rename(os.path.join(d, "Phi.in"), os.path.join(d, "Phi.bak"))
    res = run_driver(d, stdin)
    assert res.returncode == 1 and "[Error]:" in res.stderr and "Phi.in" in res.stderr
    os.rename(os.path.join(d, "Phi.bak"), os.path.join(d, "Phi.in"))
    # singular per-SNP system: an all-zero SNP column
    X = prob.XR.copy()
    X[:, 17] = 0.0
    np.asarray(X.ravel(order="F"), dtype="<f8").tofile(xr)
    res = run_driver(d, stdin)
    assert res.returncode == 1 and "dpotrf(V) failed (info: 4) - i: 17" in res.stderr
    # M not positive definite: h2 = 1 and a duplicated individual makes Phi singular
    Phi = prob.Phi.copy()
    Phi[:, 1] = Phi[:, 0]
    Phi[1, :] = Phi[0, :]
    np.asarray(Phi.ravel(order="F"), dtype="<f8").tofile(os.path.join(d, "Phi.in"))
    np.asarray([1.0], dtype="<f8").tofile(os.path.join(d, "H.in"))
    res = run_driver(d, stdin)
    assert res.returncode == 1 and "dpotrf(M) failed" in res.stderr


@pytest.mark.parametrize("n,m,t,x_b", [(1000, 5000, 1, 256), (333, 517, 2, 100)])
def test_fgls_chol_gpu_mixed_precision(pkg, gpu, tmp_path, n, m, t, x_b):
    """precision=mixed through the file pipeline: int8 genotypes on the
    device, same files out; within the mixed path's tolerance of the oracle
    (b relative to max(|b|, se) 2e-11, var(b) element-wise 1e-12)."""
    print(f"fgls_chol_gpu precision=mixed n={n} m={m} t={t}", flush=True)
    wXL = 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=n + m + 7, t=t)
    d = str(tmp_path)
    H = np.linspace(0.35, 0.65, t)
    Sig = np.linspace(0.9, 1.3, t)
    write_problem_dir(prob, d, H=H, Sig=Sig)
    p = wXL + 1
    f = lambda s: os.path.join(d, s)
    pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, x_b, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                      f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"),
                      options={"precision": "mixed", "block": 2 * x_b})
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=1)
    B = read_out(d, "B.out", t * m * p)
    V = read_out(d, "V.out", t * m * p * p)
    e = orc.parity_errors(B, V, Bo, Vo, p)
    assert e["B_scaled"] < 2e-11 and e["var_rel"] < 1e-12 and e["V_upper_zero"], e
    # slices=6: the six-digit form of L^-1, to the same bar
    pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, x_b, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                      f("XL.in"), f("XR.in"), f("Y.in"), f("B6.out"), f("V6.out"),
                      options={"precision": "mixed", "block": 2 * x_b, "slices": 6})
    e6 = orc.parity_errors(read_out(d, "B6.out", t * m * p), read_out(d, "V6.out", t * m * p * p), Bo, Vo, p)
    assert e6["B_scaled"] < 2e-11 and e6["var_rel"] < 1e-12 and e6["V_upper_zero"], e6
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, x_b, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B6.out"), f("V6.out"),
                          options={"precision": "mixed", "slices": 7})
    assert ei.value.code == pkg.FGLS_ERR_ARG and "slices" in ei.value.msg


def test_fgls_chol_gpu_mixed_rejects_dosages(pkg, gpu, tmp_path):
    """Imputed (non-integer) dosages cannot take the int8 path: a defined error."""
    prob = make_problem(n=200, wXL=3, m=300, seed=4, dosage="imputed")
    d = str(tmp_path)
    write_problem_dir(prob, d, H=np.array([0.5]), Sig=np.array([1.0]))
    f = lambda s: os.path.join(d, s)
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(200, 4, 300, 1, 3, 1, 100, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"),
                          options={"precision": "mixed"})
    assert ei.value.code == pkg.FGLS_ERR_ARG and "precision=fp64" in ei.value.msg


def test_fgls_chol_gpu_host_and_device_factor_agree(pkg, gpu, tmp_path):
    """potrf=host (the reference's host dpotrf) and the default device
    factorisation give the same files to the parity bar."""
    n, m, t, wXL = 500, 700, 2, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=77, t=t)
    d = str(tmp_path)
    H, Sig = np.array([0.3, 0.65]), np.array([1.0, 2.0])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    p = wXL + 1
    f = lambda s: os.path.join(d, s)
    out = {}
    for mode in ("gpu", "host"):
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 128, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f(f"B_{mode}.out"), f(f"V_{mode}.out"),
                          options={"potrf": mode})
        out[mode] = (read_out(d, f"B_{mode}.out", t * m * p), read_out(d, f"V_{mode}.out", t * m * p * p))
    e = orc.parity_errors(out["gpu"][0], out["gpu"][1], out["host"][0], out["host"][1], p)
    assert e["B_scaled"] < 1e-11 and e["V_scaled"] < 1e-11 and e["V_upper_zero"], e
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=1)
    e = orc.parity_errors(out["gpu"][0], out["gpu"][1], Bo, Vo, p)
    assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10, e


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_int8_genotype_file(pkg, gpu, tmp_path, precision):
    """xr_format=i8 (compact X_R file, 1 byte per value) gives exactly the
    files of the FP64 X_R file with the same integer genotypes."""
    print(f"int8 X_R file, precision={precision}", flush=True)
    n, m, t, wXL = 700, 1100, 2, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=21, t=t)
    d = str(tmp_path)
    H, Sig = np.array([0.4, 0.6]), np.array([1.0, 1.5])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    f = lambda s: os.path.join(d, s)
    prob.XR.astype(np.int8).ravel(order="F").tofile(f("XR8.in"))
    p = wXL + 1
    out = {}
    for fmt, xr in (("f64", "XR.in"), ("i8", "XR8.in")):
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f(xr), f("Y.in"), f(f"B_{fmt}.out"), f(f"V_{fmt}.out"),
                          options={"precision": precision, "xr_format": fmt, "block": 512})
        out[fmt] = (read_out(d, f"B_{fmt}.out", t * m * p), read_out(d, f"V_{fmt}.out", t * m * p * p))
    assert np.array_equal(out["i8"][0], out["f64"][0])
    assert np.array_equal(out["i8"][1], out["f64"][1])
    # a short int8 file is a defined error
    np.zeros(n * (m - 1), dtype=np.int8).tofile(f("XR8s.in"))
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(n, p, m, 1, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR8s.in"), f("Y.in"), f("B_s.out"), f("V_s.out"),
                          options={"precision": precision, "xr_format": "i8"})
    assert ei.value.code == pkg.FGLS_ERR_IO and "not read completely" in ei.value.msg
    # -128 is outside the int8 product's range: a defined error on the mixed
    # path; the FP64 path converts it exactly and compares with the oracle
    X = prob.XR.astype(np.int8)
    X[12, 1033] = -128
    X.ravel(order="F").tofile(f("XR8neg.in"))
    opts = {"precision": precision, "xr_format": "i8", "block": 512}
    if precision == "mixed":
        with pytest.raises(pkg.FglsError) as ei:
            pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                              f("XL.in"), f("XR8neg.in"), f("Y.in"), f("B_n.out"), f("V_n.out"), options=opts)
        assert ei.value.code == pkg.FGLS_ERR_ARG and "1 genotype values of SNPs 1024" in ei.value.msg
    else:
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR8neg.in"), f("Y.in"), f("B_n.out"), f("V_n.out"), options=opts)
        XR = prob.XR.copy()
        XR[12, 1033] = -128.0
        Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, XR, prob.Y, wXR=1)
        e = orc.parity_errors(read_out(d, "B_n.out", t * m * p), read_out(d, "V_n.out", t * m * p * p),
                              Bo, Vo, p)
        assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10, e


def test_fgls_chol_gpu_mixed_dosages(pkg, gpu, tmp_path):
    """precision=mixed with dosage_scale=1000 on imputed dosages (3 decimals)
    through the file pipeline, against the oracle."""
    print("fgls_chol_gpu precision=mixed dosage_scale=1000", flush=True)
    n, m, t, wXL = 900, 1300, 2, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=99, t=t, dosage="imputed")
    d = str(tmp_path)
    H, Sig = np.array([0.3, 0.55]), np.array([1.0, 0.8])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    p = wXL + 1
    f = lambda s: os.path.join(d, s)
    pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                      f("XL.in"), f("XR.in"), f("Y.in"), f("B.out"), f("V.out"),
                      options={"precision": "mixed", "dosage_scale": 1000, "block": 512})
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=1)
    e = orc.parity_errors(read_out(d, "B.out", t * m * p), read_out(d, "V.out", t * m * p * p), Bo, Vo, p)
    assert e["B_scaled"] < 2e-11 and e["var_rel"] < 1e-12 and e["V_upper_zero"], e
    with pytest.raises(pkg.FglsError) as ei:      # the same data as genotypes: not int8-exact
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR.in"), f("Y.in"), f("B2.out"), f("V2.out"),
                          options={"precision": "mixed"})
    assert ei.value.code == pkg.FGLS_ERR_ARG

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_u2_genotype_file(pkg, gpu, tmp_path, precision):
    """xr_format=u2 (2-bit X_R file, ceil(n/4) bytes per SNP) gives exactly the
    files of the int8 X_R file with the same genotypes (n not a multiple of 4,
    ragged last block); a code 3 is a defined error."""
    print(f"2-bit X_R file, precision={precision}", flush=True)
    n, m, t, wXL = 701, 1100, 2, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=23, t=t)
    d = str(tmp_path)
    H, Sig = np.array([0.4, 0.6]), np.array([1.0, 1.5])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    f = lambda s: os.path.join(d, s)
    prob.XR.astype(np.int8).ravel(order="F").tofile(f("XR8.in"))
    P = pkg.pack_u2(prob.XR)
    assert P.shape == (m, (n + 3) // 4)
    P.tofile(f("XR2.in"))
    p = wXL + 1
    out = {}
    for fmt, xr in (("i8", "XR8.in"), ("u2", "XR2.in")):
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f(xr), f("Y.in"), f(f"B_{fmt}.out"), f(f"V_{fmt}.out"),
                          options={"precision": precision, "xr_format": fmt, "block": 512})
        out[fmt] = (read_out(d, f"B_{fmt}.out", t * m * p), read_out(d, f"V_{fmt}.out", t * m * p * p))
    assert np.array_equal(out["u2"][0], out["i8"][0])
    assert np.array_equal(out["u2"][1], out["i8"][1])
    # the eigen variant reads the same file
    pkg.fgls_eigen(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"),
                   f("XR2.in"), f("Y.in"), f("Be.out"), f("Ve.out"),
                   options={"precision": precision, "xr_format": "u2"})
    Bo, Vo = orc.fgls_eigen(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y)
    e = orc.parity_errors(read_out(d, "Be.out", t * m * p), read_out(d, "Ve.out", t * m * p * p), Bo, Vo, p)
    assert e["B_scaled"] < 2e-11 and e["V_scaled"] < 2e-11, e
    # a code 3 (SNP 900, individual 10: byte 2, bits 4-5) fails the run with a defined error
    P[900, 2] |= 0x30
    P.tofile(f("XR2bad.in"))
    with pytest.raises(pkg.FglsError) as ei:
        pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f("XR2bad.in"), f("Y.in"), f("B.out"), f("V.out"),
                          options={"precision": precision, "xr_format": "u2", "block": 512})
    assert ei.value.code == pkg.FGLS_ERR_ARG and "1 2-bit codes" in ei.value.msg


def test_u2_unpack_kernel(pkg, gpu):
    """fgls_hip_u2_to_i8 against the packed values, every padding case of n % 4."""
    import torch
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 5, 63, 64, 65, 1001):
        X = rng.integers(0, 3, size=(n, 37))
        P = torch.from_numpy(pkg.pack_u2(X)).cuda()
        X8 = pkg.u2_to_i8(P, n)
        got = X8.cpu().numpy()
        assert np.array_equal(got[:, :n].T, X.astype(np.int8))
        assert not got[:, n:].any()


@pytest.mark.parametrize("variant", ["chol", "eigen"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pipeline_edge_sizes(pkg, gpu, tmp_path, variant, precision):
    """m = 0 writes empty B.out / V.out; one SNP; a tiny n (7 < 16) with one
    covariate; both variants and precisions against the oracle."""
    run = pkg.fgls_chol_gpu if variant == "chol" else pkg.fgls_eigen
    f = lambda d, s: os.path.join(d, s)
    for n, wXL, m, t in ((40, 2, 0, 1), (40, 2, 1, 2), (7, 1, 9, 1)):
        prob = make_problem(n=n, wXL=wXL, m=max(m, 1), seed=n + m, t=t)
        d = str(tmp_path / f"{variant}_{precision}_{n}_{m}")
        os.makedirs(d)
        H, Sig = np.linspace(0.3, 0.6, t), np.ones(t)
        write_problem_dir(prob, d, H=H, Sig=Sig)
        if m == 0:
            open(f(d, "XR.in"), "wb").close()
        p = wXL + 1
        run(n, p, m, t, wXL, 1, 256, 1, 2, f(d, "Phi.in"), f(d, "H.in"), f(d, "Sig.in"), f(d, "XL.in"),
            f(d, "XR.in"), f(d, "Y.in"), f(d, "B.out"), f(d, "V.out"), options={"precision": precision})
        B = np.fromfile(f(d, "B.out"), dtype="<f8")
        V = np.fromfile(f(d, "V.out"), dtype="<f8")
        assert B.size == t * m * p and V.size == t * m * p * p
        if m:
            Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR[:, :m], prob.Y)
            e = orc.parity_errors(B, V, Bo, Vo, p)
            assert e["B_scaled"] < 1e-10 and e["V_scaled"] < 1e-10 and e["V_upper_zero"], (n, m, e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_i16_dosage_file(pkg, gpu, tmp_path, precision):
    """xr_format=i16 (16-bit dosage file, v = 1000 x) gives exactly the files
    of the FP64 dosage file (x = round(U(0, 2), 3)), for both variants; a value
    out of range is a defined error."""
    print(f"16-bit dosage file, precision={precision}", flush=True)
    n, m, t, wXL = 333, 700, 2, 3
    prob = make_problem(n=n, wXL=wXL, m=m, seed=29, t=t, dosage="imputed")
    d = str(tmp_path)
    H, Sig = np.array([0.3, 0.55]), np.array([1.0, 1.2])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    f = lambda s: os.path.join(d, s)
    pkg.pack_i16(prob.XR, 1000.0).tofile(f("XR16.in"))
    assert os.path.getsize(f("XR16.in")) == 2 * n * m
    p = wXL + 1
    base = {"precision": precision, "block": 256}
    if precision == "mixed":
        base["dosage_scale"] = 1000
    for run in (pkg.fgls_chol_gpu, pkg.fgls_eigen):
        out = {}
        for fmt, xr in (("f64", "XR.in"), ("i16", "XR16.in")):
            opts = dict(base, xr_format=fmt)
            if fmt == "i16":
                opts["dosage_scale"] = 1000
            run(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"), f(xr),
                f("Y.in"), f(f"B_{fmt}.out"), f(f"V_{fmt}.out"), options=opts)
            out[fmt] = (read_out(d, f"B_{fmt}.out", t * m * p), read_out(d, f"V_{fmt}.out", t * m * p * p))
        assert np.array_equal(out["i16"][0], out["f64"][0]), run.__name__
        assert np.array_equal(out["i16"][1], out["f64"][1]), run.__name__
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y)
    e = orc.parity_errors(out["i16"][0], out["i16"][1], Bo, Vo, p)
    assert e["B_scaled"] < 2e-11 and e["V_scaled"] < 2e-11, e
    if precision == "mixed":
        bad = np.fromfile(f("XR16.in"), dtype=np.int16)
        bad[5 * n + 7] = 20000
        bad.tofile(f("XR16bad.in"))
        with pytest.raises(pkg.FglsError) as ei:
            pkg.fgls_chol_gpu(n, p, m, t, wXL, 1, 256, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"),
                              f("XR16bad.in"), f("Y.in"), f("B.out"), f("V.out"),
                              options=dict(base, xr_format="i16", dosage_scale=1000))
        assert ei.value.code == pkg.FGLS_ERR_ARG and "outside" in ei.value.msg


def test_fgls_chol_gpu_mixed_several_columns_per_snp(pkg, gpu, tmp_path):
    """precision=mixed with wXR = 2 through the file pipeline: the FP64, int8
    and 2-bit X_R files give bitwise the same files, within the mixed bound
    of the oracle; dosages and the eigen variant with wXR > 1 in the mixed
    mode are defined errors."""
    print("fgls_chol_gpu precision=mixed wXR=2", flush=True)
    n, m, t, wXL, wXR = 611, 700, 2, 3, 2
    prob = make_problem(n=n, wXL=wXL, m=m, seed=41, t=t, wXR=wXR)
    d = str(tmp_path)
    H, Sig = np.array([0.45, 0.6]), np.array([1.0, 1.4])
    write_problem_dir(prob, d, H=H, Sig=Sig)
    f = lambda s: os.path.join(d, s)
    prob.XR.astype(np.int8).ravel(order="F").tofile(f("XR8.in"))
    pkg.pack_u2(prob.XR).tofile(f("XR2.in"))
    p = wXL + wXR
    out = {}
    for fmt, xr in (("f64", "XR.in"), ("i8", "XR8.in"), ("u2", "XR2.in")):
        pkg.fgls_chol_gpu(n, p, m, t, wXL, wXR, 128, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"),
                          f("XL.in"), f(xr), f("Y.in"), f(f"B_{fmt}.out"), f(f"V_{fmt}.out"),
                          options={"precision": "mixed", "xr_format": fmt, "block": 256})
        out[fmt] = (read_out(d, f"B_{fmt}.out", t * m * p), read_out(d, f"V_{fmt}.out", t * m * p * p))
    for fmt in ("i8", "u2"):
        assert np.array_equal(out[fmt][0], out["f64"][0]) and np.array_equal(out[fmt][1], out["f64"][1]), fmt
    Bo, Vo = orc.fgls_chol(prob.Phi, H, Sig, prob.XL, prob.XR, prob.Y, wXR=wXR)
    e = orc.parity_errors(out["f64"][0], out["f64"][1], Bo, Vo, p)
    assert e["B_scaled"] < 2e-11 and e["V_scaled"] < 2e-11 and e["var_rel"] < 1e-12 and e["V_upper_zero"], e
    for call, opts in ((pkg.fgls_chol_gpu, {"precision": "mixed", "dosage_scale": 1000}),
                       (pkg.fgls_eigen, {"precision": "mixed"})):
        with pytest.raises(pkg.FglsError) as ei:
            call(n, p, m, t, wXL, wXR, 128, 1, 4, f("Phi.in"), f("H.in"), f("Sig.in"), f("XL.in"),
                 f("XR.in"), f("Y.in"), f("Be.out"), f("Ve.out"), options=opts)
        assert ei.value.code == pkg.FGLS_ERR_ARG and "wXR = 1" in ei.value.msg
