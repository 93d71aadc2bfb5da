"""The host C code under AddressSanitizer and UndefinedBehaviorSanitizer.

lib/driver_san.x is the CLI driver with driver.c and fgls_host.c (option
parsing, file reads, the reader threads, pinned slots, the block and trait
loops, output writes) built by gcc with -fsanitize=address,undefined and
-fno-sanitize-recover=all, linked to the regular HIP objects.  Each run must
end without a sanitizer report, and its B.out and V.out must equal those of
the regular driver.x bitwise: the reference fixture, the compact X_R files,
the host factor, the serial and multi-threaded readers, several traits per
pass and the eigen variant; a truncated X_R file, a missing input file and a
singular per-SNP system must end in their defined errors.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.fgls_data import make_problem, write_problem_dir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cugwas-proto_amd", "lib")
DRIVER = os.path.join(LIB, "driver.x")
DRIVER_SAN = os.path.join(LIB, "driver_san.x")
# leaks are not checked: the HIP runtime keeps its allocations until exit
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:verify_asan_link_order=0",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}
SAN_MARKS = ("ERROR: AddressSanitizer", "runtime error:", "UndefinedBehaviorSanitizer")


def run(exe, d, stdin, variant="g", options=None):
    env = dict(os.environ)
    env.update(SAN_ENV)
    if options:
        env["FGLS_OPTIONS"] = options
    else:
        env.pop("FGLS_OPTIONS", None)
    return subprocess.run([exe, variant, d, "1"], input=stdin, capture_output=True, text=True, env=env,
                          timeout=300)


def clean(res):
    return not any(m in res.stderr for m in SAN_MARKS)


@pytest.fixture(scope="module")
def san():
    if not os.path.exists(DRIVER_SAN):
        pytest.fail("lib/driver_san.x is not built (cugwas-proto_amd/build.py)")
    # the executable carries the sanitizer runtimes (linked statically)
    with open(DRIVER_SAN, "rb") as fh:
        image = fh.read()
    assert b"AddressSanitizer" in image and b"UndefinedBehaviorSanitizer" in image
    return DRIVER_SAN


def both(d, stdin, variant="g", options=None):
    """Run the sanitized and the regular driver on d; return their outputs."""
    outs = []
    for exe in (DRIVER_SAN, DRIVER):
        print(f"{os.path.basename(exe)} {variant} options={options}", flush=True)
        res = run(exe, d, stdin, variant, options)
        assert res.returncode == 0, res.stderr[-3000:]
        assert clean(res), res.stderr[-3000:]
        outs.append((np.fromfile(os.path.join(d, "B.out")), np.fromfile(os.path.join(d, "V.out"))))
    return outs


def test_sanitized_driver_fixture(gpu, san, tmp_path):
    from tests.golden.test_xl_fixture import write_dir
    d = str(tmp_path / "test_XL")
    write_dir(d)
    stdin = open(os.path.join(d, "input.in")).read()
    (Bs, Vs), (B, V) = both(d, stdin)
    assert np.array_equal(Bs, B) and np.array_equal(Vs, V)


@pytest.mark.parametrize("variant,options", [
    ("g", None),
    ("g", "precision=mixed"),
    ("g", "potrf=host"),
    ("g", "serial=1,block=128"),
    ("g", "io_threads=3,block=192,traits_per_pass=1"),
    ("g", "xr_format=i8,precision=mixed"),
    ("g", "xr_format=u2,precision=mixed,block=256"),
    ("g", "xr_format=u2"),
    ("e", None),
    ("e", "precision=mixed,xr_format=u2"),
])
def test_sanitized_driver_options(gpu, san, tmp_path, variant, options):
    n, wXL, m, t, x_b = 300, 3, 701, 3, 128
    prob = make_problem(n=n, wXL=wXL, m=m, seed=17, t=t)
    d = str(tmp_path)
    write_problem_dir(prob, d, H=np.array([0.3, 0.55, 0.8]), Sig=np.array([1.0, 1.5, 0.7]))
    fmt = dict(kv.split("=") for kv in (options or "").split(",") if kv).get("xr_format", "f64")
    if fmt != "f64":
        # the compact file replaces XR.in under the same name
        X8 = prob.XR.astype(np.int8)
        if fmt == "i8":
            X8.ravel(order="F").tofile(os.path.join(d, "XR.in"))
        else:
            nb = (n + 3) // 4
            q = np.zeros((m, 4 * nb), dtype=np.uint8)
            q[:, :n] = X8.T
            q = q.reshape(m, nb, 4)
            (q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)).astype(np.uint8).tofile(
                os.path.join(d, "XR.in"))
    stdin = f"{n} {wXL + 1} {m} {t} {wXL} {x_b} 1 4"
    (Bs, Vs), (B, V) = both(d, stdin, variant, options)
    assert Bs.size == t * m * (wXL + 1)
    assert np.array_equal(Bs, B) and np.array_equal(Vs, V)


def test_sanitized_driver_error_paths(gpu, san, tmp_path):
    """Defined errors under the sanitizers: a truncated X_R file, a missing
    input file, a singular per-SNP system."""
    n, wXL, m = 200, 3, 300
    prob = make_problem(n=n, wXL=wXL, m=m, seed=5)
    d = str(tmp_path)
    write_problem_dir(prob, d)
    stdin = f"{n} {wXL + 1} {m} 1 {wXL} 64 1 2"
    xr = os.path.join(d, "XR.in")
    data = open(xr, "rb").read()
    open(xr, "wb").write(data[: len(data) - 8 * n * 3])
    res = run(DRIVER_SAN, d, stdin)
    assert res.returncode == 1 and "not read completely" in res.stderr, res.stderr[-2000:]
    assert clean(res), res.stderr[-3000:]
    open(xr, "wb").write(data)
    os.rename(os.path.join(d, "Phi.in"), os.path.join(d, "Phi.bak"))
    res = run(DRIVER_SAN, d, stdin)
    assert res.returncode == 1 and "Phi.in" in res.stderr, res.stderr[-2000:]
    assert clean(res), res.stderr[-3000:]
    os.rename(os.path.join(d, "Phi.bak"), os.path.join(d, "Phi.in"))
    X = prob.XR.copy()
    X[:, 17] = 0.0
    np.asarray(X.ravel(order="F"), dtype="<f8").tofile(xr)
    res = run(DRIVER_SAN, d, stdin)
    assert res.returncode == 1 and "dpotrf(V) failed" in res.stderr, res.stderr[-2000:]
    assert clean(res), res.stderr[-3000:]
