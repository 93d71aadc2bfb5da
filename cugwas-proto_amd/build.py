"""Build recipe of the package's native code (no CMake; hipcc + gcc).

Outputs (git-ignored, shipped with the tree):
  lib/libfgls_amd.so   HIP kernels + C-ABI (include/fgls_hip.h) + the C host
                       solver fgls_chol_gpu (include/fgls_chol_gpu.h)
  lib/driver.x         the reference's CLI contract (driver.c) over that solver
  lib/driver_san.x     the same driver with the host C code (driver.c,
                       fgls_host.c) built with AddressSanitizer and
                       UndefinedBehaviorSanitizer (gcc, runtimes linked
                       statically); the HIP objects are the regular ones, so
                       only host code is instrumented
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
ARCH = "gfx950"

# All HIP code is compiled without packed FP32 instructions.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def openblas_lib() -> str:
    """LP64 OpenBLAS bundled with scipy (exports scipy_-prefixed LAPACK)."""
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__) + ".libs", "libscipy_openblas*.so"))
    libs = [p for p in libs if "64_" not in os.path.basename(p)]
    if not libs:
        raise RuntimeError("scipy's bundled OpenBLAS not found")
    return sorted(libs)[0]


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return res


def build(verbose: bool = False) -> dict:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    blas = openblas_lib()
    blasdir = os.path.dirname(blas)
    blasname = os.path.basename(blas)

    # one hipcc per translation unit, run concurrently
    hip_units = {
        "fgls_kernels": [],
        # the int8 kernel keeps its MFMA accumulators in VGPRs (see fgls_i8.hip)
        "fgls_i8": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
        "fgls_adjoint": [],
        # the 64 x 64 leaf of the factorisation keeps its row and column in
        # registers: its loops must unroll completely (see fgls_potrf.hip)
        "fgls_potrf": ["-mllvm", "-pragma-unroll-threshold=500000"],
        "fgls_eigen": [],
    }
    hip_objs = []
    cmds = []
    for name, extra in hip_units.items():
        obj = os.path.join(objdir, name + ".o")
        cmds.append([hipcc(), *HIP_FLAGS, *extra, "-I", INCLUDE, "-c", os.path.join(CSRC, name + ".hip"),
                     "-o", obj])
        hip_objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        for f in [pool.submit(run, c, verbose) for c in cmds]:
            f.result()

    host_objs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.c"))):
        if os.path.basename(src) == "driver.c":
            continue
        obj = os.path.join(objdir, os.path.basename(src)[:-2] + ".o")
        run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-fopenmp", "-I", INCLUDE, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
             "-c", src, "-o", obj], verbose)
        host_objs.append(obj)

    so = os.path.join(LIBDIR, "libfgls_amd.so")
    tmp = so + ".tmp"
    run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *hip_objs, *host_objs,
         "-L", blasdir, f"-l:{blasname}", f"-Wl,-rpath,{blasdir}", "-fopenmp", "-lpthread", "-lrt", "-ldl"],
        verbose)
    os.replace(tmp, so)

    out = {"lib": so}
    drv = os.path.join(CSRC, "driver.c")
    if os.path.exists(drv):
        exe = os.path.join(LIBDIR, "driver.x")
        run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INCLUDE, drv, "-o", exe,
             "-L", LIBDIR, "-lfgls_amd", "-Wl,-rpath,$ORIGIN",
             "-Wl,-rpath-link," + blasdir], verbose)
        out["driver"] = exe
        out["driver_san"] = build_sanitized_driver(objdir, hip_objs, blasdir, blasname, verbose)
    return out


# host-only sanitizers: every instrumented object is compiled and linked by
# gcc; the device objects come from the regular hipcc build above
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def build_sanitized_driver(objdir, hip_objs, blasdir, blasname, verbose):
    objs = []
    for name in ("fgls_host", "driver"):
        obj = os.path.join(objdir, name + "_san.o")
        run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-Wall", "-Wno-unused-parameter", *SAN_FLAGS, "-fopenmp",
             "-I", INCLUDE, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
             "-c", os.path.join(CSRC, name + ".c"), "-o", obj], verbose)
        objs.append(obj)
    exe = os.path.join(LIBDIR, "driver_san.x")
    run(["g++", *SAN_FLAGS, "-static-libasan", "-static-libubsan", "-o", exe, *objs, *hip_objs,
         "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
         "-L", blasdir, f"-l:{blasname}", f"-Wl,-rpath,{blasdir}", "-fopenmp", "-lpthread", "-lrt", "-ldl"],
        verbose)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
