"""ctypes runner of the reference's own CPU solver built by oracle/Makefile
(oracle/_ref/libfgls_ref.so) -- TEST INFRASTRUCTURE ONLY (checker and CPU
baseline; never on the product path).

fgls_chol (src/fgls_chol.c:29-544) reads and writes the driver's files; the
reference exits the process on any error (common.c:11-17), so callers pass
valid inputs only.
"""
from __future__ import annotations

import ctypes
import glob
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libfgls_ref.so")

_lib = None
_blas = None


def available() -> bool:
    return os.path.exists(REF_LIB)


def _load():
    global _lib, _blas
    if _lib is None:
        _lib = ctypes.CDLL(REF_LIB)
        _lib.fgls_chol.restype = ctypes.c_int
        _lib.fgls_chol.argtypes = [ctypes.c_int] * 9 + [ctypes.c_char_p] * 8
        import scipy
        libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__) + ".libs", "libscipy_openblas-*.so"))
        _blas = ctypes.CDLL(sorted(libs)[0]) if libs else None
    return _lib


def set_blas_threads(n: int):
    """The reference's setenv of *_NUM_THREADS happens after the BLAS has
    initialised and has no effect; the thread count is set explicitly."""
    _load()
    if _blas is not None:
        _blas.scipy_openblas_set_num_threads(int(n))


def fgls_chol_dir(d: str, n: int, p: int, m: int, t: int, wXL: int, x_b: int, threads: int = 1,
                  out_prefix: str = "") -> float:
    """Run the reference fgls_chol on directory d (driver.c:56-66 file names).
    Returns the wall time in seconds of the call."""
    lib = _load()
    set_blas_threads(threads)
    f = lambda s: os.path.join(d, s).encode()
    t0 = time.perf_counter()
    lib.fgls_chol(n, p, m, t, wXL, p - wXL, x_b, 1, threads, f("Phi.in"), f("H.in"), f("Sig.in"),
                  f("XL.in"), f("XR.in"), f("Y.in"), f(out_prefix + "B.out"), f(out_prefix + "V.out"))
    return time.perf_counter() - t0


def fgls_eigen_dir(d: str, n: int, p: int, m: int, t: int, wXL: int, x_b: int, threads: int = 1,
                   out_prefix: str = "") -> float:
    """Run the reference's eigenbasis solver fgls_eigen (src/fgls_eigen.c:37-236) on
    directory d.  It writes its temporaries XL.in.tmp, XR.in.tmp and Y.in.tmp
    next to the inputs (common.c:98-100).  Returns the wall time in seconds."""
    lib = _load()
    lib.fgls_eigen.restype = ctypes.c_int
    lib.fgls_eigen.argtypes = [ctypes.c_int] * 9 + [ctypes.c_char_p] * 8
    set_blas_threads(threads)
    f = lambda s: os.path.join(d, s).encode()
    t0 = time.perf_counter()
    lib.fgls_eigen(n, p, m, t, wXL, p - wXL, x_b, 1, threads, f("Phi.in"), f("H.in"), f("Sig.in"),
                   f("XL.in"), f("XR.in"), f("Y.in"), f(out_prefix + "B.out"), f(out_prefix + "V.out"))
    return time.perf_counter() - t0
