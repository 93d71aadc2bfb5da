"""Python host mirror of the reference's FGLS interface over the C-ABI.

The native library (lib/libfgls_amd.so, built by build.py) is required: there
is no CPU fallback, and every entry point raises ``FglsError`` when the HIP
library or a GPU is missing.

Two levels are exposed:

* ``fgls_chol_gpu(n, p, m, t, wXL, wXR, x_b, y_b, num_threads, Phi_path, ...)``
  -- the reference solver's 17-argument entry point (fgls_chol_gpu.h:4-10),
  files in, files out, running the C host pipeline of the library.
* ``DeviceContext`` -- the device layer (include/fgls_hip.h) on torch CUDA
  tensors: per-trait precompute, fused block solve, TRSM, TRMM and the
  per-SNP stage.
* ``fgls_eigen(...)`` and ``EigenContext`` -- the same for the eigenbasis
  variant (fgls_eigen.h:68-74): one eigendecomposition of Phi, one product
  Z' X_R per block for all traits.

Matrix convention on tensors: a column-major n x k matrix with leading
dimension ld is a contiguous tensor of shape (k, ld) (row c = column c).
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libfgls_amd.so")

FGLS_OK = 0
FGLS_ERR_ARG = -1
FGLS_ERR_HIP = -2
FGLS_ERR_NOMEM = -3
FGLS_ERR_NOTSPD = -4
FGLS_ERR_IO = -5
FGLS_ERR_STATE = -6

FGLS_MODE_FP64 = 0
FGLS_MODE_MIXED_I8 = 1
# digit form of the slices of L^-1 in the mixed mode (fgls_hip_ctx_set_slice_form)
FGLS_SLICES_AUTO = 0
FGLS_SLICES_SIX = 6

_lib = None


class RunStats(ctypes.Structure):
    """fgls_run_stats (include/fgls_chol_gpu.h)."""
    _fields_ = [("total_s", ctypes.c_double), ("factor_s", ctypes.c_double), ("trait_s", ctypes.c_double),
                ("loop_s", ctypes.c_double), ("read_busy_s", ctypes.c_double),
                ("h2d_busy_s", ctypes.c_double), ("xr_bytes", ctypes.c_double),
                ("h2d_bytes", ctypes.c_double), ("trait_snps", ctypes.c_int64), ("block", ctypes.c_int64),
                ("passes", ctypes.c_int), ("traits_per_pass", ctypes.c_int), ("io_threads", ctypes.c_int)]


class FglsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def lib() -> ctypes.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FglsError(FGLS_ERR_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
        sig = {
            "fgls_hip_last_error": (ctypes.c_char_p, []),
            "fgls_hip_device_count": (i32, []),
            "fgls_hip_ctx_create": (i32, [ctypes.POINTER(vp), i32, i32, i32, i32, i32]),
            "fgls_hip_ctx_destroy": (i32, [vp]),
            "fgls_hip_ctx_ldn": (i32, [vp]),
            "fgls_hip_set_trait": (i32, [vp, dp, i64, dp, i64, dp, vp]),
            "fgls_hip_get_trait": (i32, [vp, dp, dp, dp]),
            "fgls_hip_solve_block": (i32, [vp, dp, i64, i32, dp, dp, i64, vp]),
            "fgls_hip_solve_block_noalloc": (i32, [vp, dp, i64, i32, dp, dp, i64, vp]),
            "fgls_hip_block_reduce": (i32, [vp, dp, i64, i32, vp]),
            "fgls_hip_block_finalize": (i32, [vp, i32, dp, dp, i64, vp]),
            "fgls_hip_ctx_set_mode": (i32, [vp, i32]),
            "fgls_hip_ctx_set_slice_form": (i32, [vp, i32]),
            "fgls_hip_ctx_slice_info": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]),
            "fgls_hip_block_reduce_i8": (i32, [vp, dp, i64, i32, vp]),
            "fgls_hip_solve_block_i8": (i32, [vp, dp, i64, i32, dp, dp, i64, vp]),
            "fgls_hip_to_i8": (i32, [dp, i64, i32, i32, dp, i64, dp, vp]),
            "fgls_hip_trsm": (i32, [vp, dp, i64, i32, vp]),
            "fgls_hip_trmm": (i32, [vp, dp, i64, i32, dp, i64, vp]),
            "fgls_hip_snp_stage": (i32, [vp, dp, i64, i32, dp, dp, i64, vp]),
            "fgls_hip_check": (i32, [vp, vp, ctypes.POINTER(i64), ctypes.POINTER(i32)]),
            "fgls_hip_to_dosage_i8": (i32, [dp, i64, i32, i32, ctypes.c_double, dp, i64, dp, vp]),
            "fgls_hip_block_reduce_dosage": (i32, [vp, dp, i64, i32, ctypes.c_double, vp]),
            "fgls_hip_solve_block_dosage": (i32, [vp, dp, i64, i32, ctypes.c_double, dp, dp, i64, vp]),
            "fgls_hip_i8_to_f64": (i32, [dp, i64, i32, i32, dp, i64, vp]),
            "fgls_hip_check_i8": (i32, [dp, i64, i32, i32, dp, vp]),
            "fgls_hip_u2_to_i8": (i32, [dp, i64, i32, i32, dp, i64, dp, vp]),
            "fgls_hip_i16_to_dosage_i8": (i32, [dp, i64, i32, i32, dp, i64, dp, vp]),
            "fgls_hip_i16_to_f64": (i32, [dp, i64, i32, i32, ctypes.c_double, dp, i64, vp]),
            "fgls_hip_x8t_bytes": (ctypes.c_size_t, [i32, i32, i32]),
            "fgls_hip_tile_i8": (i32, [dp, i64, i32, i32, i32, dp, vp]),
            "fgls_hip_block_reduce_i8t": (i32, [vp, dp, i32, i32, ctypes.c_double, vp]),
            "fgls_hip_solve_block_i8t": (i32, [vp, dp, i32, i32, ctypes.c_double, dp, dp, i64, vp]),
            "fgls_hip_eig_solve_block_i8t": (i32, [vp, dp, i32, i32, ctypes.c_double, dp, i64, dp, i64,
                                                   i64, vp]),
            "fgls_hip_x2t_bytes": (ctypes.c_size_t, [i32, i32]),
            "fgls_hip_tile_u2": (i32, [dp, i64, i32, i32, dp, dp, vp]),
            "fgls_hip_tile_u2_packed": (i32, [dp, i64, i32, i32, dp, dp, vp]),
            "fgls_hip_block_reduce_u2t": (i32, [vp, dp, i32, vp]),
            "fgls_hip_solve_block_u2t": (i32, [vp, dp, i32, dp, dp, i64, vp]),
            "fgls_hip_eig_solve_block_u2t": (i32, [vp, dp, i32, dp, i64, dp, i64, i64, vp]),
            "fgls_hip_set_trait_phi": (i32, [vp, dp, i64, ctypes.c_double, ctypes.c_double, dp, i64, dp, vp]),
            "fgls_hip_get_factor": (i32, [vp, dp, i64, vp]),
            "fgls_hip_adjoint_begin": (i32, [vp, dp, i64, ctypes.c_double, vp]),
            "fgls_hip_adjoint_block": (i32, [vp, dp, i64, i32, dp, i64, vp]),
            "fgls_hip_adjoint_block_i8": (i32, [vp, dp, i64, i32, dp, i64, vp]),
            "fgls_hip_adjoint_end": (i32, [vp, dp, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), vp]),
            "fgls_factor_M": (i32, [i32, dp, i64, ctypes.c_double, ctypes.c_double, dp, i64, i32]),
            "fgls_last_run_stats": (i32, [ctypes.POINTER(RunStats)]),
            "fgls_chol_gpu": (i32, [i32] * 9 + [ctypes.c_char_p] * 8),
            "fgls_chol_gpu_ex": (i32, [i32] * 9 + [ctypes.c_char_p] * 8 + [ctypes.c_void_p]),
            "fgls_eigen": (i32, [i32] * 9 + [ctypes.c_char_p] * 8),
            "fgls_eigen_gpu_ex": (i32, [i32] * 9 + [ctypes.c_char_p] * 8 + [ctypes.c_void_p]),
            "fgls_eigen_decompose": (i32, [i32, dp, i64, dp, i64, dp, i32]),
            "fgls_hip_eigen_decompose": (i32, [i32, dp, i64, dp, i64, dp, i32, vp]),
            "fgls_hip_eig_create": (i32, [ctypes.POINTER(vp), i32, i32, i32, i32, i32]),
            "fgls_hip_eig_destroy": (i32, [vp]),
            "fgls_hip_eig_ldn": (i32, [vp]),
            "fgls_hip_eig_set_mode": (i32, [vp, i32]),
            "fgls_hip_eig_set_basis": (i32, [vp, dp, i64, dp, dp, i64, vp]),
            "fgls_hip_eig_set_traits": (i32, [vp, i32, dp, dp, dp, i64, vp]),
            "fgls_hip_eig_get_trait": (i32, [vp, i32, dp, dp, dp]),
            "fgls_hip_eig_solve_block": (i32, [vp, dp, i64, i32, dp, i64, dp, i64, i64, vp]),
            "fgls_hip_eig_solve_block_wide": (i32, [vp, dp, i64, i32, i32, dp, i64, dp, i64, i64, vp]),
            "fgls_hip_eig_solve_block_i8": (i32, [vp, dp, i64, i32, dp, i64, dp, i64, i64, vp]),
            "fgls_hip_eig_solve_block_dosage": (i32, [vp, dp, i64, i32, ctypes.c_double, dp, i64, dp, i64,
                                                      i64, vp]),
            "fgls_hip_eig_check": (i32, [vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i64),
                                         ctypes.POINTER(i32)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return [s for s in ("fgls_hip_last_error", "fgls_hip_device_count", "fgls_hip_ctx_create",
                        "fgls_hip_ctx_destroy", "fgls_hip_ctx_ldn", "fgls_hip_set_trait",
                        "fgls_hip_get_trait", "fgls_hip_solve_block", "fgls_hip_solve_block_noalloc",
                        "fgls_hip_block_reduce",
                        "fgls_hip_block_finalize", "fgls_hip_ctx_set_mode", "fgls_hip_block_reduce_i8",
                        "fgls_hip_ctx_set_slice_form", "fgls_hip_ctx_slice_info",
                        "fgls_hip_solve_block_i8", "fgls_hip_to_i8", "fgls_hip_trsm",
                        "fgls_hip_trmm", "fgls_hip_snp_stage", "fgls_hip_check",
                        "fgls_hip_adjoint_begin", "fgls_hip_adjoint_block", "fgls_hip_adjoint_end",
                        "fgls_hip_adjoint_block_i8",
                        "fgls_hip_set_trait_phi", "fgls_hip_get_factor", "fgls_hip_to_dosage_i8",
                        "fgls_hip_block_reduce_dosage", "fgls_hip_solve_block_dosage",
                        "fgls_hip_i8_to_f64", "fgls_hip_check_i8", "fgls_hip_u2_to_i8",
                        "fgls_hip_x8t_bytes",
                        "fgls_hip_i16_to_dosage_i8", "fgls_hip_i16_to_f64",
                        "fgls_hip_tile_i8", "fgls_hip_block_reduce_i8t", "fgls_hip_solve_block_i8t",
                        "fgls_hip_eig_solve_block_i8t", "fgls_hip_x2t_bytes", "fgls_hip_tile_u2",
                        "fgls_hip_tile_u2_packed", "fgls_hip_block_reduce_u2t", "fgls_hip_solve_block_u2t",
                        "fgls_hip_eig_solve_block_u2t",
                        "fgls_factor_M", "fgls_last_run_stats", "fgls_eigen", "fgls_eigen_gpu_ex",
                        "fgls_eigen_decompose",
                        "fgls_hip_eigen_decompose",
                        "fgls_hip_eig_create", "fgls_hip_eig_destroy", "fgls_hip_eig_ldn",
                        "fgls_hip_eig_set_mode", "fgls_hip_eig_set_basis", "fgls_hip_eig_set_traits",
                        "fgls_hip_eig_get_trait", "fgls_hip_eig_solve_block", "fgls_hip_eig_solve_block_wide",
                        "fgls_hip_eig_solve_block_i8", "fgls_hip_eig_solve_block_dosage",
                        "fgls_hip_eig_check")
            if hasattr(lib(), s)]


def last_error() -> str:
    return (lib().fgls_hip_last_error() or b"").decode()


def _check(rc: int):
    if rc != FGLS_OK:
        raise FglsError(rc, last_error())


def device_count() -> int:
    return int(lib().fgls_hip_device_count())


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _stream(stream) -> int:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _need_cuda_f64(t, name: str):
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float64:
        raise FglsError(FGLS_ERR_ARG, f"{name} must be a float64 CUDA tensor")
    if not t.is_contiguous():
        raise FglsError(FGLS_ERR_ARG, f"{name} must be contiguous")


def colmajor(a, ld: int | None = None, device="cuda"):
    """numpy n x k (any order) -> contiguous float64 tensor of shape (k, ld)
    holding the matrix column-major with leading dimension ld (zero padded)."""
    import numpy as np
    import torch
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    n, k = a.shape
    ld = n if ld is None else ld
    out = torch.zeros((k, ld), dtype=torch.float64)
    out[:, :n] = torch.from_numpy(np.ascontiguousarray(a.T))
    return out.to(device)


class DeviceContext:
    """One device context (include/fgls_hip.h, fgls_hip_ctx_*)."""

    def __init__(self, n: int, wXL: int, wXR: int = 1, max_block: int = 4096, device: int = 0):
        self.n, self.wXL, self.wXR, self.p = n, wXL, wXR, wXL + wXR
        self.max_block = max_block
        self.device = device
        h = ctypes.c_void_p()
        _check(lib().fgls_hip_ctx_create(ctypes.byref(h), n, wXL, wXR, max_block, device))
        self._h = h
        self.ldn = int(lib().fgls_hip_ctx_ldn(h))

    def close(self):
        if getattr(self, "_h", None):
            lib().fgls_hip_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- per trait ---------------------------------------------------------
    def set_trait(self, L, XL, y, stream=None):
        """L: (n, ldl) column-major lower factor; XL: (wXL, ldxl) or None; y: (>=n,)."""
        _need_cuda_f64(L, "L")
        _need_cuda_f64(y, "y")
        ldxl = self.n
        if self.wXL > 0:
            _need_cuda_f64(XL, "XL")
            ldxl = XL.shape[1]
        _check(lib().fgls_hip_set_trait(self._h, _ptr(L), L.shape[1], _ptr(XL) if self.wXL else 0,
                                        ldxl, _ptr(y), _stream(stream)))

    def set_trait_phi(self, Phi, h2: float, sigma2: float, XL, y, stream=None):
        """Build M = sigma2 (h2 Phi + (1 - h2) I) and factor it on the device, then
        the per-trait precompute.  Phi: (n, ldphi) column-major (symmetric)."""
        _need_cuda_f64(Phi, "Phi")
        _need_cuda_f64(y, "y")
        ldxl = self.n
        if self.wXL > 0:
            _need_cuda_f64(XL, "XL")
            ldxl = XL.shape[1]
        _check(lib().fgls_hip_set_trait_phi(self._h, _ptr(Phi), Phi.shape[1], float(h2), float(sigma2),
                                            _ptr(XL) if self.wXL else 0, ldxl, _ptr(y), _stream(stream)))

    def factor(self, stream=None):
        """The current trait's factor L as an (n, n) tensor holding L column-major
        (row c = column c; entries above the diagonal are unspecified)."""
        import torch
        L = torch.empty((self.n, self.n), dtype=torch.float64, device="cuda")
        _check(lib().fgls_hip_get_factor(self._h, _ptr(L), self.n, _stream(stream)))
        return L

    def trait(self):
        import numpy as np
        Bt = np.zeros(max(self.wXL, 1))
        Vtl = np.zeros(max(self.wXL * self.wXL, 1))
        res = np.zeros(1)
        _check(lib().fgls_hip_get_trait(self._h, Bt.ctypes.data, Vtl.ctypes.data, res.ctypes.data))
        return Bt[:self.wXL], Vtl[:self.wXL * self.wXL].reshape(self.wXL, self.wXL, order="F"), float(res[0])

    # -- per block ---------------------------------------------------------
    def solve_block(self, XR, count: int | None = None, B=None, V=None, snp_base: int = 0, stream=None):
        """XR: (count*wXR, ldx) tensor.  Returns (B (count, p), V (count, p*p))."""
        import torch
        _need_cuda_f64(XR, "XR")
        if count is None:
            count = XR.shape[0] // self.wXR
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=XR.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=XR.device)
        _check(lib().fgls_hip_solve_block(self._h, _ptr(XR), XR.shape[1], count, _ptr(B), _ptr(V),
                                          snp_base, _stream(stream)))
        return B, V

    def solve_block_noalloc(self, XR, B, V, count: int | None = None, snp_base: int = 0, stream=None):
        """fgls_hip_solve_block_noalloc: solve_block with its scratch in the
        library's static device array (no allocation; B and V are given)."""
        _need_cuda_f64(XR, "XR")
        _need_cuda_f64(B, "B")
        _need_cuda_f64(V, "V")
        if count is None:
            count = XR.shape[0] // self.wXR
        _check(lib().fgls_hip_solve_block_noalloc(self._h, _ptr(XR), XR.shape[1], count, _ptr(B), _ptr(V),
                                                  snp_base, _stream(stream)))
        return B, V

    def block_reduce(self, XR, count: int | None = None, stream=None):
        _need_cuda_f64(XR, "XR")
        count = XR.shape[0] // self.wXR if count is None else count
        _check(lib().fgls_hip_block_reduce(self._h, _ptr(XR), XR.shape[1], count, _stream(stream)))
        return count

    def block_finalize(self, count: int, B, V, snp_base: int = 0, stream=None):
        _check(lib().fgls_hip_block_finalize(self._h, count, _ptr(B), _ptr(V), snp_base, _stream(stream)))
        return B, V

    # -- mixed mode (int8 genotypes, sliced L^-1) ---------------------------
    def set_mode(self, mode: int):
        _check(lib().fgls_hip_ctx_set_mode(self._h, mode))
        self.mode = mode

    def set_slice_form(self, form: int):
        """FGLS_SLICES_AUTO (five digits of L^-1 where the error bound allows) or FGLS_SLICES_SIX."""
        _check(lib().fgls_hip_ctx_set_slice_form(self._h, form))

    def slice_info(self) -> tuple[int, float]:
        """(digits of the slices in use: 5, 6, or 0 before any; loss of the five-digit form, -1 if not tried)."""
        ns, loss = ctypes.c_int(0), ctypes.c_double(0.0)
        _check(lib().fgls_hip_ctx_slice_info(self._h, ctypes.byref(ns), ctypes.byref(loss)))
        return ns.value, loss.value

    @property
    def ld8(self) -> int:
        """Smallest valid leading dimension of an int8 genotype block."""
        return (self.n + 63) // 64 * 64

    def to_i8(self, X, ncols: int | None = None, X8=None, stream=None):
        """FP64 block (ncols, ldx) -> int8 block (ncols, ld8); raises if a value
        is not an integer in [-127, 127]."""
        return to_i8(X, self.n, ncols, X8, stream)

    def to_dosage_i8(self, X, scale: float = 1000.0, ncols: int | None = None, X8=None, stream=None):
        """FP64 dosages (ncols, ldx) -> two int8 digit planes, a (2 * ncols, ld8)
        tensor (lo plane first); raises if a value times scale is not an integer
        in [-16256, 16383]."""
        import torch
        _need_cuda_f64(X, "X")
        ncols = X.shape[0] if ncols is None else ncols
        if X8 is None:
            X8 = torch.empty((2 * ncols, self.ld8), dtype=torch.int8, device=X.device)
        bad = torch.zeros(1, dtype=torch.int32, device=X.device)
        _check(lib().fgls_hip_to_dosage_i8(_ptr(X), X.shape[1], self.n, ncols, float(scale), _ptr(X8),
                                           X8.shape[1], _ptr(bad), _stream(stream)))
        nb = int(bad.item())
        if nb:
            raise FglsError(FGLS_ERR_ARG, f"{nb} dosage values are not multiples of 1/{scale:g} in range")
        return X8

    def block_reduce_dosage(self, X8, count: int | None = None, scale: float = 1000.0, stream=None):
        count = X8.shape[0] // 2 if count is None else count
        _check(lib().fgls_hip_block_reduce_dosage(self._h, _ptr(X8), X8.shape[1], count, float(scale),
                                                  _stream(stream)))
        return count

    def solve_block_dosage(self, X8, scale: float = 1000.0, count: int | None = None, B=None, V=None,
                           snp_base: int = 0, stream=None):
        """X8: (2 * count, ld8) digit planes from to_dosage_i8."""
        import torch
        if count is None:
            count = X8.shape[0] // 2
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=X8.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=X8.device)
        _check(lib().fgls_hip_solve_block_dosage(self._h, _ptr(X8), X8.shape[1], count, float(scale),
                                                 _ptr(B), _ptr(V), snp_base, _stream(stream)))
        return B, V

    def block_reduce_i8t(self, X8T, count: int, planes: int = 1, scale: float = 1.0, stream=None):
        """The int8 product of tiled genotypes (tile_i8)."""
        _check(lib().fgls_hip_block_reduce_i8t(self._h, _ptr(X8T), count, planes, float(scale),
                                               _stream(stream)))
        return count

    def solve_block_i8t(self, X8T, count: int, planes: int = 1, scale: float = 1.0, B=None, V=None,
                        snp_base: int = 0, stream=None):
        import torch
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=X8T.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=X8T.device)
        _check(lib().fgls_hip_solve_block_i8t(self._h, _ptr(X8T), count, planes, float(scale), _ptr(B),
                                              _ptr(V), snp_base, _stream(stream)))
        return B, V

    def block_reduce_u2t(self, X2T, count: int, stream=None):
        """The int8 product of 2-bit tiled genotypes (tile_u2 / tile_u2_packed)."""
        _check(lib().fgls_hip_block_reduce_u2t(self._h, _ptr(X2T), count, _stream(stream)))
        return count

    def solve_block_u2t(self, X2T, count: int, B=None, V=None, snp_base: int = 0, stream=None):
        import torch
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=X2T.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=X2T.device)
        _check(lib().fgls_hip_solve_block_u2t(self._h, _ptr(X2T), count, _ptr(B), _ptr(V), snp_base,
                                              _stream(stream)))
        return B, V

    def block_reduce_i8(self, X8, count: int | None = None, stream=None):
        """X8: (count*wXR, ld8) int8 genotypes, SNP i's columns at rows i*wXR .."""
        count = X8.shape[0] // self.wXR if count is None else count
        _check(lib().fgls_hip_block_reduce_i8(self._h, _ptr(X8), X8.shape[1], count, _stream(stream)))
        return count

    def solve_block_i8(self, X8, count: int | None = None, B=None, V=None, snp_base: int = 0, stream=None):
        """X8: (count*wXR, ld8) int8 genotypes.  Returns (B (count, p), V (count, p*p))."""
        import torch
        if count is None:
            count = X8.shape[0] // self.wXR
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=X8.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=X8.device)
        _check(lib().fgls_hip_solve_block_i8(self._h, _ptr(X8), X8.shape[1], count, _ptr(B), _ptr(V),
                                             snp_base, _stream(stream)))
        return B, V

    def trsm(self, X, ncols: int | None = None, stream=None):
        """In place X <- L^-1 X (X: (ncols, ldx))."""
        _need_cuda_f64(X, "X")
        ncols = X.shape[0] if ncols is None else ncols
        _check(lib().fgls_hip_trsm(self._h, _ptr(X), X.shape[1], ncols, _stream(stream)))
        return X

    def trmm(self, X, ncols: int | None = None, Z=None, stream=None):
        """Z = L^-1 X through the explicit inverse; Z has shape (ncols, ldn)."""
        import torch
        _need_cuda_f64(X, "X")
        ncols = X.shape[0] if ncols is None else ncols
        if Z is None:
            Z = torch.empty((ncols, self.ldn), dtype=torch.float64, device=X.device)
        _check(lib().fgls_hip_trmm(self._h, _ptr(X), X.shape[1], ncols, _ptr(Z), Z.shape[1], _stream(stream)))
        return Z

    def snp_stage(self, Z, count: int | None = None, B=None, V=None, snp_base: int = 0, stream=None):
        import torch
        _need_cuda_f64(Z, "Z")
        if count is None:
            count = Z.shape[0] // self.wXR
        if B is None:
            B = torch.empty((count, self.p), dtype=torch.float64, device=Z.device)
        if V is None:
            V = torch.empty((count, self.p * self.p), dtype=torch.float64, device=Z.device)
        _check(lib().fgls_hip_snp_stage(self._h, _ptr(Z), Z.shape[1], count, _ptr(B), _ptr(V),
                                        snp_base, _stream(stream)))
        return B, V

    # -- adjoint of l = sum_i w_i b_i[p-1]^2 (FP64, wXR = 1) ----------------
    def adjoint_begin(self, Phi, sigma2: float, stream=None):
        """Phi: (n, ldphi) column-major (symmetric) on the device."""
        _need_cuda_f64(Phi, "Phi")
        _check(lib().fgls_hip_adjoint_begin(self._h, _ptr(Phi), Phi.shape[1], float(sigma2),
                                            _stream(stream)))

    def adjoint_block(self, XR, w, count: int | None = None, snp_base: int = 0, stream=None):
        """XR: (count, ldx) genotype block; w: (count,) loss weights."""
        _need_cuda_f64(XR, "XR")
        _need_cuda_f64(w, "w")
        count = XR.shape[0] if count is None else count
        if w.numel() < count:
            raise FglsError(FGLS_ERR_ARG, "w has fewer than count weights")
        _check(lib().fgls_hip_adjoint_block(self._h, _ptr(XR), XR.shape[1], count, _ptr(w), snp_base,
                                            _stream(stream)))

    def adjoint_block_i8(self, X8, w, count: int | None = None, snp_base: int = 0, stream=None):
        """Mixed mode: X8 (count, ld8) int8 genotypes (to_i8); w: (count,) weights."""
        _need_cuda_f64(w, "w")
        count = X8.shape[0] if count is None else count
        if w.numel() < count:
            raise FglsError(FGLS_ERR_ARG, "w has fewer than count weights")
        _check(lib().fgls_hip_adjoint_block_i8(self._h, _ptr(X8), X8.shape[1], count, _ptr(w), snp_base,
                                               _stream(stream)))

    def adjoint_end(self, grad_y=None, stream=None):
        """Returns (loss, dl/dy as a (n,) tensor, dl/dh2)."""
        import torch
        if grad_y is None:
            grad_y = torch.empty(self.n, dtype=torch.float64, device="cuda")
        _need_cuda_f64(grad_y, "grad_y")
        loss = ctypes.c_double(0.0)
        gh = ctypes.c_double(0.0)
        _check(lib().fgls_hip_adjoint_end(self._h, _ptr(grad_y), ctypes.byref(loss), ctypes.byref(gh),
                                          _stream(stream)))
        return loss.value, grad_y, gh.value

    def check(self, stream=None):
        """Raise FglsError(FGLS_ERR_NOTSPD) carrying .bad_snp / .info if a per-SNP
        factorisation failed since the last check."""
        snp = ctypes.c_int64(-1)
        info = ctypes.c_int(0)
        rc = lib().fgls_hip_check(self._h, _stream(stream), ctypes.byref(snp), ctypes.byref(info))
        if rc != FGLS_OK:
            err = FglsError(rc, last_error())
            err.bad_snp = int(snp.value)
            err.info = int(info.value)
            raise err


class EigenContext:
    """The eigenbasis variant on the device (include/fgls_hip.h, fgls_hip_eig_*):
    one eigendecomposition Phi = Z diag(W) Z' serves every trait; a block of
    SNPs is solved for all the traits of the last set_traits at once."""

    def __init__(self, n: int, wXL: int, max_traits: int = 1, max_block: int = 4096, device: int = 0):
        self.n, self.wXL, self.p = n, wXL, wXL + 1
        self.max_traits, self.max_block, self.device = max_traits, max_block, device
        self.ntr = 0
        h = ctypes.c_void_p()
        _check(lib().fgls_hip_eig_create(ctypes.byref(h), n, wXL, max_traits, max_block, device))
        self._h = h
        self.ldn = int(lib().fgls_hip_eig_ldn(h))

    def close(self):
        if getattr(self, "_h", None):
            lib().fgls_hip_eig_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ld8(self) -> int:
        return (self.n + 63) // 64 * 64

    def set_mode(self, mode: int):
        _check(lib().fgls_hip_eig_set_mode(self._h, mode))

    def set_basis(self, Z, W, XL, stream=None):
        """Z: (n, ldz) tensor holding the eigenvectors column-major (row r =
        eigenvector r); W: (n,) eigenvalues; XL: (wXL, ldxl) or None."""
        _need_cuda_f64(Z, "Z")
        _need_cuda_f64(W, "W")
        ldxl = self.n
        if self.wXL > 0:
            _need_cuda_f64(XL, "XL")
            ldxl = XL.shape[1]
        _check(lib().fgls_hip_eig_set_basis(self._h, _ptr(Z), Z.shape[1], _ptr(W),
                                            _ptr(XL) if self.wXL else 0, ldxl, _stream(stream)))

    def set_traits(self, h2, sigma2, Y, stream=None):
        """h2, sigma2: sequences of ntr floats; Y: (ntr, ldy) tensor (row k = trait k)."""
        import numpy as np
        _need_cuda_f64(Y, "Y")
        h = np.ascontiguousarray(np.atleast_1d(np.asarray(h2, dtype=np.float64)))
        sg = np.ascontiguousarray(np.atleast_1d(np.asarray(sigma2, dtype=np.float64)))
        if sg.size == 1 and h.size > 1:
            sg = np.full(h.size, float(sg[0]))
        ntr = int(h.size)
        if sg.size != ntr or Y.shape[0] < ntr:
            raise FglsError(FGLS_ERR_ARG, "h2, sigma2 and Y disagree on the number of traits")
        _check(lib().fgls_hip_eig_set_traits(self._h, ntr, h.ctypes.data, sg.ctypes.data, _ptr(Y),
                                             Y.shape[1], _stream(stream)))
        self.ntr = ntr

    def trait(self, k: int = 0):
        import numpy as np
        Bt = np.zeros(max(self.wXL, 1))
        Vtl = np.zeros(max(self.wXL * self.wXL, 1))
        res = np.zeros(1)
        _check(lib().fgls_hip_eig_get_trait(self._h, k, Bt.ctypes.data, Vtl.ctypes.data, res.ctypes.data))
        return Bt[:self.wXL], Vtl[:self.wXL * self.wXL].reshape(self.wXL, self.wXL, order="F"), float(res[0])

    def _outputs(self, count, B, V, device):
        import torch
        if B is None:
            B = torch.empty((self.ntr, count, self.p), dtype=torch.float64, device=device)
        if V is None:
            V = torch.empty((self.ntr, count, self.p * self.p), dtype=torch.float64, device=device)
        return B, V

    def solve_block(self, XR, count: int | None = None, B=None, V=None, snp_base: int = 0, stream=None):
        """XR: (count, ldx).  Returns B (ntr, count, p) and V (ntr, count, p*p)."""
        _need_cuda_f64(XR, "XR")
        count = XR.shape[0] if count is None else count
        B, V = self._outputs(count, B, V, XR.device)
        _check(lib().fgls_hip_eig_solve_block(self._h, _ptr(XR), XR.shape[1], count, _ptr(B),
                                              B.stride(0), _ptr(V), V.stride(0), snp_base, _stream(stream)))
        return B, V

    def solve_block_wide(self, XR, wXR: int, count: int | None = None, B=None, V=None, snp_base: int = 0,
                         stream=None):
        """Several columns per SNP (FP64): XR (count*wXR, ldx), SNP i's columns
        at rows i*wXR .. i*wXR + wXR - 1.  Returns B (ntr, count, p) and
        V (ntr, count, p*p) with p = wXL + wXR."""
        import torch
        _need_cuda_f64(XR, "XR")
        count = XR.shape[0] // wXR if count is None else count
        p = self.wXL + wXR
        if B is None:
            B = torch.empty((self.ntr, count, p), dtype=torch.float64, device=XR.device)
        if V is None:
            V = torch.empty((self.ntr, count, p * p), dtype=torch.float64, device=XR.device)
        _check(lib().fgls_hip_eig_solve_block_wide(self._h, _ptr(XR), XR.shape[1], count, wXR, _ptr(B),
                                                   B.stride(0), _ptr(V), V.stride(0), snp_base,
                                                   _stream(stream)))
        return B, V

    def solve_block_i8(self, X8, count: int | None = None, B=None, V=None, snp_base: int = 0, stream=None):
        count = X8.shape[0] if count is None else count
        B, V = self._outputs(count, B, V, X8.device)
        _check(lib().fgls_hip_eig_solve_block_i8(self._h, _ptr(X8), X8.shape[1], count, _ptr(B),
                                                 B.stride(0), _ptr(V), V.stride(0), snp_base,
                                                 _stream(stream)))
        return B, V

    def solve_block_dosage(self, X8, scale: float = 1000.0, count: int | None = None, B=None, V=None,
                           snp_base: int = 0, stream=None):
        """X8: (2 * count, ld8) digit planes (DeviceContext.to_dosage_i8)."""
        count = X8.shape[0] // 2 if count is None else count
        B, V = self._outputs(count, B, V, X8.device)
        _check(lib().fgls_hip_eig_solve_block_dosage(self._h, _ptr(X8), X8.shape[1], count, float(scale),
                                                     _ptr(B), B.stride(0), _ptr(V), V.stride(0), snp_base,
                                                     _stream(stream)))
        return B, V

    def solve_block_i8t(self, X8T, count: int, planes: int = 1, scale: float = 1.0, B=None, V=None,
                        snp_base: int = 0, stream=None):
        """Tiled genotypes (tile_i8)."""
        B, V = self._outputs(count, B, V, X8T.device)
        _check(lib().fgls_hip_eig_solve_block_i8t(self._h, _ptr(X8T), count, planes, float(scale), _ptr(B),
                                                  B.stride(0), _ptr(V), V.stride(0), snp_base,
                                                  _stream(stream)))
        return B, V

    def solve_block_u2t(self, X2T, count: int, B=None, V=None, snp_base: int = 0, stream=None):
        """2-bit tiled genotypes (tile_u2 / tile_u2_packed)."""
        B, V = self._outputs(count, B, V, X2T.device)
        _check(lib().fgls_hip_eig_solve_block_u2t(self._h, _ptr(X2T), count, _ptr(B), B.stride(0), _ptr(V),
                                                  V.stride(0), snp_base, _stream(stream)))
        return B, V

    def check(self, stream=None):
        tr = ctypes.c_int(-1)
        snp = ctypes.c_int64(-1)
        info = ctypes.c_int(0)
        rc = lib().fgls_hip_eig_check(self._h, _stream(stream), ctypes.byref(tr), ctypes.byref(snp),
                                      ctypes.byref(info))
        if rc != FGLS_OK:
            err = FglsError(rc, last_error())
            err.trait, err.bad_snp, err.info = int(tr.value), int(snp.value), int(info.value)
            raise err


def eigen_decompose(Phi, threads: int = 0):
    """Host Phi = Z diag(W) Z' by LAPACK dsyevr (fgls_eigen.c:780-806): numpy in,
    (Z as an n x n Fortran-order array, W ascending) out."""
    import numpy as np
    Phi = np.asfortranarray(Phi, dtype=np.float64)
    n = Phi.shape[0]
    Z = np.empty((n, n), dtype=np.float64, order="F")
    W = np.empty(n, dtype=np.float64)
    _check(lib().fgls_eigen_decompose(n, Phi.ctypes.data, n, Z.ctypes.data, n, W.ctypes.data, int(threads)))
    return Z, W


def pack_u2(X):
    """Genotypes X (n x m, values 0, 1, 2; column j = SNP j) -> the 2-bit
    layout of xr_format=u2 / fgls_hip_u2_to_i8: an (m, ceil(n/4)) uint8 array,
    individual r in bits 2(r%4)..2(r%4)+1 of byte r/4.  Values outside
    {0, 1, 2} raise."""
    import numpy as np
    X = np.asarray(X)
    if X.ndim == 1:
        X = X[:, None]
    n, m = X.shape
    c = np.rint(X).astype(np.int64)
    if not np.array_equal(c, X) or c.min(initial=0) < 0 or c.max(initial=0) > 2:
        raise FglsError(FGLS_ERR_ARG, "2-bit genotypes must be 0, 1 or 2")
    nb = (n + 3) // 4
    pad = np.zeros((nb * 4, m), dtype=np.uint8)
    pad[:n] = c.astype(np.uint8)
    q = pad.reshape(nb, 4, m)
    packed = q[:, 0] | (q[:, 1] << 2) | (q[:, 2] << 4) | (q[:, 3] << 6)
    return np.ascontiguousarray(packed.T)


def tile_i8(X8, n: int, count: int | None = None, planes: int = 1, X8T=None, stream=None):
    """Column-major int8 genotypes X8 ((planes * count, ld8) tensor; dosage digit
    planes as from DeviceContext.to_dosage_i8) -> the tiled layout of the int8
    product, a flat int8 tensor of x8t_bytes(n, count, planes) bytes."""
    import torch
    count = X8.shape[0] // planes if count is None else count
    nbytes = int(lib().fgls_hip_x8t_bytes(n, count, planes))
    if X8T is None:
        X8T = torch.empty(nbytes, dtype=torch.int8, device=X8.device)
    if X8T.numel() < nbytes:
        raise FglsError(FGLS_ERR_ARG, f"X8T needs {nbytes} bytes")
    _check(lib().fgls_hip_tile_i8(_ptr(X8), X8.shape[1], n, count, planes, _ptr(X8T), _stream(stream)))
    return X8T


def x2t_bytes(n: int, count: int) -> int:
    """Bytes of the 2-bit tiled genotypes of count SNPs (fgls_hip_x2t_bytes)."""
    return int(lib().fgls_hip_x2t_bytes(n, count))


def tile_u2(X8, n: int, count: int | None = None, X2T=None, bad=None, stream=None):
    """Column-major int8 hard-call genotypes X8 ((count, ld8) tensor, values
    0..3) -> the 2-bit tiles of the int8 product (a flat uint8 tensor of
    x2t_bytes(n, count) bytes).  Raises if a value is outside 0..3, unless a
    device counter `bad` (int32, caller-cleared) is given, which then
    receives the count without a synchronisation."""
    import torch
    count = X8.shape[0] if count is None else count
    nbytes = x2t_bytes(n, count)
    if X2T is None:
        X2T = torch.empty(nbytes, dtype=torch.uint8, device=X8.device)
    if X2T.numel() < nbytes:
        raise FglsError(FGLS_ERR_ARG, f"X2T needs {nbytes} bytes")
    own = bad is None
    if own:
        bad = torch.zeros(1, dtype=torch.int32, device=X8.device)
    _check(lib().fgls_hip_tile_u2(_ptr(X8), X8.shape[1], n, count, _ptr(X2T), _ptr(bad), _stream(stream)))
    if own and int(bad.item()):
        raise FglsError(FGLS_ERR_ARG, f"{int(bad.item())} genotype values are outside 0..3")
    return X2T


def tile_u2_packed(P, n: int, count: int | None = None, X2T=None, bad=None, stream=None):
    """2-bit genotype file columns P ((count, ldp) uint8 tensor, pack_u2's
    layout) -> the 2-bit tiles of the int8 product.  Raises if a code is 3,
    unless a device counter `bad` (int32, caller-cleared) is given, which then
    receives the count without a synchronisation."""
    import torch
    count = P.shape[0] if count is None else count
    nbytes = x2t_bytes(n, count)
    if X2T is None:
        X2T = torch.empty(nbytes, dtype=torch.uint8, device=P.device)
    own = bad is None
    if own:
        bad = torch.zeros(1, dtype=torch.int32, device=P.device)
    _check(lib().fgls_hip_tile_u2_packed(_ptr(P), P.shape[1], n, count, _ptr(X2T), _ptr(bad), _stream(stream)))
    if own and int(bad.item()):
        raise FglsError(FGLS_ERR_ARG, f"{int(bad.item())} 2-bit codes are 3 (not a genotype)")
    return X2T


def pack_i16(X, scale: float = 1000.0):
    """Dosages X (n x m; column j = SNP j) with a fixed number of decimals ->
    the 16-bit layout of xr_format=i16: an (m, n) int16 array of
    v = rint(x * scale); raises unless x * scale is within 8 eps max(1, |x * scale|)
    of v (eps = 2^-52, the device conversion's test) and v fits in
    [-16256, 16383]."""
    import numpy as np
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    y = X * scale
    v = np.rint(y)
    tol = 8.0 * np.finfo(np.float64).eps * np.maximum(1.0, np.abs(y))
    if np.any(np.abs(y - v) > tol) or v.min(initial=0) < -16256 or v.max(initial=0) > 16383:
        raise FglsError(FGLS_ERR_ARG, f"dosages must be multiples of 1/{scale:g} within [-16256, 16383]/scale")
    return np.ascontiguousarray(v.astype(np.int16).T)


def to_i8(X, n: int, ncols: int | None = None, X8=None, stream=None):
    """FP64 genotypes X ((ncols, ldx) CUDA tensor, rows 0..n-1 of each column
    used) -> int8 (ncols, ceil(n/64)*64) column-major block, the input of the
    int8 paths (fgls_hip_to_i8); raises if a value is not an integer in
    [-127, 127]."""
    import torch
    _need_cuda_f64(X, "X")
    ncols = X.shape[0] if ncols is None else ncols
    if X8 is None:
        X8 = torch.empty((ncols, (n + 63) // 64 * 64), dtype=torch.int8, device=X.device)
    bad = torch.zeros(1, dtype=torch.int32, device=X.device)
    _check(lib().fgls_hip_to_i8(_ptr(X), X.shape[1], n, ncols, _ptr(X8), X8.shape[1], _ptr(bad),
                                _stream(stream)))
    nb = int(bad.item())
    if nb:
        raise FglsError(FGLS_ERR_ARG, f"{nb} genotype values are not integers in [-127, 127]")
    return X8


def u2_to_i8(P, n: int, X8=None, stream=None):
    """Device 2-bit genotypes P ((ncols, ldp) uint8 tensor) -> int8 (ncols, ld8)
    column-major block; raises if a code is 3."""
    import torch
    ncols = P.shape[0]
    ld8 = (n + 63) // 64 * 64
    if X8 is None:
        X8 = torch.empty((ncols, ld8), dtype=torch.int8, device=P.device)
    bad = torch.zeros(1, dtype=torch.int32, device=P.device)
    _check(lib().fgls_hip_u2_to_i8(_ptr(P), P.shape[1], n, ncols, _ptr(X8), X8.shape[1], _ptr(bad),
                                   _stream(stream)))
    nb = int(bad.item())
    if nb:
        raise FglsError(FGLS_ERR_ARG, f"{nb} 2-bit codes are 3 (not a genotype)")
    return X8


def eigen_decompose_device(Phi, stream=None, device: int = 0):
    """Device Phi = Z diag(W) Z' (rocSOLVER dsyevd): Phi an (n, ldphi) CUDA
    tensor holding Phi column-major; returns (Z as an (n, n) tensor whose row r
    is eigenvector r, i.e. Z column-major, W ascending)."""
    import torch
    _need_cuda_f64(Phi, "Phi")
    n = Phi.shape[0]
    Z = torch.empty((n, n), dtype=torch.float64, device=Phi.device)
    W = torch.empty(n, dtype=torch.float64, device=Phi.device)
    _check(lib().fgls_hip_eigen_decompose(n, _ptr(Phi), Phi.shape[1], _ptr(Z), n, _ptr(W), device,
                                          _stream(stream)))
    return Z, W


def factor_M(Phi, h2: float, sigma2: float, threads: int = 0):
    """Host Cholesky factor of M = sigma2 (h2 Phi + (1 - h2) I) (fgls_chol_gpu.c:406-430):
    numpy n x n in, numpy lower factor (column-major, Fortran order) out."""
    import numpy as np
    Phi = np.asfortranarray(Phi, dtype=np.float64)
    n = Phi.shape[0]
    L = np.empty((n, n), dtype=np.float64, order="F")
    _check(lib().fgls_factor_M(n, Phi.ctypes.data, n, float(h2), float(sigma2), L.ctypes.data, n,
                               int(threads)))
    return np.tril(L)


def fgls_chol_gpu(n, p, m, t, wXL, wXR, x_b, y_b, num_threads,
                  Phi_path, h2_path, sigma2_path, XL_path, XR_path, Y_path, B_path, V_path,
                  options: dict | None = None) -> int:
    """The reference entry point (fgls_chol_gpu.h:4-10), files in and out.

    Unlike the reference (which prints "[Error]: ..." and exits the process,
    common.c:11-17) this Python mirror raises FglsError; the C symbol used by
    the CLI driver keeps the exit behaviour."""
    L = lib()
    args = [int(v) for v in (n, p, m, t, wXL, wXR, x_b, y_b, num_threads)]
    paths = [os.fsencode(s) for s in (Phi_path, h2_path, sigma2_path, XL_path, XR_path, Y_path,
                                      B_path, V_path)]
    opts = None
    if options:
        opts = ",".join(f"{k}={v}" for k, v in options.items()).encode()
    rc = L.fgls_chol_gpu_ex(*args, *paths, opts)
    _check(rc)
    return rc


def last_run_stats() -> dict:
    """Timings of this thread's last successful fgls_chol_gpu / fgls_eigen
    call (fgls_last_run_stats), as a dict of the fgls_run_stats fields."""
    st = RunStats()
    _check(lib().fgls_last_run_stats(ctypes.byref(st)))
    return {name: getattr(st, name) for name, _ in RunStats._fields_}


def fgls_eigen(n, p, m, t, wXL, wXR, x_b, y_b, num_threads,
               Phi_path, h2_path, sigma2_path, XL_path, XR_path, Y_path, B_path, V_path,
               options: dict | None = None) -> int:
    """The reference's eigenbasis entry point (fgls_eigen.h:68-74) on the GPU,
    files in and out; raises FglsError where the C symbol exits."""
    L = lib()
    args = [int(v) for v in (n, p, m, t, wXL, wXR, x_b, y_b, num_threads)]
    paths = [os.fsencode(s) for s in (Phi_path, h2_path, sigma2_path, XL_path, XR_path, Y_path,
                                      B_path, V_path)]
    opts = None
    if options:
        opts = ",".join(f"{k}={v}" for k, v in options.items()).encode()
    rc = L.fgls_eigen_gpu_ex(*args, *paths, opts)
    _check(rc)
    return rc
