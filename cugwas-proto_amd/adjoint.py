"""Differentiable GWAS scan loss over the FGLS solver.

    l(y, h2) = sum_i w_i * b_i[p-1]**2

where b_i is the per-SNP GLS estimate of the reference's Cholesky FGLS path
(fgls_chol.c:386-439; M = sigma2 (h2 Phi + (1 - h2) I), fgls_chol_gpu.c:406-414)
and b_i[p-1] is the SNP's own coefficient.  The reference computes no
gradients; this is an addition on top of its forward path.

``SnpLoss`` is a ``torch.autograd.Function`` over the C-ABI:

* forward: M = sigma2 (h2 Phi + (1 - h2) I) built and factored on the device
  together with the per-trait precompute (``fgls_hip_set_trait_phi``; no
  host LAPACK call), then the fused block solve (``fgls_hip_solve_block``);
  nothing of L^-1 X_R is kept;
* backward: ``fgls_hip_adjoint_begin/block/end`` recompute L^-1 X_R block by
  block and return dl/dy and dl/dh2.

With m = 0 SNPs the loss is 0 and both gradients are zero.

With ``precision="mixed"`` the genotypes go to the device as int8 and the
block products of both passes run on the int8 MFMA
(``fgls_hip_solve_block_i8``, ``fgls_hip_adjoint_block_i8``).  Gradients
flow to ``y`` and ``h2`` only.
"""
from __future__ import annotations

import numpy as np
import torch

from .fgls import DeviceContext, FglsError, FGLS_ERR_ARG, FGLS_MODE_MIXED_I8


def _as_f64_cuda(t, name):
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    t = t.detach().to(device="cuda", dtype=torch.float64).contiguous()
    return t


def _phi_device(Phi):
    """Phi (n, n), symmetric, numpy or torch -> contiguous float64 CUDA tensor
    (its row-major storage is the column-major matrix)."""
    if not isinstance(Phi, torch.Tensor):
        Phi = torch.from_numpy(np.ascontiguousarray(Phi, dtype=np.float64))
    return Phi.detach().to(device="cuda", dtype=torch.float64).contiguous()


def _genotypes(XR, n):
    """(m, ldx) CUDA float64 with ldx >= n and ldx % 4 == 0 (zero-padded copy
    when the given leading dimension does not qualify)."""
    XR = _as_f64_cuda(XR, "XR")
    if XR.shape[1] < n:
        raise FglsError(FGLS_ERR_ARG, "XR must be (m, ldx >= n)")
    if XR.shape[1] % 4 or XR.data_ptr() % 32:
        ld = (XR.shape[1] + 3) // 4 * 4
        out = torch.zeros((XR.shape[0], ld), dtype=torch.float64, device="cuda")
        out[:, :XR.shape[1]] = XR
        XR = out
    return XR


class SnpLoss(torch.autograd.Function):
    """l = sum_i w_i b_i[p-1]^2; see the module docstring.

    Arguments of ``apply``: y (n,), h2 (0-dim), XR (m, ldx) genotypes
    (SNP i in row i, ldx >= n; copied with zero padding unless ldx % 4 == 0),
    w (m,), Phi
    (n, n) symmetric, XL (wXL, n) covariates (row k = column k of X_L),
    sigma2 (float), block (SNPs per device call), precision ("fp64", or
    "mixed": int8 genotypes for the forward and backward products).
    """

    @staticmethod
    def forward(ctx, y, h2, XR, w, Phi, XL, sigma2: float = 1.0, block: int = 2048,
                precision: str = "fp64"):
        n = int(y.shape[0])
        wXL = int(XL.shape[0])
        m = int(XR.shape[0])
        if tuple(Phi.shape) != (n, n):
            raise FglsError(FGLS_ERR_ARG, f"Phi must be {n} x {n}")
        if XR.shape[1] < n or w.shape[0] != m:
            raise FglsError(FGLS_ERR_ARG, "XR must be (m, ldx >= n) and w (m,)")
        if precision not in ("fp64", "mixed"):
            raise FglsError(FGLS_ERR_ARG, "precision must be fp64 or mixed")
        if block < 1:
            raise FglsError(FGLS_ERR_ARG, "block must be positive")
        ctx.y_device, ctx.h2_like, ctx.dc, ctx.n = y.device, h2, None, n
        if m == 0:
            return torch.zeros((), dtype=torch.float64, device=y.device)
        Phid = _phi_device(Phi)
        dc = DeviceContext(n, wXL, 1, max_block=min(block, m))
        try:
            if precision == "mixed":
                dc.set_mode(FGLS_MODE_MIXED_I8)
            dc.set_trait_phi(Phid, float(h2), float(sigma2), _as_f64_cuda(XL, "XL"), _as_f64_cuda(y, "y"))
            XRd = _genotypes(XR, n)
            wd = _as_f64_cuda(w, "w")
            p = wXL + 1
            beta = torch.empty(m, dtype=torch.float64, device="cuda")
            B = torch.empty((dc.max_block, p), dtype=torch.float64, device="cuda")
            V = torch.empty((dc.max_block, p * p), dtype=torch.float64, device="cuda")
            X8 = dc.to_i8(XRd) if precision == "mixed" else None
            for i0 in range(0, m, dc.max_block):
                cnt = min(dc.max_block, m - i0)
                if X8 is not None:
                    dc.solve_block_i8(X8[i0:i0 + cnt], cnt, B, V, snp_base=i0)
                else:
                    dc.solve_block(XRd[i0:i0 + cnt], cnt, B, V, snp_base=i0)
                beta[i0:i0 + cnt] = B[:cnt, p - 1]
            dc.check()
        except BaseException:
            dc.close()
            raise
        loss = torch.dot(wd, beta * beta)
        ctx.dc = dc
        ctx.Phi = Phid
        ctx.XR, ctx.X8, ctx.w, ctx.sigma2 = XRd, X8, wd, float(sigma2)
        return loss.to(y.device)

    @staticmethod
    def backward(ctx, grad_out):
        dc = ctx.dc
        h2 = ctx.h2_like
        if dc is None:                        # m = 0: the loss does not depend on y or h2
            grad_y = torch.zeros(ctx.n, dtype=torch.float64, device=ctx.y_device)
            grad_h2 = torch.zeros_like(h2) if isinstance(h2, torch.Tensor) else None
            return grad_y, grad_h2, None, None, None, None, None, None, None
        ctx.dc = None
        try:
            m = ctx.XR.shape[0]
            dc.adjoint_begin(ctx.Phi, ctx.sigma2)
            for i0 in range(0, m, dc.max_block):
                cnt = min(dc.max_block, m - i0)
                if ctx.X8 is not None:
                    dc.adjoint_block_i8(ctx.X8[i0:i0 + cnt], ctx.w[i0:i0 + cnt], cnt, snp_base=i0)
                else:
                    dc.adjoint_block(ctx.XR[i0:i0 + cnt], ctx.w[i0:i0 + cnt], cnt, snp_base=i0)
            _, gy, gh = dc.adjoint_end()
            dc.check()
        finally:
            dc.close()
        go = grad_out.detach().to(device="cuda", dtype=torch.float64)
        grad_y = (gy * go).to(ctx.y_device)
        if isinstance(h2, torch.Tensor):
            grad_h2 = (go.to(h2.device) * gh).reshape(h2.shape).to(h2.dtype)
        else:
            grad_h2 = None
        return grad_y, grad_h2, None, None, None, None, None, None, None


def snp_loss(y, h2, XR, w, Phi, XL, sigma2: float = 1.0, block: int = 2048, precision: str = "fp64"):
    """Differentiable l(y, h2) = sum_i w_i b_i[p-1]^2 (see SnpLoss)."""
    return SnpLoss.apply(y, h2, XR, w, Phi, XL, sigma2, block, precision)


def snp_loss_grad(y, h2: float, XR, w, Phi, XL, sigma2: float = 1.0, block: int = 2048,
                  precision: str = "fp64"):
    """Loss and both gradients in one pass without autograd bookkeeping:
    returns (l, dl/dy (n,) CUDA tensor, dl/dh2).  M is built and factored on
    the device (fgls_hip_set_trait_phi).  precision="mixed" runs the
    per-block products on int8 genotypes (XR must hold integers in
    [-127, 127])."""
    if precision not in ("fp64", "mixed"):
        raise FglsError(FGLS_ERR_ARG, "precision must be fp64 or mixed")
    n = int(y.shape[0])
    m = int(XR.shape[0])
    if tuple(Phi.shape) != (n, n) or XR.shape[1] < n or w.shape[0] != m or block < 1:
        raise FglsError(FGLS_ERR_ARG, "Phi must be (n, n), XR (m, ldx >= n), w (m,) and block > 0")
    if m == 0:
        return 0.0, torch.zeros(n, dtype=torch.float64, device="cuda"), 0.0
    Phid = _phi_device(Phi)
    dc = DeviceContext(n, int(XL.shape[0]), 1, max_block=min(block, m))
    try:
        if precision == "mixed":
            dc.set_mode(FGLS_MODE_MIXED_I8)
        dc.set_trait_phi(Phid, float(h2), float(sigma2), _as_f64_cuda(XL, "XL"), _as_f64_cuda(y, "y"))
        dc.adjoint_begin(Phid, float(sigma2))
        XRd = _genotypes(XR, n)
        wd = _as_f64_cuda(w, "w")
        X8 = dc.to_i8(XRd) if precision == "mixed" else None
        for i0 in range(0, m, dc.max_block):
            cnt = min(dc.max_block, m - i0)
            if X8 is not None:
                dc.adjoint_block_i8(X8[i0:i0 + cnt], wd[i0:i0 + cnt], cnt, snp_base=i0)
            else:
                dc.adjoint_block(XRd[i0:i0 + cnt], wd[i0:i0 + cnt], cnt, snp_base=i0)
        loss, gy, gh = dc.adjoint_end()
        dc.check()
        return loss, gy, gh
    finally:
        dc.close()
