"""Differentiable GWAS scan loss over the FGLS solver.

    l(y, h2) = sum_i w_i * b_i[p-1]**2

where b_i is the per-SNP GLS estimate of the reference's Cholesky FGLS path
(fgls_chol.c:386-439; M = sigma2 (h2 Phi + (1 - h2) I), fgls_chol_gpu.c:406-414)
and b_i[p-1] is the SNP's own coefficient.  The reference computes no
gradients; this is an addition on top of its forward path.

``SnpLoss`` is a ``torch.autograd.Function`` over the C-ABI:

* forward: host Cholesky of M (``fgls_factor_M``), per-trait precompute and
  the fused block solve on the device (``fgls_hip_solve_block``); nothing of
  L^-1 X_R is kept;
* backward: ``fgls_hip_adjoint_begin/block/end`` recompute L^-1 X_R block by
  block and return dl/dy and dl/dh2.

With ``precision="mixed"`` the genotypes go to the device as int8 and the
block products of both passes run on the int8 MFMA
(``fgls_hip_solve_block_i8``, ``fgls_hip_adjoint_block_i8``).  Gradients
flow to ``y`` and ``h2`` only.
"""
from __future__ import annotations

import numpy as np
import torch

from .fgls import DeviceContext, FglsError, FGLS_ERR_ARG, FGLS_MODE_MIXED_I8, factor_M


def _as_f64_cuda(t, name):
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    t = t.detach().to(device="cuda", dtype=torch.float64).contiguous()
    return t


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
        Phi_h = Phi.detach().cpu().numpy() if isinstance(Phi, torch.Tensor) else np.asarray(Phi)
        if Phi_h.shape != (n, n):
            raise FglsError(FGLS_ERR_ARG, f"Phi must be {n} x {n}")
        wXL = int(XL.shape[0])
        m = int(XR.shape[0])
        if XR.shape[1] < n or w.shape[0] != m:
            raise FglsError(FGLS_ERR_ARG, "XR must be (m, ldx >= n) and w (m,)")
        if precision not in ("fp64", "mixed"):
            raise FglsError(FGLS_ERR_ARG, "precision must be fp64 or mixed")
        h2v = float(h2)
        L = factor_M(Phi_h, h2v, float(sigma2))
        dc = DeviceContext(n, wXL, 1, max_block=min(block, m))
        if precision == "mixed":
            dc.set_mode(FGLS_MODE_MIXED_I8)
        Ld = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()      # column-major (n, n)
        XLd = _as_f64_cuda(XL, "XL")
        yd = _as_f64_cuda(y, "y")
        XRd = _genotypes(XR, n)
        wd = _as_f64_cuda(w, "w")
        dc.set_trait(Ld, XLd, yd)
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
        loss = torch.dot(wd, beta * beta)
        ctx.dc = dc
        ctx.Phi = torch.from_numpy(np.ascontiguousarray(Phi_h.T)).cuda()
        ctx.XR, ctx.X8, ctx.w, ctx.sigma2, ctx.h2_like = XRd, X8, wd, float(sigma2), h2
        ctx.y_device = y.device
        return loss.to(y.device)

    @staticmethod
    def backward(ctx, grad_out):
        dc = ctx.dc
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
        go = grad_out.detach().to(device="cuda", dtype=torch.float64)
        grad_y = (gy * go).to(ctx.y_device)
        h2 = ctx.h2_like
        if isinstance(h2, torch.Tensor):
            grad_h2 = (go.to(h2.device) * gh).reshape(h2.shape).to(h2.dtype)
        else:
            grad_h2 = None
        ctx.dc = None
        dc.close()
        return grad_y, grad_h2, None, None, None, None, None, None, None


def snp_loss(y, h2, XR, w, Phi, XL, sigma2: float = 1.0, block: int = 2048, precision: str = "fp64"):
    """Differentiable l(y, h2) = sum_i w_i b_i[p-1]^2 (see SnpLoss)."""
    return SnpLoss.apply(y, h2, XR, w, Phi, XL, sigma2, block, precision)


def snp_loss_grad(y, h2: float, XR, w, Phi, XL, sigma2: float = 1.0, block: int = 2048,
                  precision: str = "fp64"):
    """Loss and both gradients in one pass without autograd bookkeeping:
    returns (l, dl/dy (n,) CUDA tensor, dl/dh2).  precision="mixed" runs the
    per-block products on int8 genotypes (XR must hold integers in
    [-127, 127])."""
    if precision not in ("fp64", "mixed"):
        raise FglsError(FGLS_ERR_ARG, "precision must be fp64 or mixed")
    n = int(y.shape[0])
    Phi_h = Phi.detach().cpu().numpy() if isinstance(Phi, torch.Tensor) else np.asarray(Phi)
    L = factor_M(Phi_h, float(h2), float(sigma2))
    m = int(XR.shape[0])
    dc = DeviceContext(n, int(XL.shape[0]), 1, max_block=min(block, m))
    try:
        if precision == "mixed":
            dc.set_mode(FGLS_MODE_MIXED_I8)
        dc.set_trait(torch.from_numpy(np.ascontiguousarray(L.T)).cuda(), _as_f64_cuda(XL, "XL"),
                     _as_f64_cuda(y, "y"))
        dc.adjoint_begin(torch.from_numpy(np.ascontiguousarray(Phi_h.T)).cuda(), float(sigma2))
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
