"""Dense FP64 torch restatement of the GWAS scan loss, for torch.autograd.

l(y, h) = sum_i w_i b_i[p-1]^2 with b_i = (X_i' M^-1 X_i)^-1 X_i' M^-1 y,
X_i = [X_L, x_i], M = sigma (h Phi + (1 - h) I) (fgls_chol.c:251-257 and
:386-439), computed with an explicit inverse of M and batched dense solves --
a different route from the oracle and the device code, used only to check
their gradients.
"""
from __future__ import annotations

import numpy as np
import torch


def dense_loss(y, h, Phi, XL, XR, w, sigma=1.0):
    """y: (n,) tensor, h: 0-dim tensor; Phi n x n, XL n x wXL, XR n x m,
    w (m,) as float64 CPU tensors.  Returns the 0-dim loss tensor.

    S_i = X_i' M^-1 X_i and r_i = X_i' M^-1 y are assembled from the two
    dense products M^-1 X_L and M^-1 X_R (blocks shared by every SNP, and one
    column per SNP), then solved per SNP."""
    n = Phi.shape[0]
    M = sigma * (h * Phi + (1.0 - h) * torch.eye(n, dtype=torch.float64))
    Minv = torch.linalg.inv(M)
    m, wXL = XR.shape[1], XL.shape[1]
    ML, MR = Minv @ XL, Minv @ XR                                  # n x wXL, n x m
    A = (XL.T @ ML).expand(m, wXL, wXL)                            # X_L' M^-1 X_L
    c = XR.T @ ML                                                  # m x wXL: x_i' M^-1 X_L
    q = (XR * MR).sum(0)                                           # m: x_i' M^-1 x_i
    S = torch.cat([torch.cat([A, c.unsqueeze(2)], 2),
                   torch.cat([c.unsqueeze(1), q.view(m, 1, 1)], 2)], 1)   # m x p x p
    r = torch.cat([(ML.T @ y).expand(m, wXL), (MR.T @ y).unsqueeze(1)], 1).unsqueeze(2)
    b = torch.linalg.solve(S, r)[..., 0]
    beta = b[:, -1]
    return (w * beta * beta).sum()


def dense_grads(y, h, Phi, XL, XR, w, sigma=1.0):
    """(loss, dl/dy, dl/dh) by torch.autograd on dense_loss (numpy in/out)."""
    t = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
    yt = t(y).requires_grad_(True)
    ht = torch.tensor(float(h), dtype=torch.float64, requires_grad=True)
    loss = dense_loss(yt, ht, t(Phi), t(XL), t(XR), t(w), sigma)
    gy, gh = torch.autograd.grad(loss, (yt, ht))
    return float(loss.detach()), gy.detach().numpy(), float(gh.detach())
