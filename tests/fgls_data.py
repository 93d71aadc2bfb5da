"""Seeded synthetic FGLS problems with the generators of the benchmark configs.

Phi = Z Z' / (2n) with Z ~ N(0,1) of shape n x 2n, rescaled to unit diagonal;
X_L = [1, N(0,1) ...] (the intercept is written into X_L: the reference adds
none, fgls_chol.c:146-148); X_R dosages ~ Binomial(2, f_j), f_j ~ U(0.05, 0.5)
(or 3-decimal U(0,2) dosages); y = X_L beta + eps with eps ~ N(0, M),
M = h2 Phi + (1 - h2) I.  All draws come from one numpy Generator.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


@dataclass
class Problem:
    n: int
    wXL: int
    m: int
    h2: float
    Phi: np.ndarray      # n x n
    XL: np.ndarray       # n x wXL
    XR: np.ndarray       # n x m
    y: np.ndarray        # n
    Y: np.ndarray        # n x t


def make_phi(rng, n):
    Z = rng.standard_normal((n, 2 * n))
    Phi = (Z @ Z.T) / (2 * n)
    d = np.sqrt(np.diag(Phi))
    Phi = Phi / d[:, None] / d[None, :]
    np.fill_diagonal(Phi, 1.0)
    return Phi


def make_problem(n=1000, wXL=3, m=2000, seed=0, h2=0.4, t=1, dosage="binomial", wXR=1):
    rng = np.random.default_rng(seed)
    Phi = make_phi(rng, n)
    XL = np.empty((n, wXL))
    if wXL > 0:
        XL[:, 0] = 1.0
        XL[:, 1:] = rng.standard_normal((n, wXL - 1))
    cols = m * wXR
    if dosage == "binomial":
        f = rng.uniform(0.05, 0.5, size=cols)
        XR = rng.binomial(2, f[None, :], size=(n, cols)).astype(np.float64)
    elif dosage == "imputed":
        XR = np.round(rng.uniform(0.0, 2.0, size=(n, cols)), 3)
    else:
        raise ValueError(dosage)
    M = h2 * Phi + (1.0 - h2) * np.eye(n)
    Lm = np.linalg.cholesky(M)
    Y = np.empty((n, t))
    for j in range(t):
        beta = rng.standard_normal(wXL)
        Y[:, j] = XL @ beta + Lm @ rng.standard_normal(n)
    return Problem(n=n, wXL=wXL, m=m, h2=h2, Phi=Phi, XL=XL, XR=XR, y=Y[:, 0].copy(), Y=Y)


def write_problem_dir(prob: Problem, d: str, H=None, Sig=None):
    """Write the reference's input files (driver.c:56-66) for `prob`."""
    os.makedirs(d, exist_ok=True)
    t = prob.Y.shape[1]
    H = np.full(t, prob.h2) if H is None else np.asarray(H, dtype=np.float64)
    Sig = np.ones(t) if Sig is None else np.asarray(Sig, dtype=np.float64)
    files = {"Phi.in": prob.Phi.ravel(order="F"), "H.in": H, "Sig.in": Sig,
             "XL.in": prob.XL.ravel(order="F"), "XR.in": prob.XR.ravel(order="F"),
             "Y.in": prob.Y.ravel(order="F")}
    for name, arr in files.items():
        np.asarray(arr, dtype="<f8").tofile(os.path.join(d, name))


@dataclass
class DeviceProblem:
    """A configuration-scale problem drawn on the device (torch, float64).

    Phi (n, n) symmetric (its row-major storage is the column-major matrix);
    XL (wXL, n), row k = covariate column k; Y (t, n), row j = trait j;
    H (t,) the traits' h2; X8 (m, ld8) int8 genotypes (column-major blocks,
    rows >= n zero) or, for dosages, V16 (m, n) int16 values 1000 x;
    X (m, ldx) the exact FP64 values (ldx = ceil(n/4)*4, rows >= n zero)."""
    n: int
    wXL: int
    m: int
    H: list
    Phi: object
    XL: object
    Y: object
    X: object
    X8: object = None
    V16: object = None

    def host_phi(self):
        return self.Phi.cpu().numpy()

    def host_xl(self):
        return self.XL.T.cpu().numpy().copy()

    def host_x(self, idx):
        """n x len(idx) FP64 columns of the selected SNPs (numpy)."""
        import torch
        sel = torch.as_tensor(np.asarray(idx), device=self.X.device)
        return self.X[sel, :self.n].cpu().numpy().T.copy()


def make_problem_device(n, wXL, m, seed, H, dosage="binomial", device="cuda"):
    """The configurations' generators at their full n, drawn with a seeded
    torch generator on the device (the n x 2n product for Phi is 1e14 flop at
    n = 30,000: minutes on the host, seconds on the device):
    Phi = Z Z' / (2n), Z ~ N(0,1) n x 2n, rescaled to unit diagonal;
    X_L = [1, N(0,1) ...]; X_R ~ Binomial(2, f_j), f_j ~ U(0.05, 0.5)
    (dosage="binomial") or round(U(0, 2), 3) (dosage="imputed");
    y_j = X_L beta_j + L_j eps_j with M_j = h_j Phi + (1 - h_j) I = L_j L_j'
    (torch's Cholesky, independent of the code under test)."""
    import torch
    f64 = torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    Z = torch.randn((n, 2 * n), dtype=f64, device=device, generator=g)
    Phi = (Z @ Z.T) / (2 * n)
    del Z
    Phi = 0.5 * (Phi + Phi.T)
    d = Phi.diagonal().sqrt()
    Phi = Phi / d[:, None] / d[None, :]
    Phi = torch.tril(Phi) + torch.tril(Phi, -1).T     # exactly symmetric (the scaling rounds per side)
    Phi.fill_diagonal_(1.0)
    XL = torch.empty((wXL, n), dtype=f64, device=device)
    if wXL > 0:
        XL[0] = 1.0
        XL[1:] = torch.randn((wXL - 1, n), dtype=f64, device=device, generator=g)
    H = [float(h) for h in H]
    Y = torch.empty((len(H), n), dtype=f64, device=device)
    for j, h in enumerate(H):
        M = h * Phi
        M.diagonal().add_(1.0 - h)
        L = torch.linalg.cholesky(M)
        del M
        beta = torch.randn(wXL, dtype=f64, device=device, generator=g)
        Y[j] = XL.T @ beta + L @ torch.randn(n, dtype=f64, device=device, generator=g)
        del L
    ldx = (n + 3) // 4 * 4
    ld8 = (n + 63) // 64 * 64
    X = torch.zeros((m, ldx), dtype=f64, device=device)
    X8 = V16 = None
    if dosage == "binomial":
        X8 = torch.zeros((m, ld8), dtype=torch.int8, device=device)
        f = torch.rand(m, dtype=torch.float32, device=device, generator=g) * 0.45 + 0.05
        C = 8192
        for i0 in range(0, m, C):
            fc = f[i0:i0 + C, None]
            cnt = fc.shape[0]
            a = torch.rand((cnt, n), dtype=torch.float32, device=device, generator=g) < fc
            b = torch.rand((cnt, n), dtype=torch.float32, device=device, generator=g) < fc
            X8[i0:i0 + cnt, :n] = a.to(torch.int8) + b.to(torch.int8)
            X[i0:i0 + cnt, :n] = X8[i0:i0 + cnt, :n].to(f64)
    elif dosage == "imputed":
        V16 = torch.randint(0, 2001, (m, n), dtype=torch.int16, device=device, generator=g)
        # the nearest double to v / 1000, as round(U, 3) gives: a true division
        # (torch divides by a Python scalar as a product with its reciprocal,
        # which is one ulp off for some v)
        C = 8192
        for i0 in range(0, m, C):
            v = V16[i0:i0 + C].to(f64)
            X[i0:i0 + C, :n] = v / torch.full_like(v, 1000.0)
    else:
        raise ValueError(dosage)
    return DeviceProblem(n=n, wXL=wXL, m=m, H=H, Phi=Phi, XL=XL, Y=Y, X=X, X8=X8, V16=V16)


def write_device_problem_dir(prob: DeviceProblem, d: str, xr_formats=("f64",), Sig=None, m=None):
    """The reference's input files (driver.c:56-66) for a device problem (the
    first m SNPs, default all); X_R in each requested format: f64 -> XR.in
    (8 n bytes per SNP), i8 -> XR8.in, u2 -> XR2.in (2-bit, pack_u2's
    layout), i16 -> XR16.in (1000 x)."""
    os.makedirs(d, exist_ok=True)
    t = len(prob.H)
    n = prob.n
    m = prob.m if m is None else min(m, prob.m)
    Sig = np.ones(t) if Sig is None else np.asarray(Sig, dtype=np.float64)
    np.ascontiguousarray(prob.Phi.cpu().numpy(), dtype="<f8").tofile(os.path.join(d, "Phi.in"))
    np.asarray(prob.H, dtype="<f8").tofile(os.path.join(d, "H.in"))
    Sig.astype("<f8").tofile(os.path.join(d, "Sig.in"))
    prob.XL.cpu().numpy().astype("<f8").tofile(os.path.join(d, "XL.in"))
    prob.Y.cpu().numpy().astype("<f8").tofile(os.path.join(d, "Y.in"))
    C = 8192
    for fmt in xr_formats:
        name = {"f64": "XR.in", "i8": "XR8.in", "u2": "XR2.in", "i16": "XR16.in"}[fmt]
        with open(os.path.join(d, name), "wb") as fh:
            for i0 in range(0, m, C):
                i1 = min(i0 + C, m)
                if fmt == "f64":
                    fh.write(np.ascontiguousarray(prob.X[i0:i1, :n].cpu().numpy(), dtype="<f8").tobytes())
                elif fmt == "i8":
                    fh.write(np.ascontiguousarray(prob.X8[i0:i1, :n].cpu().numpy()).tobytes())
                elif fmt == "u2":
                    fh.write(pack_u2_rows(prob.X8[i0:i1, :n].cpu().numpy()).tobytes())
                else:
                    fh.write(np.ascontiguousarray(prob.V16[i0:i1].cpu().numpy(), dtype="<i2").tobytes())


def pack_u2_rows(x):
    """(cols, n) integer genotypes 0/1/2, one SNP per row -> (cols, ceil(n/4))
    bytes of the 2-bit layout (individual r in bits 2(r%4)..2(r%4)+1 of byte
    r/4)."""
    cols, n = x.shape
    nb = (n + 3) // 4
    pad = np.zeros((cols, nb * 4), dtype=np.uint8)
    pad[:, :n] = x.astype(np.uint8)
    q = pad.reshape(cols, nb, 4)
    return q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)
