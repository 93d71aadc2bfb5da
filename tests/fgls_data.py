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
