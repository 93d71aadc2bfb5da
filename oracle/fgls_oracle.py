"""FP64 CPU oracle of the Cholesky FGLS path -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the MI355X implementation.  Only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
it; the product path never calls it and has no CPU fallback.

It restates, with numpy/scipy LAPACK calls in the same order, the reference
solver ``fgls_chol`` (src/fgls_chol.c) and the identical arithmetic of the GPU
variant (src/fgls_chol_gpu.c):

* ``build_M``           fgls_chol.c:251-257  (M = sigma*(h*Phi) + sigma*(1-h) I)
* ``trait_precompute``  fgls_chol.c:259-325  (dpotrf, dtrsm, dtrsv, dgemv, dsyrk,
                                              null-model res_sigma)
* ``solve_snps``        fgls_chol.c:365-368  (dtrsm of the X_R block) and
                        fgls_chol.c:386-439  (per-SNP stage)
* ``fgls_chol``         fgls_chol.c:223-491  (trait loop, output layout
                                              fgls_chol.c:458-474)
* ``read_doubles`` / ``read_inputs``  io.c:7-19 raw little-endian doubles (the
                        output files are written by the tests with numpy, io.c:21-33).
* ``eigen_decompose``, ``eigen_trait``, ``eigen_solve_snps``, ``fgls_eigen``
                        the eigenbasis solver src/fgls_eigen.c: dsyevr of Phi
                        (:780-806), Winv = sqrt(1/(sigma h W + sigma (1-h)))
                        (:455-468), the scaled Z'X_L, Z'y, B_t, V_tl (:470-491),
                        residual_sigma (:997-1152) and the per-SNP loop
                        (:494-563); the per-SNP stage is per_snp_from_z.
* ``snp_loss_adjoint``  no reference counterpart: gradients of
                        l = sum_i w_i b_i[p-1]^2 w.r.t. y and h (an addition
                        on top of the per-SNP stage); pinned only by
                        torch.autograd through a dense FP64 torch restatement
                        and by central finite differences (tests/test_oracle.py).

How each function is pinned (beyond this restatement itself) is listed in
tests/test_oracle.py: the reference fixture test_cases/test_XL/B_exp.out
(B only, through a synthetic X_R), closed forms (brute-force inverse of M at
small n, GLS normal equations), invariances (sigma scale, SNP scaling and
permutation, intercept shift), and LAPACK identities.  V has no reference
fixture: its layout and values are pinned by the closed form
V = res_sigma * (X' M^-1 X)^-1 (dense brute force) only.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla


@dataclass
class Trait:
    """Per-trait quantities of fgls_chol.c:259-325."""

    L: np.ndarray          # lower Cholesky factor of M
    XLt: np.ndarray        # L^-1 X_L  (n x wXL)
    yt: np.ndarray         # L^-1 y    (n)
    B_t: np.ndarray        # X~_L' y~  (wXL)
    V_tl: np.ndarray       # lower(X~_L' X~_L), upper 0 (wXL x wXL)
    res_sigma: float


def build_M(Phi: np.ndarray, h: float, sigma: float) -> np.ndarray:
    """M := sigma*h*Phi; M_ii += sigma*(1-h)  (fgls_chol.c:251-257).

    H.in holds h^2 directly; the code uses it linearly (the comment at
    fgls_chol.c:251 is misleading, the arithmetic is what counts)."""
    M = (h * sigma) * np.array(Phi, dtype=np.float64, copy=True)
    M[np.diag_indices_from(M)] += (1.0 - h) * sigma
    return M


def cholesky_lower(M: np.ndarray) -> np.ndarray:
    """dpotrf_("L") (fgls_chol.c:262); raises on a non-SPD matrix."""
    return sla.cholesky(M, lower=True, check_finite=False)


def trait_precompute(L: np.ndarray, XL: np.ndarray, y: np.ndarray) -> Trait:
    """fgls_chol.c:278-325 for one trait.

    XL is n x wXL (column k = covariate k), y has n entries."""
    n = L.shape[0]
    wXL = XL.shape[1]
    XLt = sla.solve_triangular(L, XL, lower=True, check_finite=False)   # dtrsm  :279
    yt = sla.solve_triangular(L, y, lower=True, check_finite=False)     # dtrsv  :291
    B_t = XLt.T @ yt                                                    # dgemv  :294
    V_tl = np.tril(XLt.T @ XLt)                                         # dsyrk  :297 (upper stays 0)
    # sigma2.score, fgls_chol.c:301-325
    if wXL > 0:
        C = np.linalg.cholesky(V_tl + np.tril(V_tl, -1).T)              # dpotrf :307
        bL = sla.solve_triangular(C, B_t, lower=True)                   # dtrsv  :314
        bL = sla.solve_triangular(C.T, bL, lower=False)                 # dtrsv  :315
        r = y - XL @ bL                                                 # dgemv  :317-320
    else:
        r = np.array(y, dtype=np.float64, copy=True)
    rt = sla.solve_triangular(L, r, lower=True, check_finite=False)     # dtrsv  :322-323
    res = float(rt @ rt) / (n - wXL)                                    # ddot   :324
    return Trait(L=L, XLt=XLt, yt=yt, B_t=B_t, V_tl=V_tl, res_sigma=res)


def per_snp_from_z(Z: np.ndarray, tr: Trait, wXR: int):
    """Per-SNP stage (fgls_chol.c:386-439) for transformed columns Z = L^-1 X_R.

    Z is n x (count*wXR), SNP i's columns at [i*wXR, (i+1)*wXR).
    Returns B (count x p) and V (count x p x p) where V[i] is stored so that
    V[i].ravel(order="F") is the on-disk order (column-major, lower)."""
    n = Z.shape[0]
    wXL = tr.XLt.shape[1]
    p = wXL + wXR
    count = Z.shape[1] // wXR
    Zs = Z.reshape(n, count, wXR).transpose(1, 0, 2)           # count x n x wXR
    S = np.zeros((count, p, p))
    S[:, :wXL, :wXL] = tr.V_tl + np.tril(tr.V_tl, -1).T        # V_TL (symmetrised copy)
    VBL = np.einsum("cna,nk->cak", Zs, tr.XLt)                 # dgemm "T","N" :409-412
    S[:, wXL:, :wXL] = VBL
    S[:, :wXL, wXL:] = VBL.transpose(0, 2, 1)
    S[:, wXL:, wXL:] = np.einsum("cna,cnb->cab", Zs, Zs)       # dsyrk :414-417
    rhs = np.empty((count, p))
    rhs[:, :wXL] = tr.B_t                                      # memcpy B_t :398
    rhs[:, wXL:] = np.einsum("cna,n->ca", Zs, tr.yt)           # dgemv "T" :400-403
    C = np.linalg.cholesky(S)                                  # dpotrf :420
    b = np.linalg.solve(C, rhs[..., None])                     # dtrsv N :427
    b = np.linalg.solve(np.swapaxes(C, 1, 2), b)[..., 0]       # dtrsv T :428
    Ci = np.linalg.inv(C)
    Sinv = np.swapaxes(Ci, 1, 2) @ Ci                          # dpotri :431
    V = tr.res_sigma * np.tril(Sinv)                           # dscal over p^2, upper stays 0
    idx = np.arange(p)
    V[:, idx, idx] = np.sqrt(V[:, idx, idx])                   # sqrt of the diagonal :437-438
    return b, V


def solve_snps(XR_block: np.ndarray, tr: Trait, wXR: int):
    """dtrsm of one X_R block (fgls_chol.c:365-368) followed by the per-SNP
    stage.  XR_block is n x (count*wXR) column-major data as a numpy array."""
    Z = sla.solve_triangular(tr.L, XR_block, lower=True, check_finite=False)
    return per_snp_from_z(Z, tr, wXR)


def fgls_chol(Phi, H, Sig, XL, XR, Y, wXR: int = 1, x_b: int | None = None):
    """Whole solver over t traits, outputs in the reference layout.

    XL: n x wXL, XR: n x (m*wXR), Y: n x t.  Returns flat B (t*m*p) and
    V (t*m*p*p) exactly as B.out / V.out are laid out (fgls_chol.c:458-474):
    B[(j*m + i)*p + k], V[(j*m + i)*p*p + c*p + r]."""
    n = Phi.shape[0]
    wXL = XL.shape[1]
    p = wXL + wXR
    m = XR.shape[1] // wXR
    t = Y.shape[1]
    x_b = m if x_b is None else x_b
    Bout = np.empty(t * m * p)
    Vout = np.empty(t * m * p * p)
    for j in range(t):
        L = cholesky_lower(build_M(Phi, float(H[j]), float(Sig[j])))
        tr = trait_precompute(L, XL, Y[:, j])
        for ib in range(0, m, x_b):
            cnt = min(x_b, m - ib)
            b, V = solve_snps(XR[:, ib * wXR:(ib + cnt) * wXR], tr, wXR)
            o = (j * m + ib) * p
            Bout[o:o + cnt * p] = b.ravel()
            ov = (j * m + ib) * p * p
            Vout[ov:ov + cnt * p * p] = V.transpose(0, 2, 1).ravel()
    return Bout, Vout


def eigen_decompose(Phi: np.ndarray):
    """eigenDec (fgls_eigen.c:780-806): dsyevr("V", "A", "L") of Phi.
    Returns (Z, W) with Phi = Z diag(W) Z', W ascending."""
    W, Z = sla.eigh(np.asarray(Phi, dtype=np.float64), lower=True, driver="evr", check_finite=False)
    return Z, W


def eigen_trait(Z: np.ndarray, W: np.ndarray, XL: np.ndarray, y: np.ndarray, h: float,
                sigma: float) -> Trait:
    """Per-trait quantities of the eigenbasis solver for one trait.

    Winv = sqrt(1 / (sigma h W + sigma (1 - h)))          fgls_eigen.c:455-468
    X~_L = Winv * Z'X_L, y~ = Winv * Z'y                   :476-487
    B_t = X~_L' y~, V_tl = lower(X~_L' X~_L)               :490-500
    res_sigma: b_L from V_tl, r = y - X_L b_L,
               |Winv * Z'r|^2 / (n - wXL)                  residual_sigma, :1119-1152
    The returned Trait has L = None; XLt / yt hold the scaled X~_L, y~."""
    n = Z.shape[0]
    wXL = XL.shape[1]
    alpha, beta = sigma * h, sigma * (1.0 - h)
    Winv = np.sqrt(1.0 / (alpha * W + beta))
    XLt = Winv[:, None] * (Z.T @ XL)
    yt = Winv * (Z.T @ y)
    B_t = XLt.T @ yt
    V_tl = np.tril(XLt.T @ XLt)
    if wXL > 0:
        C = np.linalg.cholesky(V_tl + np.tril(V_tl, -1).T)
        bL = sla.solve_triangular(C, B_t, lower=True)
        bL = sla.solve_triangular(C.T, bL, lower=False)
        r = y - XL @ bL
    else:
        r = np.array(y, dtype=np.float64, copy=True)
    rt = Winv * (Z.T @ r)
    res = float(rt @ rt) / (n - wXL)
    tr = Trait(L=None, XLt=XLt, yt=yt, B_t=B_t, V_tl=V_tl, res_sigma=res)
    tr.Winv = Winv
    return tr


def eigen_solve_snps(Z: np.ndarray, tr: Trait, XR_block: np.ndarray, wXR: int):
    """x~ = Winv * Z'x for the columns of one block (the out-of-core gemm,
    fgls_eigen.c:727-757, and the Winv scaling, :506-509), then the per-SNP
    stage (:510-563, the same arithmetic as fgls_chol.c:386-439)."""
    return per_snp_from_z(tr.Winv[:, None] * (Z.T @ XR_block), tr, wXR)


def fgls_eigen(Phi, H, Sig, XL, XR, Y, wXR: int = 1, x_b: int | None = None):
    """The eigenbasis solver over t traits; arguments and output layout as
    fgls_chol (fgls_eigen.c:591-611 writes the same offsets)."""
    n = Phi.shape[0]
    wXL = XL.shape[1]
    p = wXL + wXR
    m = XR.shape[1] // wXR
    t = Y.shape[1]
    x_b = m if x_b is None else x_b
    Z, W = eigen_decompose(Phi)
    Bout = np.empty(t * m * p)
    Vout = np.empty(t * m * p * p)
    for j in range(t):
        tr = eigen_trait(Z, W, XL, Y[:, j], float(H[j]), float(Sig[j]))
        for ib in range(0, m, x_b):
            cnt = min(x_b, m - ib)
            b, V = eigen_solve_snps(Z, tr, XR[:, ib * wXR:(ib + cnt) * wXR], wXR)
            o = (j * m + ib) * p
            Bout[o:o + cnt * p] = b.ravel()
            ov = (j * m + ib) * p * p
            Vout[ov:ov + cnt * p * p] = V.transpose(0, 2, 1).ravel()
    return Bout, Vout


def snp_loss_adjoint(Phi, h: float, sigma: float, XL, XR, y, w):
    """l = sum_i w_i b_i[p-1]^2 and (dl/dy, dl/dh), wXR = 1.

    XL is n x wXL, XR n x m (column i = SNP i).  With X_i = [X_L, x_i],
    S_i = X_i' M^-1 X_i, u_i = S_i^-1 e_p, g_i = 2 w_i b_i[p-1] u_i,
    r_i = y - X_i b_i and G = L^-1 (Phi - I) L^-T:
        dl/dy = M^-1 sum_i X_i g_i = L^-T sum_i p~_i,   p~_i = L^-1 X_i g_i
        dl/dh = -sigma sum_i p~_i' G q~_i,              q~_i = L^-1 r_i
    (from d b_i = -S_i^-1 X_i' M^-1 dM M^-1 r_i and dM = sigma (Phi - I) dh)."""
    Phi = np.asarray(Phi, dtype=np.float64)
    XL = np.asarray(XL, dtype=np.float64).reshape(Phi.shape[0], -1)
    XR = np.asarray(XR, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    n, wXL = XL.shape
    p = wXL + 1
    L = cholesky_lower(build_M(Phi, h, sigma))
    tr = trait_precompute(L, XL, y)
    Z = sla.solve_triangular(L, XR, lower=True, check_finite=False)     # n x m
    b, V = per_snp_from_z(Z, tr, 1)
    m = Z.shape[1]
    S = np.zeros((m, p, p))
    S[:, :wXL, :wXL] = tr.XLt.T @ tr.XLt
    S[:, wXL, :wXL] = (Z.T @ tr.XLt)
    S[:, :wXL, wXL] = S[:, wXL, :wXL]
    S[:, wXL, wXL] = np.einsum("nm,nm->m", Z, Z)
    u = np.linalg.solve(S, np.broadcast_to(np.eye(p)[:, -1], (m, p))[..., None])[..., 0]
    beta = b[:, -1]
    g = 2.0 * (w * beta)[:, None] * u                                   # m x p
    Pt = tr.XLt @ g[:, :wXL].T + Z * g[:, -1]                           # n x m
    Qt = tr.yt[:, None] - tr.XLt @ b[:, :wXL].T - Z * beta              # n x m
    grad_y = sla.solve_triangular(L.T, Pt.sum(axis=1), lower=False, check_finite=False)
    T = sla.solve_triangular(L, Phi - np.eye(n), lower=True, check_finite=False)
    G = sla.solve_triangular(L, T.T, lower=True, check_finite=False)
    grad_h = -sigma * float(np.einsum("nm,nm->", Pt, G @ Qt))
    loss = float(w @ (beta * beta))
    return loss, grad_y, grad_h


def read_doubles(path: str, count: int, offset: int = 0) -> np.ndarray:
    """sync_read (io.c:7-19): raw little-endian doubles, no header."""
    a = np.fromfile(path, dtype="<f8", count=count, offset=offset * 8)
    if a.size != count:
        raise IOError(f"{path}: read {a.size} elements - {count} expected")
    return a


def read_inputs(d: str, n: int, p: int, m: int, t: int, wXL: int):
    """The driver's fixed file names (driver.c:56-66) and layouts."""
    wXR = p - wXL
    Phi = read_doubles(os.path.join(d, "Phi.in"), n * n).reshape(n, n, order="F")
    H = read_doubles(os.path.join(d, "H.in"), t)
    Sig = read_doubles(os.path.join(d, "Sig.in"), t)
    XL = read_doubles(os.path.join(d, "XL.in"), wXL * n).reshape(n, wXL, order="F")
    XR = read_doubles(os.path.join(d, "XR.in"), m * wXR * n).reshape(n, m * wXR, order="F")
    Y = read_doubles(os.path.join(d, "Y.in"), t * n).reshape(n, t, order="F")
    return Phi, H, Sig, XL, XR, Y


def max_rel_err(a: np.ndarray, b: np.ndarray, p: int | None = None, lower_only: bool = False) -> float:
    """driver.c:159-177: max_i |a_i - b_i| / |b_i|.

    With lower_only the comparison of a flat V is restricted to the lower
    triangle of each p x p matrix (the strict upper is 0 in both)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if lower_only:
        assert p is not None
        r, c = np.tril_indices(p)
        mask = np.zeros((p, p), dtype=bool)
        mask[c, r] = True            # flat index c*p + r with r >= c
        sel = np.tile(mask.ravel(), a.size // (p * p))
        a, b = a[sel], b[sel]
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.abs(a - b) / np.abs(b)
    rel = rel[~np.isnan(rel)]      # 0/0 is skipped, as NaN > max is false in driver.c:171
    return float(rel.max()) if rel.size else 0.0


def parity_errors(B, V, Bref, Vref, p: int) -> dict:
    """Errors of (B, V) against (Bref, Vref), both in the on-disk layout
    (B: ... x p, V: ... x p*p with entry c*p + r).

    * ``B_rel``/``V_rel``: the reference driver's element-wise metric
      (driver.c:159-177; V over its lower triangle).  Entries that are tiny
      against their natural scale (coefficients near 0, near-zero
      covariances) make it grow without bound even between two correct FP64
      codes -- the reference's own fgls_chol against a LAPACK-faithful
      restatement reaches 2.4e-10 on V this way.
    * ``B_scaled``: max |b - b_ref| / max(|b_ref|, se), se the coefficient's
      standard error (the diagonal of V_ref).
    * ``V_scaled``: diagonal (standard errors) element-wise; off-diagonal
      covariances as |V - V_ref| / max(|V_ref|, se_r * se_c).
    * ``var_rel``: element-wise relative error of the variances se^2, i.e.
      of var(b_ik), never near zero.
    * ``B_norm``: per SNP the normwise relative error of its coefficient
      vector, max_i |b_i - b_ref_i|_2 / |b_ref_i|_2; ``V_norm`` the same for
      the lower triangle of V_i (Frobenius norm).
    The scaled and normwise forms are the parity bars; all are reported."""
    B = np.asarray(B, dtype=np.float64).reshape(-1, p)
    Bref = np.asarray(Bref, dtype=np.float64).reshape(-1, p)
    V = np.asarray(V, dtype=np.float64).reshape(-1, p, p)        # [i][c][r]
    Vref = np.asarray(Vref, dtype=np.float64).reshape(-1, p, p)
    se = np.einsum("icc->ic", Vref)
    eb = np.abs(B - Bref) / np.maximum(np.abs(Bref), se)
    scale = np.maximum(np.abs(Vref), se[:, :, None] * se[:, None, :])
    ev = np.abs(V - Vref) / scale
    r, c = np.tril_indices(p)
    se_new = np.einsum("icc->ic", V)
    var_rel = np.abs(se_new ** 2 - se ** 2) / (se ** 2)
    bn = np.linalg.norm(B - Bref, axis=1) / np.linalg.norm(Bref, axis=1)
    VL, VLr = V[:, c, r], Vref[:, c, r]
    vn = np.linalg.norm(VL - VLr, axis=1) / np.linalg.norm(VLr, axis=1)
    return {"B_rel": max_rel_err(B, Bref), "V_rel": max_rel_err(V, Vref, p=p, lower_only=True),
            "var_rel": float(np.max(var_rel)) if var_rel.size else 0.0,
            "B_scaled": float(np.max(eb)) if eb.size else 0.0,
            "V_scaled": float(np.max(ev[:, c, r])) if ev.size else 0.0,
            "B_norm": float(np.max(bn)) if bn.size else 0.0,
            "V_norm": float(np.max(vn)) if vn.size else 0.0,
            "V_upper_zero": bool(np.all(V[:, r[r != c], c[r != c]] == 0.0)) if p > 1 else True}


def error_stats(B, V, Bref, Vref, p: int) -> dict:
    """Distribution of the element-wise relative errors (driver.c:159-177's
    |a - b| / |b|, entry by entry) of the coefficients b and of the variances
    var(b) = se^2 (the squared diagonal of V), both in the on-disk layout.

    Returns, for "b" and "var": the maximum, the median and the 99.9th
    percentile over all entries, and for "b" also the maximum over the
    entries with |b_ref| >= 1e-3 se (coefficients not negligible against
    their standard error; nearer zero the relative error of any two FP64
    computations grows without bound).  A zero reference entry counts as
    error 0 when the other entry is zero too, else as infinity."""
    B = np.asarray(B, dtype=np.float64).reshape(-1, p)
    Bref = np.asarray(Bref, dtype=np.float64).reshape(-1, p)
    V = np.asarray(V, dtype=np.float64).reshape(-1, p, p)
    Vref = np.asarray(Vref, dtype=np.float64).reshape(-1, p, p)
    se = np.einsum("icc->ic", Vref)
    var_ref, var_new = se ** 2, np.einsum("icc->ic", V) ** 2

    def rel(a, b):
        d = np.abs(a - b)
        with np.errstate(divide="ignore", invalid="ignore"):
            r = d / np.abs(b)
        r[(b == 0) & (d == 0)] = 0.0
        return r.ravel()

    def dist(r):
        if r.size == 0:
            return {"max": 0.0, "median": 0.0, "p999": 0.0}
        return {"max": float(r.max()), "median": float(np.median(r)), "p999": float(np.quantile(r, 0.999))}

    eb = rel(B, Bref)
    big = (np.abs(Bref) >= 1e-3 * se).ravel()
    out = {"b": dist(eb), "var": dist(rel(var_new, var_ref)), "entries": int(eb.size)}
    out["b"]["max_if_b_ge_1e-3_se"] = float(eb[big].max()) if big.any() else 0.0
    out["b"]["share_b_ge_1e-3_se"] = float(big.mean()) if big.size else 0.0
    return out
