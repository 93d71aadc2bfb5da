"""Known-answer fixture built on the reference's test_cases/test_XL.

The reference ships that directory's inputs (Phi.in, H.in, Sig.in, XL.in,
Y.in), the driver's stdin (input.in) and the expected coefficients
B_exp.out, copied unchanged into tests/golden/test_XL/.  Its XR.in and
V_exp.out are not shipped.

b_ij depends on the SNP column x_i only through c = X~_L' z, d_j = y~_j' z
and q = z'z, with z = L^-1 x (fgls_chol.c:394-428).  Since H and Sig are
constant across the 20 traits, M (and L) are shared, so for every SNP the
expected B rows fix c (the first wXL equations are rank one in c) and d_j up
to the one free value q.  This module rebuilds, for every SNP, a column x
that is consistent with all 20 expected B rows: the minimum-norm z meeting
[X~_L | Y~]' z = [c ; d(q)] plus a component orthogonal to that span, with q
at the vertex of the feasible interval, and x = L z.  Recomputing B from that
X_R reproduces B_exp.out, which makes test_XL an end-to-end known-answer test
for B.  V depends on the free q and has no reference values: parity of V is
pinned elsewhere (closed form), not here.

Run as a script to write XR.in next to the fixture inputs (8 MB, not
committed): ``python tests/golden/test_xl_fixture.py [outdir]``.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE_DIR = os.path.join(HERE, "test_XL")


def read_stdin_params(d: str = FIXTURE_DIR):
    """input.in holds the driver's 8 stdin integers (driver.c:35-49)."""
    vals = [int(v) for v in open(os.path.join(d, "input.in")).read().split()]
    keys = ["n", "p", "m", "t", "wXL", "x_b", "y_b", "threads"]
    return dict(zip(keys, vals))


def load_fixture(d: str = FIXTURE_DIR):
    prm = read_stdin_params(d)
    n, p, m, t, wXL = prm["n"], prm["p"], prm["m"], prm["t"], prm["wXL"]

    def rd(name, count):
        a = np.fromfile(os.path.join(d, name), dtype="<f8", count=count)
        assert a.size == count, name
        return a

    Phi = rd("Phi.in", n * n).reshape(n, n, order="F")
    H = rd("H.in", t)
    Sig = rd("Sig.in", t)
    XL = rd("XL.in", wXL * n).reshape(n, wXL, order="F")
    Y = rd("Y.in", t * n).reshape(n, t, order="F")
    Bexp = rd("B_exp.out", t * m * p).reshape(t, m, p)
    return prm, Phi, H, Sig, XL, Y, Bexp


def synth_xr(seed: int = 0):
    """Return (prm, Phi, H, Sig, XL, XR, Y, Bexp) with a synthetic XR (n x m)."""
    prm, Phi, H, Sig, XL, Y, Bexp = load_fixture()
    n, p, m, t, wXL = prm["n"], prm["p"], prm["m"], prm["t"], prm["wXL"]
    assert p - wXL == 1
    assert np.all(H == H[0]) and np.all(Sig == Sig[0]), "M must be shared by all traits"
    M = (H[0] * Sig[0]) * Phi
    M[np.diag_indices(n)] += (1.0 - H[0]) * Sig[0]
    L = sla.cholesky(M, lower=True)
    XLt = sla.solve_triangular(L, XL, lower=True)
    Yt = sla.solve_triangular(L, Y, lower=True)
    A = XLt.T @ XLt
    Bt = XLt.T @ Yt                                  # wXL x t
    G = np.hstack([XLt, Yt])                         # n x (wXL + t)
    Gi = np.linalg.inv(G.T @ G)
    Q, _ = np.linalg.qr(G)
    rng = np.random.default_rng(seed)
    XR = np.empty((n, m))
    for i in range(m):
        btop = Bexp[:, i, :wXL].T                    # wXL x t
        bs = Bexp[:, i, wXL]                         # t
        # A btop_j + c bs_j = Bt_j  ->  least squares for c over the traits
        R = Bt - A @ btop                            # wXL x t
        c = (R @ bs) / (bs @ bs)
        d0 = c @ btop                                # t  (d_j = c'btop_j + q bs_j)
        g0 = np.concatenate([c, d0])
        g1 = np.concatenate([np.zeros(wXL), bs])
        a2 = g1 @ Gi @ g1
        a1 = 2.0 * (g0 @ Gi @ g1)
        a0 = g0 @ Gi @ g0
        q = (1.0 - a1) / (2.0 * a2)                  # vertex of q - |z0(q)|^2
        g = g0 + q * g1
        z0 = G @ (Gi @ g)
        slack = q - z0 @ z0
        if slack < 0:
            raise ValueError(f"SNP {i}: no consistent column (slack {slack})")
        u = rng.standard_normal(n)
        u -= Q @ (Q.T @ u)
        u /= np.linalg.norm(u)
        z = z0 + np.sqrt(slack) * u
        XR[:, i] = L @ z
    return prm, Phi, H, Sig, XL, XR, Y, Bexp


def write_dir(out: str, seed: int = 0) -> dict:
    """Write a complete driver data directory (reference file names)."""
    prm, Phi, H, Sig, XL, XR, Y, Bexp = synth_xr(seed)
    os.makedirs(out, exist_ok=True)
    files = {"Phi.in": Phi.ravel(order="F"), "H.in": H, "Sig.in": Sig,
             "XL.in": XL.ravel(order="F"), "XR.in": XR.ravel(order="F"),
             "Y.in": Y.ravel(order="F"), "B_exp.out": Bexp.ravel()}
    for name, arr in files.items():
        np.asarray(arr, dtype="<f8").tofile(os.path.join(out, name))
    with open(os.path.join(out, "input.in"), "w") as f:
        f.write(open(os.path.join(FIXTURE_DIR, "input.in")).read())
    return prm


if __name__ == "__main__":
    write_dir(sys.argv[1] if len(sys.argv) > 1 else os.path.join(FIXTURE_DIR, "..", "test_XL_full"))
