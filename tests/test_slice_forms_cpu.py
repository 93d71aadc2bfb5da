"""The arithmetic of the two digit forms of L^-1 used by the mixed path,
restated in numpy on the CPU (the device kernel k_slice_tri follows the same
steps for both forms; its outputs are checked on the GPU against the oracle
in test_gpu_slice_forms.py).

Six digits: row i scaled by sc_i = 127 / max_i, balanced radix-254 digits
a_s = rint(r_s), r_{s+1} = 254 (r_s - a_s).  Five digits with a top digit:
sc_i = min(127 / off_i, 127 * 254 / max_i) with off_i the largest |L^-1_ik|
outside the row's diagonal 64-column block; inside that block a top digit
t = rint(r / 254) first.  Pinned here: every digit is an int8 in
[-127, 127]; the top digit is zero off the diagonal block; the
reconstruction error per entry is at most 127 * 254^-S / sc_i (the bound the
error analysis uses), i.e. loss_i = 254 off_i / max_i times the six-digit
bound; the integer products of the digits with genotypes in 0..3 reproduce
L^-1 x to that bound.
"""
import numpy as np
import scipy.linalg as sla

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem

R = 254.0


def digits(r, S):
    out = []
    for _ in range(S):
        a = np.clip(np.rint(r), -127, 127)
        out.append(a)
        r = (r - a) * R
    return out


def horner(ds):
    t = ds[-1].copy()
    for a in ds[-2::-1]:
        t = t / R + a
    return t


def six_digit_form(Li):
    sc = 127.0 / np.abs(Li).max(axis=1)
    ds = digits(Li * sc[:, None], 6)
    return ds, sc


def five_digit_form(Li):
    n = Li.shape[0]
    k = np.arange(n)[None, :]
    d0 = (np.arange(n) // 64 * 64)[:, None]
    diag = (k >= d0) & (k < d0 + 64)
    mx = np.abs(Li).max(axis=1)
    off = np.where(diag, 0.0, np.abs(Li)).max(axis=1)
    with np.errstate(divide="ignore"):
        sc = np.minimum(np.where(off > 0, 127.0 / off, np.inf), 127.0 * R / mx)
    r = Li * sc[:, None]
    top = np.where(diag, np.clip(np.rint(r / R), -127, 127), 0.0)
    ds = digits(r - R * top, 5)
    loss = R * off / mx
    return top, ds, sc, loss, diag


def test_digit_forms_bounds_and_products():
    prob = make_problem(n=700, wXL=3, m=64, seed=17, h2=0.5)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    Li = sla.solve_triangular(L, np.eye(prob.n), lower=True)

    ds6, sc6 = six_digit_form(Li)
    for a in ds6:
        assert np.all(np.abs(a) <= 127) and np.array_equal(a, np.rint(a))
    # FP64 rounding of the scaling and of the recombination: a few ulp of the row maximum
    fp = 8 * np.finfo(float).eps * np.abs(Li).max(axis=1)[:, None]
    err6 = np.abs(horner(ds6) / sc6[:, None] - Li)
    bound6 = 127.0 * R ** -6 / sc6
    assert np.all(err6 <= bound6[:, None] + fp)

    top, ds5, sc5, loss, diag = five_digit_form(Li)
    assert np.all(np.abs(top) <= 127) and np.all(top[~diag] == 0)
    for a in ds5:
        assert np.all(np.abs(a) <= 127) and np.array_equal(a, np.rint(a))
    rec5 = (R * top + horner(ds5)) / sc5[:, None]
    err5 = np.abs(rec5 - Li)
    bound5 = 127.0 * R ** -5 / sc5
    assert np.all(err5 <= bound5[:, None] + fp)
    # the five-digit bound is max(1, loss_i) times the six-digit one
    ratio = bound5 / (127.0 * R ** -6 / sc6)
    assert np.allclose(ratio, np.maximum(loss, 1.0), rtol=1e-12)

    # the integer products (what the int8 MFMA sums exactly) reproduce L^-1 x
    X = prob.XR[:, :32]
    z = Li @ X
    acc6 = [a @ X for a in ds6]
    z6 = horner(acc6) / sc6[:, None]
    acc5 = [a @ X for a in ds5]
    acc5[0] = acc5[0] + R * (top @ X)             # the top digit folded into digit 0
    z5 = horner(acc5) / sc5[:, None]
    colsum = np.abs(X).sum(axis=0)[None, :]
    zf = 8 * np.finfo(float).eps * np.abs(Li).max(axis=1)[:, None] * colsum + 1e-13 * np.abs(z)
    assert np.all(np.abs(z6 - z) <= bound6[:, None] * colsum + zf)
    assert np.all(np.abs(z5 - z) <= bound5[:, None] * colsum + zf)


def test_planted_slip_fails_the_bound():
    """Dropping the top digit of the diagonal block (a plausible mistake)
    violates the five-digit reconstruction bound."""
    prob = make_problem(n=300, wXL=3, m=8, seed=3, h2=0.6)
    L = orc.cholesky_lower(orc.build_M(prob.Phi, prob.h2, 1.0))
    Li = sla.solve_triangular(L, np.eye(prob.n), lower=True)
    top, ds5, sc5, loss, diag = five_digit_form(Li)
    rec_wrong = horner(ds5) / sc5[:, None]          # top digit dropped
    bound5 = 127.0 * R ** -5 / sc5
    fp = 8 * np.finfo(float).eps * np.abs(Li).max(axis=1)[:, None]
    assert np.any(np.abs(rec_wrong - Li) > bound5[:, None] + fp)
