"""fgls_hip_solve_block_noalloc (the solve for callers that cannot allocate)
against fgls_hip_solve_block, through the C-ABI.

The no-alloc entry point runs the same kernels with its scratch (partial
sums for wXR = 1, transformed columns for wXR > 1) in the library's static
device array: its b and V must equal fgls_hip_solve_block's bit for bit, also
on the second of two calls in a row with a smaller input (stale scratch from
the first call must not leak into the second), and match the FP64 oracle to
the FP64 path's bar.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import fgls_oracle as orc
from tests.fgls_data import make_problem
from tests.test_gpu_parity import check_bv, setup

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20     # FGLS_NOALLOC_SCRATCH_BYTES (include/fgls_hip.h)


def noalloc(pkg, ctx, X, count, B, V, base=0):
    """The C entry point itself (no Python-side allocation either)."""
    return pkg.lib().fgls_hip_solve_block_noalloc(ctx._h, X.data_ptr(), X.shape[1], count, B.data_ptr(),
                                                  V.data_ptr(), base,
                                                  torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("n,wXL,wXR", [(1000, 3, 1), (1031, 5, 1), (600, 2, 3)])
def test_noalloc_equals_solve_block(pkg, gpu, n, wXL, wXR):
    print(f"no-alloc solve n={n} wXL={wXL} wXR={wXR}", flush=True)
    m1, m2 = 700, 333
    p = wXL + wXR
    prob = make_problem(n=n, wXL=wXL, m=m1, seed=n + wXR, wXR=wXR)
    tr, ctx = setup(pkg, prob, wXR=wXR, max_block=1024)
    X = pkg.colmajor(prob.XR, ld=(n + 3) // 4 * 4)
    X2 = X[:m2 * wXR].contiguous()
    R1 = ctx.solve_block(X, count=m1, snp_base=11)
    R2 = ctx.solve_block(X2, count=m2, snp_base=3)
    torch.cuda.synchronize()
    out = []
    for count in (m1, m2):
        out.append((torch.full((count, p), float("nan"), dtype=torch.float64, device="cuda"),
                    torch.full((count, p * p), float("nan"), dtype=torch.float64, device="cuda")))
    # two calls in a row on one stream, the second smaller
    assert noalloc(pkg, ctx, X, m1, *out[0], 11) == pkg.FGLS_OK, pkg.last_error()
    assert noalloc(pkg, ctx, X2, m2, *out[1], 3) == pkg.FGLS_OK, pkg.last_error()
    torch.cuda.synchronize()
    ctx.check()
    for (B, V), (Br, Vr) in zip(out, (R1, R2)):
        assert torch.equal(B, Br) and torch.equal(V, Vr)
    b_ref, V_ref = orc.solve_snps(prob.XR, tr, wXR)
    check_bv(out[0][0], out[0][1], b_ref, V_ref, p)
    check_bv(out[1][0], out[1][1], b_ref[:m2], V_ref[:m2], p)
    ctx.close()


def test_noalloc_keeps_pending_reduce_and_records_failures(pkg, gpu):
    """A no-alloc call between block_reduce and block_finalize leaves the
    context's partial sums alone; a singular SNP is recorded in the context
    (fgls_hip_check) as by fgls_hip_solve_block."""
    prob = make_problem(n=500, wXL=3, m=200, seed=5)
    prob.XR[:, 150] = 0.0
    tr, ctx = setup(pkg, prob, max_block=256)
    X = pkg.colmajor(prob.XR)
    Bs, Vs = ctx.solve_block(X[:100].contiguous(), count=100)
    ctx.block_reduce(X[:100].contiguous(), 100)
    B = torch.empty((200, 4), dtype=torch.float64, device="cuda")
    V = torch.empty((200, 16), dtype=torch.float64, device="cuda")
    assert noalloc(pkg, ctx, X, 200, B, V, 1000) == pkg.FGLS_OK
    Bf = torch.empty_like(Bs)
    Vf = torch.empty_like(Vs)
    ctx.block_finalize(100, Bf, Vf)
    torch.cuda.synchronize()
    assert torch.equal(Bf, Bs) and torch.equal(Vf, Vs)
    with pytest.raises(pkg.FglsError) as ei:
        ctx.check()
    assert ei.value.bad_snp == 1150 and ei.value.info == 4
    ctx.close()


def test_noalloc_rejects_counts_beyond_the_scratch(pkg, gpu):
    n, wXL = 1000, 3
    prob = make_problem(n=n, wXL=wXL, m=4, seed=2)
    _, ctx = setup(pkg, prob, max_block=64)
    cap = SCRATCH_BYTES // (((n + 127) // 128) * (wXL + 2) * 8)
    X = pkg.colmajor(prob.XR)
    B = torch.empty((4, 4), dtype=torch.float64, device="cuda")
    V = torch.empty((4, 16), dtype=torch.float64, device="cuda")
    rc = noalloc(pkg, ctx, X, cap + 1, B, V)
    assert rc == pkg.FGLS_ERR_ARG and f"[0, {cap}]" in pkg.last_error(), pkg.last_error()
    assert noalloc(pkg, ctx, X, 4, B, V) == pkg.FGLS_OK
    torch.cuda.synchronize()
    ctx.close()
