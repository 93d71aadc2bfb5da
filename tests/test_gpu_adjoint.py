"""Adjoint of l = sum_i w_i b_i[p-1]^2 on the device, through the C-ABI.

No reference fixture pins it (the reference computes no gradients): the
device gradients are checked against the oracle's closed form, against
torch.autograd through a dense FP64 torch restatement, and against central
finite differences (step 1e-6) of the device forward loss.
"""
import numpy as np
import pytest
import torch

from oracle import fgls_oracle as orc
from tests.adjoint_ref import dense_grads
from tests.fgls_data import make_problem

pytestmark = pytest.mark.gpu

TOL = 1e-9      # device vs oracle / autograd, relative to the largest entry
TOL_FD = 1e-6   # central differences with step 1e-6


def _problem(n, m, wXL=3, seed=0, h2=0.5):
    pr = make_problem(n=n, wXL=wXL, m=m, seed=seed, h2=h2)
    w = np.random.default_rng(seed + 100).uniform(0.0, 1.0, m)
    return pr, w


def _dev_inputs(pr, w):
    XR = torch.from_numpy(np.ascontiguousarray(pr.XR.T)).cuda()            # (m, n)
    XL = torch.from_numpy(np.ascontiguousarray(pr.XL.T)).cuda()            # (wXL, n)
    y = torch.from_numpy(pr.y.copy()).cuda()
    wd = torch.from_numpy(w).cuda()
    return XR, XL, y, wd


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n,m,block,wXL", [(300, 2000, 2000, 3), (257, 700, 128, 3),
                                           (200, 333, 100, 1), (130, 260, 64, 6)])
def test_adjoint_matches_oracle(gpu, pkg, n, m, block, wXL):
    print(f"adjoint vs oracle n={n} m={m} block={block} wXL={wXL}", flush=True)
    from cugwas_proto_amd.adjoint import snp_loss_grad
    pr, w = _problem(n, m, wXL=wXL, seed=n + m)
    XR, XL, y, wd = _dev_inputs(pr, w)
    loss, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, block)
    loss_o, gy_o, gh_o = orc.snp_loss_adjoint(pr.Phi, pr.h2, 1.0, pr.XL, pr.XR, pr.y, w)
    assert abs(loss - loss_o) <= TOL * abs(loss_o)
    assert _rel(gy.cpu().numpy(), gy_o) <= TOL
    assert abs(gh - gh_o) <= TOL * abs(gh_o)


@pytest.mark.parametrize("n,m,block,wXL", [(300, 2000, 2000, 3), (257, 700, 128, 3), (1000, 600, 256, 5)])
def test_mixed_adjoint_matches_oracle(gpu, pkg, n, m, block, wXL):
    """The mixed-mode adjoint (int8 products against L^-1 and H = G L^-1)
    against the FP64 oracle, to the same bar as the FP64 adjoint, and against
    the device's FP64 adjoint."""
    print(f"mixed adjoint vs oracle n={n} m={m} block={block} wXL={wXL}", flush=True)
    from cugwas_proto_amd.adjoint import snp_loss_grad
    pr, w = _problem(n, m, wXL=wXL, seed=n + m + 1)
    XR, XL, y, wd = _dev_inputs(pr, w)
    loss, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, block, precision="mixed")
    loss_o, gy_o, gh_o = orc.snp_loss_adjoint(pr.Phi, pr.h2, 1.0, pr.XL, pr.XR, pr.y, w)
    assert abs(loss - loss_o) <= TOL * abs(loss_o)
    assert _rel(gy.cpu().numpy(), gy_o) <= TOL
    assert abs(gh - gh_o) <= TOL * abs(gh_o)
    l64, gy64, gh64 = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, block)
    assert _rel(gy.cpu().numpy(), gy64.cpu().numpy()) <= TOL and abs(gh - gh64) <= TOL * abs(gh64)


def test_mixed_adjoint_needs_mixed_begin(gpu, pkg):
    """adjoint_block_i8 after an FP64-mode begin is a state error."""
    import cugwas_proto_amd.fgls as F
    pr, w = _problem(128, 64, seed=4)
    XR, XL, y, wd = _dev_inputs(pr, w)
    L = orc.cholesky_lower(orc.build_M(pr.Phi, pr.h2, 1.0))
    dc = pkg.DeviceContext(pr.n, 3, 1, max_block=64)
    dc.set_trait(torch.from_numpy(np.ascontiguousarray(L.T)).cuda(), XL, y)
    dc.adjoint_begin(torch.from_numpy(np.ascontiguousarray(pr.Phi.T)).cuda(), 1.0)
    dc.set_mode(pkg.FGLS_MODE_MIXED_I8)
    X8 = dc.to_i8(XR)
    with pytest.raises(pkg.FglsError) as ei:
        dc.adjoint_block_i8(X8, wd)
    assert ei.value.code == pkg.FGLS_ERR_STATE and "mixed mode" in ei.value.msg
    dc.close()


def test_adjoint_matches_torch_autograd_dense(gpu, pkg):
    """Dense check at m = 2000 SNPs: torch.autograd through explicit M^-1."""
    print("adjoint vs dense autograd n=160 m=2000", flush=True)
    from cugwas_proto_amd.adjoint import snp_loss_grad
    pr, w = _problem(160, 2000, seed=3, h2=0.35)
    sigma = 1.7
    XR, XL, y, wd = _dev_inputs(pr, w)
    loss, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, sigma, 512)
    loss_r, gy_r, gh_r = dense_grads(pr.y, pr.h2, pr.Phi, pr.XL, pr.XR, w, sigma)
    assert abs(loss - loss_r) <= TOL * abs(loss_r)
    assert _rel(gy.cpu().numpy(), gy_r) <= TOL
    assert abs(gh - gh_r) <= TOL * abs(gh_r)


def test_autograd_function_and_central_differences(gpu, pkg):
    """SnpLoss through torch.autograd equals the adjoint; both match central
    differences (step 1e-6) of the device forward loss."""
    print("SnpLoss autograd + FD n=200 m=600", flush=True)
    from cugwas_proto_amd.adjoint import snp_loss, snp_loss_grad
    pr, w = _problem(200, 600, seed=9, h2=0.6)
    XR, XL, y, wd = _dev_inputs(pr, w)
    yv = y.clone().requires_grad_(True)
    hv = torch.tensor(pr.h2, dtype=torch.float64, requires_grad=True)
    loss = snp_loss(yv, hv, XR, wd, pr.Phi, XL, 1.0, 256)
    gy_a, gh_a = torch.autograd.grad(loss, (yv, hv))
    loss_g, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, 256)
    assert abs(float(loss.detach()) - loss_g) <= 1e-12 * abs(loss_g)
    assert _rel(gy_a.cpu().numpy(), gy.cpu().numpy()) <= 1e-12
    assert abs(float(gh_a) - gh) <= 1e-12 * abs(gh)

    eps = 1e-6

    def f(yy, hh):
        with torch.no_grad():
            return float(snp_loss(yy, torch.tensor(hh, dtype=torch.float64), XR, wd, pr.Phi, XL, 1.0, 256))

    fd_h = (f(y, pr.h2 + eps) - f(y, pr.h2 - eps)) / (2 * eps)
    assert abs(fd_h - gh) <= TOL_FD * abs(gh)
    d = torch.from_numpy(np.random.default_rng(1).standard_normal(pr.n)).cuda()
    fd_y = (f(y + eps * d, pr.h2) - f(y - eps * d, pr.h2)) / (2 * eps)
    gd = float(torch.dot(gy, d))
    assert abs(fd_y - gd) <= TOL_FD * abs(gd)


def test_autograd_function_mixed_and_central_differences(gpu, pkg):
    """SnpLoss with precision="mixed" (int8 forward and backward products):
    its autograd gradients equal snp_loss_grad's mixed ones and match central
    differences of the mixed forward loss (the int8 path is smooth in y and
    h2: only L^-1 and H change with h2, the genotypes stay exact)."""
    print("SnpLoss mixed autograd + FD n=200 m=600", flush=True)
    from cugwas_proto_amd.adjoint import snp_loss, snp_loss_grad
    pr, w = _problem(200, 600, seed=19, h2=0.4)
    XR, XL, y, wd = _dev_inputs(pr, w)
    yv = y.clone().requires_grad_(True)
    hv = torch.tensor(pr.h2, dtype=torch.float64, requires_grad=True)
    loss = snp_loss(yv, hv, XR, wd, pr.Phi, XL, 1.0, 256, "mixed")
    gy_a, gh_a = torch.autograd.grad(loss, (yv, hv))
    loss_g, gy, gh = snp_loss_grad(y, pr.h2, XR, wd, pr.Phi, XL, 1.0, 256, precision="mixed")
    assert abs(float(loss.detach()) - loss_g) <= 1e-12 * abs(loss_g)
    assert _rel(gy_a.cpu().numpy(), gy.cpu().numpy()) <= 1e-12
    assert abs(float(gh_a) - gh) <= 1e-12 * abs(gh)
    eps = 1e-6

    def f(yy, hh):
        with torch.no_grad():
            return float(snp_loss(yy, torch.tensor(hh, dtype=torch.float64), XR, wd, pr.Phi, XL, 1.0, 256,
                                  "mixed"))

    fd_h = (f(y, pr.h2 + eps) - f(y, pr.h2 - eps)) / (2 * eps)
    assert abs(fd_h - gh) <= TOL_FD * abs(gh)
    d = torch.from_numpy(np.random.default_rng(2).standard_normal(pr.n)).cuda()
    fd_y = (f(y + eps * d, pr.h2) - f(y - eps * d, pr.h2)) / (2 * eps)
    gd = float(torch.dot(gy, d))
    assert abs(fd_y - gd) <= TOL_FD * abs(gd)


def test_adjoint_blocking_and_repeat(gpu, pkg):
    """Block size does not change the result beyond rounding, and a second
    begin/blocks/end pass on the same context reproduces the first exactly."""
    print("adjoint blocking n=256 m=900", flush=True)
    from cugwas_proto_amd import DeviceContext, factor_M
    pr, w = _problem(256, 900, seed=21)
    XR, XL, y, wd = _dev_inputs(pr, w)
    L = factor_M(pr.Phi, pr.h2, 1.0)
    Ld = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()
    Phid = torch.from_numpy(np.ascontiguousarray(pr.Phi.T)).cuda()
    out = []
    for block in (900, 300, 97):
        with DeviceContext(pr.n, pr.wXL, 1, max_block=block) as dc:
            dc.set_trait(Ld, XL, y)
            for _ in range(2):
                dc.adjoint_begin(Phid, 1.0)
                for i0 in range(0, 900, block):
                    cnt = min(block, 900 - i0)
                    dc.adjoint_block(XR[i0:i0 + cnt], wd[i0:i0 + cnt], cnt, snp_base=i0)
                loss, gy, gh = dc.adjoint_end()
                out.append((loss, gy.cpu().numpy(), gh))
            assert out[-1][0] == out[-2][0] and out[-1][2] == out[-2][2]
            assert np.array_equal(out[-1][1], out[-2][1])
    for loss, gy, gh in out[2:]:
        assert abs(loss - out[0][0]) <= 1e-12 * abs(out[0][0])
        assert _rel(gy, out[0][1]) <= 1e-11
        assert abs(gh - out[0][2]) <= 1e-11 * abs(out[0][2])


def test_adjoint_errors(gpu, pkg):
    from cugwas_proto_amd import DeviceContext, FglsError
    pr, w = _problem(64, 10, seed=2)
    XR, XL, y, wd = _dev_inputs(pr, w)
    with DeviceContext(pr.n, pr.wXL, 1, max_block=16) as dc:
        with pytest.raises(FglsError):                      # no trait yet
            dc.adjoint_begin(torch.zeros((64, 64), dtype=torch.float64, device="cuda"), 1.0)
        L = torch.from_numpy(np.ascontiguousarray(np.linalg.cholesky(
            pr.h2 * pr.Phi + (1 - pr.h2) * np.eye(64)).T)).cuda()
        dc.set_trait(L, XL, y)
        with pytest.raises(FglsError):                      # block before begin
            dc.adjoint_block(XR, wd)
        with pytest.raises(FglsError):                      # sigma2 <= 0
            dc.adjoint_begin(torch.from_numpy(pr.Phi.copy()).cuda(), 0.0)
