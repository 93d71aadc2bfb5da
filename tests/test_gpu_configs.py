"""GPU parity at the full shapes of the benchmark configurations.

The other GPU tests run at n <= 4,096; n-dependent code paths are only
reached at the configurations' sizes: the int8 path's row scaling and int32
exactness bound, the triangular tile bounds of the products, the split-K
products of the per-trait factor beyond n = 8,400, the row-tile counts of
the finalize.  Each class draws one configuration with its own generators on
the device (tests/fgls_data.make_problem_device) and checks the HIP path
through the C-ABI against the FP64 oracle (oracle/fgls_oracle.py: LAPACK
dpotrf and dtrsm on the host) on a sample of SNPs that includes every ragged
block and the SNPs where the mixed and FP64 paths differ most; the paths are
compared with each other over all SNPs.

  C1  n = 10,000, p = 4, h2 = 0.5, hard-call genotypes: device blocks of
      65,536 SNPs + a ragged block, FP64 and mixed; the file pipeline
      (FP64 and 2-bit X_R files, both precisions).
  C2  n = 20,000, p = 8, h2 = 0.3, dosages round(U(0, 2), 3): FP64 and the
      int8 digit-plane path; the file pipeline (16-bit and FP64 dosage files).
  C3  adjoint at n = 5,000, p = 4, h2 = 0.6, blocks of 2,048: against the
      oracle's closed form, central differences (step 1e-6) and torch.autograd
      through a dense FP64 reference at m = 2,000.
  C4  n = 30,000, p = 6, t = 4 traits with h2 ~ U(0.1, 0.9), a 2-bit file,
      all traits in one pass, both precisions, each trait against the oracle;
      the adjoint of one trait against central differences.

Bars: FP64 1e-10 (b relative to max(|b|, se), V scaled), mixed 2e-11 on the
same and 1e-12 on var(b) element-wise (tests/test_gpu_mixed.py states why);
the V strict upper triangle exactly 0; the adjoint 1e-9 against the oracle
and 1e-6 against central differences.
"""
import os

import numpy as np
import pytest
import torch

from oracle import fgls_oracle as orc
from tests.adjoint_ref import dense_grads
from tests.fgls_data import make_problem_device, write_device_problem_dir
from tests.test_gpu_mixed import TOL_SCALED, TOL_VAR

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL_ADJ = 1e-9
TOL_FD = 1e-6


def say(msg):
    print(f"[configs] {msg}", flush=True)


def oracle_for(prob, k, idx, sigma=1.0):
    """Oracle b (len(idx) x p) and V (on-disk order) of trait k on SNPs idx."""
    Phi = prob.host_phi()
    L = orc.cholesky_lower(orc.build_M(Phi, prob.H[k], sigma))
    del Phi
    tr = orc.trait_precompute(L, prob.host_xl(), prob.Y[k].cpu().numpy())
    b, V = orc.solve_snps(prob.host_x(idx), tr, 1)
    return b, V.transpose(0, 2, 1).reshape(len(idx), -1)


def sample(m, k, seed, must=()):
    """k random SNPs of [0, m) plus the given ones, sorted and unique."""
    r = np.random.default_rng(seed).integers(0, m, k)
    return np.unique(np.concatenate([r, np.asarray(must, dtype=np.int64)])).astype(np.int64)


def top_gaps(B, B_other, k):
    """The k SNPs with the largest element-wise relative gap of b."""
    a, b = B.cpu().numpy(), B_other.cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.nan_to_num(np.abs(a - b) / np.abs(b), nan=0.0, posinf=1e300).max(axis=1)
    return np.argsort(g)[-k:]


def check(B, V, b_ref, V_ref, p, mixed, what):
    e = orc.parity_errors(B, V, b_ref, V_ref, p)
    s = orc.error_stats(B, V, b_ref, V_ref, p)
    say(f"{what}: {e}; element-wise b max {s['b']['max']:.2e} median {s['b']['median']:.2e} "
        f"p99.9 {s['b']['p999']:.2e} (|b| >= 1e-3 se: {s['b']['max_if_b_ge_1e-3_se']:.2e}); "
        f"var max {s['var']['max']:.2e}")
    assert e["V_upper_zero"], what
    if mixed:
        assert e["B_scaled"] < TOL_SCALED and e["V_scaled"] < TOL_SCALED, (what, e)
        assert e["var_rel"] < TOL_VAR and e["B_norm"] < TOL_VAR, (what, e)
    else:
        assert e["B_scaled"] < TOL64 and e["V_scaled"] < TOL64, (what, e)
    return e


def read_out(d, name, count):
    return np.fromfile(os.path.join(d, name), dtype="<f8", count=count)


def release(prob):
    for a in ("Phi", "XL", "Y", "X", "X8", "V16"):
        setattr(prob, a, None)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
class TestC1:
    """n = 10,000, p = 4 (intercept + 2 covariates + SNP), h2 = 0.5."""
    N, WXL, H2, SEED = 10000, 3, 0.5, 1
    BLOCK = 65536
    M = 65536 + 1237
    MF, FBLOCK = 20000, 8192        # file pipeline: 2 full blocks + 3,616

    @pytest.fixture(scope="class")
    def prob(self, gpu):
        say(f"C1 data n={self.N} m={self.M}")
        pr = make_problem_device(self.N, self.WXL, self.M, self.SEED, [self.H2])
        yield pr
        release(pr)

    @pytest.fixture(scope="class")
    def device_out(self, pkg, prob):
        """FP64 and mixed b, V of all M SNPs through the device context."""
        n, p, B0 = self.N, self.WXL + 1, self.BLOCK
        out = {}
        with pkg.DeviceContext(n, self.WXL, 1, max_block=B0) as ctx:
            ctx.set_trait_phi(prob.Phi, self.H2, 1.0, prob.XL, prob.Y[0].contiguous())
            B = torch.empty((self.M, p), dtype=torch.float64, device="cuda")
            V = torch.empty((self.M, p * p), dtype=torch.float64, device="cuda")
            for i0 in (0, B0):
                cnt = min(B0, self.M - i0)
                ctx.block_reduce(prob.X[i0:i0 + cnt], cnt)
                ctx.block_finalize(cnt, B[i0:i0 + cnt], V[i0:i0 + cnt], snp_base=i0)
            ctx.check()
            out["fp64"] = (B, V)
            # the fused entry point is the same arithmetic
            Bf, Vf = ctx.solve_block(prob.X[B0:], self.M - B0)
            assert torch.equal(Bf, B[B0:]) and torch.equal(Vf, V[B0:])
            ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
            B8 = torch.empty_like(B)
            V8 = torch.empty_like(V)
            XT = pkg.tile_i8(prob.X8[:B0], n, B0)              # tiled, as the benchmark
            ctx.block_reduce_i8t(XT, B0)
            ctx.block_finalize(B0, B8[:B0], V8[:B0], snp_base=0)
            del XT
            ctx.solve_block_i8(prob.X8[B0:], self.M - B0, B8[B0:], V8[B0:], snp_base=B0)
            ctx.check()
            out["mixed"] = (B8, V8)
        torch.cuda.synchronize()
        return out

    def test_c1_device_paths_vs_oracle(self, prob, device_out):
        p = self.WXL + 1
        B, V = device_out["fp64"]
        B8, V8 = device_out["mixed"]
        e = orc.parity_errors(B8.cpu().numpy(), V8.cpu().numpy(), B.cpu().numpy(), V.cpu().numpy(), p)
        say(f"C1 mixed vs fp64 over all {self.M} SNPs: {e}")
        assert e["B_scaled"] < TOL_SCALED and e["var_rel"] < TOL_VAR and e["V_upper_zero"], e
        idx = sample(self.BLOCK, 384, 11, must=list(range(self.BLOCK, self.M, 37)) + [self.M - 1] +
                     list(top_gaps(B8, B, 64)))
        say(f"C1 oracle on {idx.size} SNPs (host dpotrf at n={self.N})")
        b_ref, V_ref = oracle_for(prob, 0, idx)
        sel = torch.as_tensor(idx, device="cuda")
        check(B[sel].cpu().numpy(), V[sel].cpu().numpy(), b_ref, V_ref, p, False, "C1 fp64 vs oracle")
        check(B8[sel].cpu().numpy(), V8[sel].cpu().numpy(), b_ref, V_ref, p, True, "C1 mixed vs oracle")

    def test_c1_file_pipeline(self, pkg, prob, device_out, tmp_path):
        """fgls_chol_gpu on files at n = 10,000: the FP64 XR.in and a 2-bit
        file of the same genotypes, both precisions, give bitwise the device
        context's results (the same kernels on other block boundaries)."""
        n, p, m = self.N, self.WXL + 1, self.MF
        d = str(tmp_path)
        say(f"C1 files: m={m} (FP64 XR.in {8 * n * m / 1e9:.1f} GB and a 2-bit file)")
        write_device_problem_dir(prob, d, xr_formats=("f64", "u2"), m=m)
        f = lambda s: os.path.join(d, s)
        for prec in ("fp64", "mixed"):
            res = {}
            for fmt, xr in (("f64", "XR.in"), ("u2", "XR2.in")):
                pkg.fgls_chol_gpu(n, p, m, 1, self.WXL, 1, 4096, 1, 16, f("Phi.in"), f("H.in"), f("Sig.in"),
                                  f("XL.in"), f(xr), f("Y.in"), f("B.out"), f("V.out"),
                                  options={"precision": prec, "xr_format": fmt, "block": self.FBLOCK})
                res[fmt] = (read_out(d, "B.out", m * p), read_out(d, "V.out", m * p * p))
            assert np.array_equal(res["f64"][0], res["u2"][0]) and np.array_equal(res["f64"][1], res["u2"][1])
            Bd, Vd = device_out[prec]
            assert np.array_equal(res["f64"][0], Bd[:m].cpu().numpy().ravel()), prec
            assert np.array_equal(res["f64"][1], Vd[:m].cpu().numpy().ravel()), prec
            say(f"C1 file pipeline {prec}: FP64 and 2-bit files bitwise equal to the device context")


# ---------------------------------------------------------------------------
class TestC2:
    """n = 20,000, p = 8 (intercept + 6 covariates + SNP), h2 = 0.3, dosages
    round(U(0, 2), 3)."""
    N, WXL, H2, SEED = 20000, 7, 0.3, 2
    BLOCK = 16384
    M = 16384 + 1001
    MF, FBLOCK = 6000, 4096

    @pytest.fixture(scope="class")
    def prob(self, gpu):
        say(f"C2 data n={self.N} m={self.M} (dosages)")
        pr = make_problem_device(self.N, self.WXL, self.M, self.SEED, [self.H2], dosage="imputed")
        yield pr
        release(pr)

    @pytest.fixture(scope="class")
    def device_out(self, pkg, prob):
        n, p, B0 = self.N, self.WXL + 1, self.BLOCK
        out = {}
        with pkg.DeviceContext(n, self.WXL, 1, max_block=B0) as ctx:
            ctx.set_trait_phi(prob.Phi, self.H2, 1.0, prob.XL, prob.Y[0].contiguous())
            B = torch.empty((self.M, p), dtype=torch.float64, device="cuda")
            V = torch.empty((self.M, p * p), dtype=torch.float64, device="cuda")
            for i0 in (0, B0):
                cnt = min(B0, self.M - i0)
                ctx.solve_block(prob.X[i0:i0 + cnt], cnt, B[i0:i0 + cnt], V[i0:i0 + cnt], snp_base=i0)
            ctx.check()
            out["fp64"] = (B, V)
            ctx.set_mode(pkg.FGLS_MODE_MIXED_I8)
            B8 = torch.empty_like(B)
            V8 = torch.empty_like(V)
            for i0 in (0, B0):
                cnt = min(B0, self.M - i0)
                D8 = ctx.to_dosage_i8(prob.X[i0:i0 + cnt], 1000.0)
                ctx.solve_block_dosage(D8, 1000.0, cnt, B8[i0:i0 + cnt], V8[i0:i0 + cnt], snp_base=i0)
                del D8
            ctx.check()
            out["mixed"] = (B8, V8)
        torch.cuda.synchronize()
        return out

    def test_c2_device_paths_vs_oracle(self, prob, device_out):
        p = self.WXL + 1
        B, V = device_out["fp64"]
        B8, V8 = device_out["mixed"]
        e = orc.parity_errors(B8.cpu().numpy(), V8.cpu().numpy(), B.cpu().numpy(), V.cpu().numpy(), p)
        say(f"C2 dosages mixed vs fp64 over all {self.M} SNPs: {e}")
        assert e["B_scaled"] < TOL_SCALED and e["var_rel"] < TOL_VAR and e["V_upper_zero"], e
        idx = sample(self.BLOCK, 192, 12, must=list(range(self.BLOCK, self.M, 31)) + [self.M - 1] +
                     list(top_gaps(B8, B, 32)))
        say(f"C2 oracle on {idx.size} SNPs (host dpotrf at n={self.N})")
        b_ref, V_ref = oracle_for(prob, 0, idx)
        sel = torch.as_tensor(idx, device="cuda")
        check(B[sel].cpu().numpy(), V[sel].cpu().numpy(), b_ref, V_ref, p, False, "C2 fp64 vs oracle")
        check(B8[sel].cpu().numpy(), V8[sel].cpu().numpy(), b_ref, V_ref, p, True, "C2 dosages vs oracle")

    def test_c2_file_pipeline(self, pkg, prob, device_out, tmp_path):
        """The 16-bit dosage file (xr_format=i16, dosage_scale=1000) and the
        reference's FP64 dosage file give bitwise the device context's
        results, FP64 and mixed."""
        n, p, m = self.N, self.WXL + 1, self.MF
        d = str(tmp_path)
        say(f"C2 files: m={m} (FP64 and 16-bit dosage files)")
        write_device_problem_dir(prob, d, xr_formats=("f64", "i16"), m=m)
        f = lambda s: os.path.join(d, s)
        runs = {("fp64", "i16"): {"xr_format": "i16", "dosage_scale": 1000},
                ("fp64", "f64"): {},
                ("mixed", "i16"): {"precision": "mixed", "xr_format": "i16", "dosage_scale": 1000},
                ("mixed", "f64"): {"precision": "mixed", "dosage_scale": 1000}}
        for (prec, fmt), opts in runs.items():
            pkg.fgls_chol_gpu(n, p, m, 1, self.WXL, 1, 4096, 1, 16, f("Phi.in"), f("H.in"), f("Sig.in"),
                              f("XL.in"), f("XR16.in" if fmt == "i16" else "XR.in"), f("Y.in"), f("B.out"),
                              f("V.out"), options=dict(opts, block=self.FBLOCK))
            Bd, Vd = device_out[prec]
            assert np.array_equal(read_out(d, "B.out", m * p), Bd[:m].cpu().numpy().ravel()), (prec, fmt)
            assert np.array_equal(read_out(d, "V.out", m * p * p), Vd[:m].cpu().numpy().ravel()), (prec, fmt)
            say(f"C2 file pipeline {prec} from the {fmt} file: bitwise equal to the device context")


# ---------------------------------------------------------------------------
class TestC3:
    """Adjoint at n = 5,000, p = 4, h2 = 0.6, blocks of 2,048 SNPs."""
    N, WXL, H2, SEED = 5000, 3, 0.6, 3
    M, BLOCK = 4096, 2048

    @pytest.fixture(scope="class")
    def prob(self, gpu):
        say(f"C3 data n={self.N} m={self.M}")
        pr = make_problem_device(self.N, self.WXL, self.M, self.SEED, [self.H2])
        pr.w = torch.from_numpy(np.random.default_rng(self.SEED + 100).uniform(0.0, 1.0, self.M)).cuda()
        yield pr
        release(pr)

    def _rel(self, a, b):
        a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    @pytest.mark.parametrize("precision", ["fp64", "mixed"])
    def test_c3_adjoint_vs_oracle(self, prob, precision):
        from cugwas_proto_amd.adjoint import snp_loss_grad
        y = prob.Y[0].contiguous()
        loss, gy, gh = snp_loss_grad(y, self.H2, prob.X, prob.w, prob.Phi, prob.XL, 1.0, self.BLOCK,
                                     precision=precision)
        say(f"C3 {precision}: loss {loss:.6e} dl/dh2 {gh:.6e}; oracle (host) ...")
        Xh = prob.host_x(np.arange(self.M))
        lo, gyo, gho = orc.snp_loss_adjoint(prob.host_phi(), self.H2, 1.0, prob.host_xl(), Xh,
                                            y.cpu().numpy(), prob.w.cpu().numpy())
        ry, rh = self._rel(gy.cpu().numpy(), gyo), abs(gh - gho) / abs(gho)
        say(f"C3 {precision} vs oracle: loss {abs(loss - lo) / abs(lo):.2e} dl/dy {ry:.2e} dl/dh2 {rh:.2e}")
        assert abs(loss - lo) <= TOL_ADJ * abs(lo) and ry <= TOL_ADJ and rh <= TOL_ADJ

    @pytest.mark.parametrize("precision", ["fp64", "mixed"])
    def test_c3_snploss_autograd_and_central_differences(self, pkg, prob, precision):
        """SnpLoss (device factor of M in the forward, adjoint in the
        backward) equals snp_loss_grad; both match central differences with
        step 1e-6 in h2 and along a random direction of y."""
        from cugwas_proto_amd import adjoint
        from cugwas_proto_amd.adjoint import snp_loss, snp_loss_grad
        assert not hasattr(adjoint, "factor_M")          # no host LAPACK on this path
        y = prob.Y[0].contiguous()
        yv = y.clone().requires_grad_(True)
        hv = torch.tensor(self.H2, dtype=torch.float64, requires_grad=True)
        loss = snp_loss(yv, hv, prob.X, prob.w, prob.Phi, prob.XL, 1.0, self.BLOCK, precision)
        gy_a, gh_a = torch.autograd.grad(loss, (yv, hv))
        lg, gy, gh = snp_loss_grad(y, self.H2, prob.X, prob.w, prob.Phi, prob.XL, 1.0, self.BLOCK,
                                   precision=precision)
        assert abs(float(loss.detach()) - lg) <= 1e-12 * abs(lg)
        assert self._rel(gy_a.cpu().numpy(), gy.cpu().numpy()) <= 1e-12
        assert abs(float(gh_a) - gh) <= 1e-12 * abs(gh)
        eps = 1e-6

        def f(yy, hh):
            with torch.no_grad():
                return float(snp_loss(yy, torch.tensor(hh, dtype=torch.float64), prob.X, prob.w, prob.Phi,
                                      prob.XL, 1.0, self.BLOCK, precision))

        fd_h = (f(y, self.H2 + eps) - f(y, self.H2 - eps)) / (2 * eps)
        dvec = torch.from_numpy(np.random.default_rng(5).standard_normal(self.N)).cuda()
        fd_y = (f(y + eps * dvec, self.H2) - f(y - eps * dvec, self.H2)) / (2 * eps)
        gd = float(torch.dot(gy, dvec))
        say(f"C3 {precision} central differences: dl/dh2 {abs(fd_h - gh) / abs(gh):.2e}, "
            f"dl/dy.d {abs(fd_y - gd) / abs(gd):.2e}")
        assert abs(fd_h - gh) <= TOL_FD * abs(gh)
        assert abs(fd_y - gd) <= TOL_FD * abs(gd)

    def test_c3_dense_autograd_m2000(self, prob):
        """torch.autograd through the dense FP64 reference (explicit M^-1) on
        the first 2,000 SNPs."""
        from cugwas_proto_amd.adjoint import snp_loss_grad
        m = 2000
        y = prob.Y[0].contiguous()
        loss, gy, gh = snp_loss_grad(y, self.H2, prob.X[:m], prob.w[:m], prob.Phi, prob.XL, 1.0, self.BLOCK)
        lr, gyr, ghr = dense_grads(y.cpu().numpy(), self.H2, prob.host_phi(), prob.host_xl(),
                                   prob.host_x(np.arange(m)), prob.w[:m].cpu().numpy())
        ry = self._rel(gy.cpu().numpy(), gyr)
        say(f"C3 dense autograd m={m}: loss {abs(loss - lr) / abs(lr):.2e} dl/dy {ry:.2e} "
            f"dl/dh2 {abs(gh - ghr) / abs(ghr):.2e}")
        assert abs(loss - lr) <= TOL_ADJ * abs(lr) and ry <= TOL_ADJ and abs(gh - ghr) <= TOL_ADJ * abs(ghr)


# ---------------------------------------------------------------------------
class TestC4:
    """n = 30,000, p = 6 (intercept + 4 covariates + SNP), t = 4 traits with
    h2 ~ U(0.1, 0.9), each with its own factor L_k, hard-call genotypes in a
    2-bit file, all four traits in one pass over X_R."""
    N, WXL, SEED, T = 30000, 5, 4, 4
    M, FBLOCK = 12000, 4096           # 2 full blocks + 3,808
    H = [float(h) for h in np.random.default_rng(4).uniform(0.1, 0.9, 4)]

    @pytest.fixture(scope="class")
    def prob(self, gpu):
        say(f"C4 data n={self.N} m={self.M} t={self.T} h2={self.H}")
        pr = make_problem_device(self.N, self.WXL, self.M, self.SEED, self.H)
        yield pr
        release(pr)

    @pytest.fixture(scope="class")
    def file_out(self, pkg, prob, tmp_path_factory):
        n, p, m, t = self.N, self.WXL + 1, self.M, self.T
        d = str(tmp_path_factory.mktemp("c4"))
        say(f"C4 files: Phi.in {8 * n * n / 1e9:.1f} GB, 2-bit X_R {m * ((n + 3) // 4) / 1e6:.0f} MB")
        write_device_problem_dir(prob, d, xr_formats=("u2",))
        f = lambda s: os.path.join(d, s)
        out = {}
        for prec in ("fp64", "mixed"):
            say(f"C4 fgls_chol_gpu {prec}, 4 traits per pass")
            pkg.fgls_chol_gpu(n, p, m, t, self.WXL, 1, 4096, 1, 16, f("Phi.in"), f("H.in"), f("Sig.in"),
                              f("XL.in"), f("XR2.in"), f("Y.in"), f("B.out"), f("V.out"),
                              options={"precision": prec, "xr_format": "u2", "block": self.FBLOCK,
                                       "traits_per_pass": t})
            out[prec] = (read_out(d, "B.out", t * m * p).reshape(t, m, p),
                         read_out(d, "V.out", t * m * p * p).reshape(t, m, p * p))
        os.remove(f("Phi.in"))
        return out

    def test_c4_mixed_vs_fp64_all(self, file_out):
        p = self.WXL + 1
        for k in range(self.T):
            e = orc.parity_errors(file_out["mixed"][0][k], file_out["mixed"][1][k], file_out["fp64"][0][k],
                                  file_out["fp64"][1][k], p)
            say(f"C4 trait {k} (h2 {self.H[k]:.3f}) mixed vs fp64 over all {self.M} SNPs: {e}")
            assert e["B_scaled"] < TOL_SCALED and e["var_rel"] < TOL_VAR and e["V_upper_zero"], (k, e)

    @pytest.mark.parametrize("k", [0, 1, 2, 3])
    def test_c4_trait_vs_oracle(self, prob, file_out, k):
        p = self.WXL + 1
        idx = sample(2 * self.FBLOCK, 40, 20 + k, must=list(range(2 * self.FBLOCK, self.M, 211)) + [self.M - 1])
        say(f"C4 trait {k}: oracle on {idx.size} SNPs (host dpotrf at n={self.N})")
        b_ref, V_ref = oracle_for(prob, k, idx)
        for prec, mixed in (("fp64", False), ("mixed", True)):
            B, V = file_out[prec]
            check(B[k][idx], V[k][idx], b_ref, V_ref, p, mixed, f"C4 trait {k} {prec} vs oracle")

    def test_c4_adjoint_one_trait(self, prob):
        """The adjoint of trait 0 at n = 30,000, p = 6 (2,048 SNPs): central
        differences with step 1e-6; the mixed adjoint within 1e-9 of FP64."""
        from cugwas_proto_amd.adjoint import snp_loss, snp_loss_grad
        m = 2048
        X = prob.X[:m]
        w = torch.from_numpy(np.random.default_rng(7).uniform(0.0, 1.0, m)).cuda()
        y, h = prob.Y[0].contiguous(), self.H[0]
        loss, gy, gh = snp_loss_grad(y, h, X, w, prob.Phi, prob.XL, 1.0, m)
        l8, gy8, gh8 = snp_loss_grad(y, h, X, w, prob.Phi, prob.XL, 1.0, m, precision="mixed")
        ry8 = float(torch.linalg.vector_norm(gy8 - gy, float("inf")) / torch.linalg.vector_norm(gy, float("inf")))
        say(f"C4 adjoint mixed vs fp64: loss {abs(l8 - loss) / abs(loss):.2e} dl/dy {ry8:.2e} "
            f"dl/dh2 {abs(gh8 - gh) / abs(gh):.2e}")
        assert abs(l8 - loss) <= TOL_ADJ * abs(loss) and ry8 <= TOL_ADJ and abs(gh8 - gh) <= TOL_ADJ * abs(gh)
        eps = 1e-6

        def f(yy, hh):
            with torch.no_grad():
                return float(snp_loss(yy, torch.tensor(hh, dtype=torch.float64), X, w, prob.Phi, prob.XL, 1.0, m))

        fd_h = (f(y, h + eps) - f(y, h - eps)) / (2 * eps)
        dvec = torch.from_numpy(np.random.default_rng(8).standard_normal(self.N)).cuda()
        fd_y = (f(y + eps * dvec, h) - f(y - eps * dvec, h)) / (2 * eps)
        gd = float(torch.dot(gy, dvec))
        say(f"C4 adjoint central differences: dl/dh2 {abs(fd_h - gh) / abs(gh):.2e}, "
            f"dl/dy.d {abs(fd_y - gd) / abs(gd):.2e}")
        assert abs(fd_h - gh) <= TOL_FD * abs(gh)
        assert abs(fd_y - gd) <= TOL_FD * abs(gd)
