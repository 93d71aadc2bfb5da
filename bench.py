#!/usr/bin/env python3
"""Throughput of the Cholesky FGLS hot path on one MI355X.

Workload (default): n = 10,000 individuals, p = 4 (intercept + 2 covariates +
1 SNP), m = 1,000,000 SNPs, one trait, h2 = 0.5 -- the streaming-scale
configuration the headline metric is quoted on, with all of X_R (80 GB FP64)
resident in HBM.  Synthetic data from the configuration's generators:
Phi = ZZ'/(2n) rescaled to unit diagonal, X_L = [1, N(0,1), N(0,1)],
X_R ~ Binomial(2, f_j), f_j ~ U(0.05, 0.5), y = X_L beta + eps, eps ~ N(0, M).

A step is one pass of the hot path over all m SNPs: per block of SNPs,
fgls_hip_block_reduce (triangular MFMA product L^-1 X_R with the per-SNP
reductions) then fgls_hip_block_finalize (p x p solves, b_i and V_i).  The
per-trait precompute (L^-1, X~_L, y~, res_sigma) is a one-off reported as
`trait_setup_s`, outside the steps.

Prints one JSON line (rank 0).  The CPU baseline is the reference's own
fgls_chol (compiled from its sources by oracle/Makefile) on a bounded sample
of the same workload, or the FP64 oracle when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SNPs/s at n=10k,p=4,m=1M: FP64 vs mixed-prec (rel err ≤1e-8); MFMA % peak"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return len(os.sched_getaffinity(0))


def make_inputs(torch, n, m, wXL, h2, seed, ldx, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    f64 = torch.float64
    Z = torch.randn((n, 2 * n), dtype=f64, device=dev, generator=g)
    Phi = (Z @ Z.T) / (2 * n)
    del Z
    Phi = 0.5 * (Phi + Phi.T)
    d = Phi.diagonal().sqrt()
    Phi = Phi / d[:, None] / d[None, :]
    Phi.fill_diagonal_(1.0)
    M = h2 * Phi
    M.diagonal().add_(1.0 - h2)
    L = torch.linalg.cholesky(M)
    del M
    XL = torch.empty((n, wXL), dtype=f64, device=dev)
    XL[:, 0] = 1.0
    XL[:, 1:] = torch.randn((n, wXL - 1), dtype=f64, device=dev, generator=g)
    beta = torch.randn(wXL, dtype=f64, device=dev, generator=g)
    y = XL @ beta + L @ torch.randn(n, dtype=f64, device=dev, generator=g)
    X = torch.empty((m, ldx), dtype=f64, device=dev)
    if ldx > n:
        X[:, n:].zero_()
    f = torch.rand(m, dtype=torch.float32, device=dev, generator=g) * 0.45 + 0.05
    C = 16384
    for i0 in range(0, m, C):
        cnt = min(C, m - i0)
        fc = f[i0:i0 + cnt, None]
        a = torch.rand((cnt, n), dtype=torch.float32, device=dev, generator=g) < fc
        b = torch.rand((cnt, n), dtype=torch.float32, device=dev, generator=g) < fc
        X[i0:i0 + cnt, :n] = a.to(f64) + b.to(f64)
        del a, b
    return Phi, L, XL, y, X


def cpu_baseline(kind, Phi, L, XL, y, X, n, wXL, h2, sample, threads):
    """Time the CPU path on `sample` SNPs of the same workload."""
    import numpy as np
    from oracle import ref_runner
    if kind == "auto":
        kind = "reference" if ref_runner.available() else "port"
    if kind == "reference":
        m0 = 4096
        with tempfile.TemporaryDirectory(prefix="fgls_cpu_") as d:
            Phi.cpu().numpy().astype("<f8").ravel(order="F").tofile(os.path.join(d, "Phi.in"))
            np.asarray([h2], "<f8").tofile(os.path.join(d, "H.in"))
            np.asarray([1.0], "<f8").tofile(os.path.join(d, "Sig.in"))
            XL.T.contiguous().cpu().numpy().astype("<f8").tofile(os.path.join(d, "XL.in"))
            y.cpu().numpy().astype("<f8").tofile(os.path.join(d, "Y.in"))
            X[:m0 + sample, :n].contiguous().cpu().numpy().astype("<f8").tofile(os.path.join(d, "XR.in"))
            # runs with m0 and m0 + sample SNPs after one warm-up run (thread
            # pool start, page cache): the difference of the two timed runs
            # excludes the per-trait one-offs (file reads, dpotrf of M,
            # precompute), which both contain.
            ref_runner.fgls_chol_dir(d, n, wXL + 1, m0, 1, wXL, 4096, threads=threads)
            t_a = ref_runner.fgls_chol_dir(d, n, wXL + 1, m0, 1, wXL, 4096, threads=threads)
            t_b = ref_runner.fgls_chol_dir(d, n, wXL + 1, m0 + sample, 1, wXL, 4096, threads=threads)
            t_a = min(t_a, ref_runner.fgls_chol_dir(d, n, wXL + 1, m0, 1, wXL, 4096, threads=threads))
        rate = sample / max(t_b - t_a, 1e-9)
        desc = (f"reference fgls_chol (oracle/_ref, OpenBLAS {threads} threads, x_b=4096): "
                f"n={n}, p={wXL + 1}; after a warm-up run, m={m0} (best of 2: {t_a:.2f}s) and "
                f"m={m0 + sample} ({t_b:.2f}s); rate = {sample} SNPs / time difference")
    else:
        from oracle import fgls_oracle as orc
        from oracle import ref_runner as rr
        rr.set_blas_threads(threads) if rr.available() else None
        Lc = L.cpu().numpy()
        tr = orc.trait_precompute(Lc, XL.cpu().numpy(), y.cpu().numpy())
        Xs = X[:sample, :n].cpu().numpy().T.copy()
        t0 = time.perf_counter()
        orc.solve_snps(Xs, tr, 1)
        rate = sample / (time.perf_counter() - t0)
        desc = f"FP64 oracle (numpy/scipy dtrsm + per-SNP stage), n={n}, p={wXL + 1}, {sample} SNPs"
    return {"value": rate, "unit": "SNPs/s", "cores": threads, "kind": kind, "sample": desc}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--m", type=int, default=1_000_000)
    ap.add_argument("--wXL", type=int, default=3)
    ap.add_argument("--h2", type=float, default=0.5)
    ap.add_argument("--block", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--check-snps", type=int, default=512)
    ap.add_argument("--cpu-baseline", choices=["auto", "reference", "port", "none"], default="auto")
    ap.add_argument("--cpu-snps", type=int, default=16384)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the SNP-range sharding across GPUs is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    if pkg.device_count() < 1 or not torch.cuda.is_available():
        raise SystemExit("no HIP device visible")
    dev = torch.device("cuda:0")
    n, m, wXL = args.n, args.m, args.wXL
    p = wXL + 1
    ldx = (n + 3) // 4 * 4

    t0 = time.perf_counter()
    Phi, L, XL, y, X = make_inputs(torch, n, m, wXL, args.h2, args.seed, ldx, dev)
    torch.cuda.synchronize()
    log(f"inputs ready: n={n} m={m} X_R {X.numel() * 8 / 1e9:.1f} GB in HBM ({time.perf_counter() - t0:.1f}s)")

    block = min(args.block, m)
    ctx = pkg.DeviceContext(n, wXL, 1, max_block=block)
    t0 = time.perf_counter()
    ctx.set_trait(L.T.contiguous(), XL.T.contiguous(), y.contiguous())
    torch.cuda.synchronize()
    trait_setup = time.perf_counter() - t0
    Bout = torch.empty((m, p), dtype=torch.float64, device=dev)
    Vout = torch.empty((m, p * p), dtype=torch.float64, device=dev)
    starts = list(range(0, m, block))
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in starts]

    def step(record):
        for bi, i0 in enumerate(starts):
            cnt = min(block, m - i0)
            if record:
                evs[bi][0].record()
            ctx.block_reduce(X[i0:i0 + cnt], cnt)
            if record:
                evs[bi][1].record()
            ctx.block_finalize(cnt, Bout[i0:i0 + cnt], Vout[i0:i0 + cnt], snp_base=i0)
            if record:
                evs[bi][2].record()

    for _ in range(args.warmup):
        step(False)
    ctx.check()
    torch.cuda.synchronize()
    reduce_ms, fin_ms, reduce_flops, nfull = 0.0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        for bi, i0 in enumerate(starts):
            cnt = min(block, m - i0)
            if cnt == block:
                reduce_ms += evs[bi][0].elapsed_time(evs[bi][1])
                fin_ms += evs[bi][1].elapsed_time(evs[bi][2])
                reduce_flops += float(n) * n * cnt
                nfull += 1
        log(f"step {k + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    ctx.check()
    # the event reads inside the loop add host time: use the event totals too
    value = m * args.steps / elapsed
    avg_reduce_ms = reduce_ms / max(nfull, 1)
    achieved_tflops = reduce_flops / (reduce_ms * 1e-3) / 1e12 if reduce_ms > 0 else 0.0

    # parity spot check against the FP64 oracle on a sample of SNPs
    check = {}
    if args.check_snps > 0:
        from oracle import fgls_oracle as orc
        idx = np.unique(np.concatenate([np.arange(min(args.check_snps // 2, m)),
                                        np.random.default_rng(0).integers(0, m, args.check_snps // 2)]))
        Lc = L.cpu().numpy()
        tr = orc.trait_precompute(Lc, XL.cpu().numpy(), y.cpu().numpy())
        Xs = X[torch.as_tensor(idx, device=dev), :n].cpu().numpy().T.copy()
        b_ref, V_ref = orc.solve_snps(Xs, tr, 1)
        Bg = Bout[torch.as_tensor(idx, device=dev)].cpu().numpy()
        Vg = Vout[torch.as_tensor(idx, device=dev)].cpu().numpy()
        check = {"snps": int(idx.size),
                 "max_rel_err_B": orc.max_rel_err(Bg, b_ref),
                 "max_rel_err_V": orc.max_rel_err(Vg, V_ref.transpose(0, 2, 1).reshape(idx.size, -1),
                                                  p=p, lower_only=True)}
        log(f"parity vs oracle on {idx.size} SNPs: {check}")

    cpu = None
    if args.cpu_baseline != "none":
        threads = host_threads()
        log(f"CPU baseline ({args.cpu_baseline}, {threads} threads, {args.cpu_snps} SNPs)")
        cpu = cpu_baseline(args.cpu_baseline, Phi, L, XL, y, X, n, wXL, args.h2, args.cpu_snps, threads)
        log(f"CPU baseline: {cpu['value']:.1f} SNPs/s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "SNPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (configuration generators, seeded); X_R resident in HBM",
        "config": {"workload": f"Cholesky FGLS, FP64 path: n={n}, p={p}, m={m}, t=1, h2={args.h2}",
                   "n": n, "p": p, "m": m, "block": block, "path": "fp64"},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "k_trmm<EPI_REDUCE> (fgls_hip_block_reduce)",
                     "avg_launch_ms": avg_reduce_ms, "snps_per_launch": block,
                     "flop_per_snp": float(n) * n},
        "cpu_baseline": cpu,
        "trait_setup_s": trait_setup,
        "finalize_avg_ms": fin_ms / max(nfull, 1),
        "parity_vs_oracle": check,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
