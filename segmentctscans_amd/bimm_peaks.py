"""The bimm-peaks driver (bimm-peaks/bimm-peaks.cpp:236-699) over the C ABI.

    python -m segmentctscans_amd.bimm_peaks in.bin --peaks "{p0,p1,...}" [--pdf-name base]
        [--tailthr T | --lotailthr L --hitailthr H] [--vtailthr V] [--maxds D] [--seed S]

Reads a float .bin volume, takes its gradient magnitude on the GPU, cuts the
intensity histogram's tails (1024 bins, dithered by one raw unit) and the
gradient histogram's upper tail, samples eligible voxels (0 < v <= the
gradient cut, intensity inside the cuts), and fits the bivariate mixture of
segmentctscans_amd/bimm.py: a simplex over 100 batches of 10,000 samples,
then over all of them.  Prints the fitted deviations, interface width,
correlation, peak intensities (raw units) and weights, as the reference does;
--pdf-name writes the fitted u marginals (full, per phase, per interface) as
text tables.

Differences from the reference: the gradient range comes from
grad_min_max (the reference's three-value binding of grad_magn at
bimm-peaks.cpp:312 no longer matches grad_magn.h, which returns the field
alone); sampling and batch order use numpy's PCG64 with --seed instead of
std::random_device; the simplex is the textbook Nelder-Mead rather than GSL's
nmsimplex2.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

from . import archive, calc_hist, grad_magn, grad_min_max, make_context, upload
from .bimm import UVSamples, fit, marginal_interface, marginal_pure, pairs
from .gmm_peaks import linearize


def tail_cutoffs(hist: np.ndarray, lo_thr: float, hi_thr: float, rs: float, re_: float) -> tuple[float, float]:
    """bimm-peaks.cpp:174-203: the lo_thr and 1 - hi_thr quantiles of the
    histogram, linear within a bin."""
    width = (re_ - rs) / hist.size
    total = float(hist.sum())

    def quantile(p):
        if p <= 0:
            return float(np.float32(rs))
        if p >= 1:
            return float(np.float32(re_))
        target, cum = p * total, 0.0
        for i, h in enumerate(hist):
            nxt = cum + float(h)
            if h > 0 and target <= nxt:
                return float(np.float32(rs + (i + (target - cum) / h) * width))
            cum = nxt
        return float(np.float32(re_))

    return quantile(lo_thr), quantile(1.0 - hi_thr)


def parse_peaks(s: str) -> list[float]:
    s = s.strip()
    if not (s.startswith("{") and s.endswith("}")):
        raise argparse.ArgumentTypeError("peaks are written {p0,p1,...}")
    vals = [float(x) for x in s[1:-1].split(",")]
    if not 2 <= len(vals) <= 6:
        raise argparse.ArgumentTypeError("2 to 6 peaks")
    return vals


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="bimm-peaks")
    ap.add_argument("input")
    ap.add_argument("--peaks", required=True, type=parse_peaks, help="initial peaks {p0,p1,...} (raw units)")
    ap.add_argument("--pdf-name", default=None)
    ap.add_argument("--tailthr", type=float, default=None)
    ap.add_argument("--lotailthr", type=float, default=0.01)
    ap.add_argument("--hitailthr", type=float, default=0.01)
    ap.add_argument("--vtailthr", type=float, default=0.01)
    ap.add_argument("--maxds", type=float, default=1.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--batch", type=int, default=10_000, help="samples per batch (reference: 10,000)")
    ap.add_argument("--batches", type=int, default=100, help="batches (reference: 100)")
    ap.add_argument("--restart-period", type=int, default=100)
    ap.add_argument("--full-iters", type=int, default=10_000)
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args(argv)
    if not os.path.exists(args.input):
        print(f"Input file {args.input} doesn't exist", file=sys.stderr)
        return 1
    lo_thr, hi_thr = (args.tailthr, args.tailthr) if args.tailthr is not None else (args.lotailthr, args.hitailthr)

    t0 = time.time()
    img = archive.load_bin(args.input)
    lo, hi = img.lo, img.hi
    ctx = make_context(0)
    d_img = upload(ctx, img)
    d_v = grad_magn(ctx, d_img)
    _, vmax = grad_min_max(ctx, d_v)
    v_all = d_v.to_host()
    # intensity tails: 1024 bins over [min, max], dithered by one raw unit
    jitter = 1.0 / (hi - lo)
    jitter = jitter if jitter * jitter < 0.5 / 1024 else 0.0
    hu, rs, re_ = calc_hist(ctx, d_img, 1024, jitter, seed=args.seed)
    lo_u, hi_u = tail_cutoffs(hu, lo_thr, hi_thr, rs, re_)
    print(f"Fitting [{lo_u * (hi - lo) + lo}, {hi_u * (hi - lo) + lo}] range of intensities")
    hv = calc_hist(ctx, d_v, 0.0, vmax, 1024, 0.0)
    lo_v, hi_v = tail_cutoffs(hv, 0.0, args.vtailthr, 0.0, vmax)
    print(f"Fitting [{lo_v * (hi - lo)}, {hi_v * (hi - lo)}] range of gradients")
    d_v.close()
    d_img.close()

    vf = img.__class__(img.mapping, v_all)
    u_lin, v_lin = linearize(img).astype(np.float64), linearize(vf).astype(np.float64)
    eligible = np.flatnonzero((v_lin > 0) & (v_lin <= hi_v) & (u_lin >= lo_u) & (u_lin <= hi_u))
    n = args.batch * args.batches
    if eligible.size < n:
        print("Not enough eligible voxels", file=sys.stderr)
        return 1
    rng = np.random.default_rng(args.seed)
    idx = rng.permutation(rng.choice(eligible, n, replace=False))
    samples = UVSamples(ctx, u_lin[idx], v_lin[idx])
    peaks = [(p - lo) / (hi - lo) for p in args.peaks]

    def log(it, cost, size, p):
        if args.quiet:
            return None
        w = " ".join(f"{x:9.6f}" for x in p.w)
        I = " ".join(f"{x:6.3f}" for x in p.I)
        print(f"{it:5} {cost:14.9f} {size:9.6f} |{w} |{I} |{math.exp(p.log_sn):7.4f}{math.exp(p.log_sb):7.4f}"
              f"{math.exp(p.log_ds):7.4f}{p.rho:7.4f}", flush=True)
        return None

    res = fit(samples, (lo_u, hi_u, hi_v), peaks, max_ds=args.maxds, batch=args.batch, batches=args.batches,
              restart_period=args.restart_period, full_iters=args.full_iters, log=log)
    samples.close()
    p = res.params
    K = p.I.size
    print(f"sigma_b = {math.exp(p.log_sb)}, sigma_n = {math.exp(p.log_sn) * (hi - lo)}, "
          f"ds = {math.exp(p.log_ds)}, rho = {p.rho}")
    for i in range(K):
        print((hi - lo) * p.I[i] + lo)
    for i in range(K):
        print(f"w[{i}] = {p.w[i]}")
    for k, (i, j) in enumerate(pairs(K)):
        print(f"w[{i}, {j}] = {p.w[K + k]}")
    if args.pdf_name:
        write_pdfs(args.pdf_name, p, lo, hi, lo_thr, hi_thr)
    print(f"[bimm-peaks] {time.time() - t0:.1f} s, {res.iterations[0]} + {res.iterations[1]} simplex iterations",
          file=sys.stderr)
    return 0


def write_pdfs(base: str, p, lo: float, hi: float, lo_thr: float, hi_thr: float, steps: int = 512) -> None:
    """bimm-peaks.cpp:600-690: the fitted u marginals at steps + 1 points of
    u = lo_thr + k (hi_thr - lo_thr) / steps (the reference's grid), in raw
    units: <base>-full.txt, <base>-NN.txt per phase, <base>-NN-MM.txt per
    interface."""
    K = p.I.size
    sn, ds = math.exp(p.log_sn), math.exp(p.log_ds)
    us = [lo_thr + k * (hi_thr - lo_thr) / steps for k in range(steps + 1)]
    comps = [(f"{base}-{c:02}.txt", lambda u, c=c: p.w[c] * marginal_pure(u, p.I[c], sn)) for c in range(K)]
    comps += [(f"{base}-{i:02}-{j:02}.txt",
               lambda u, i=i, j=j, k=k: p.w[K + k] * marginal_interface(u, p.I[i], p.I[j], sn, ds))
              for k, (i, j) in enumerate(pairs(K))]
    table = {name: [f(u) for u in us] for name, f in comps}
    with open(f"{base}-full.txt", "w") as fh:
        fh.write("u pdf\n")
        for k, u in enumerate(us):
            fh.write(f"{u * (hi - lo) + lo} {sum(v[k] for v in table.values()) / (hi - lo)}\n")
    for name, vals in table.items():
        with open(name, "w") as fh:
            fh.write("u pdf\n")
            for u, v in zip(us, vals):
                fh.write(f"{u * (hi - lo) + lo} {v / (hi - lo)}\n")


if __name__ == "__main__":
    sys.exit(main())
