#!/usr/bin/env python
"""Benchmark of the Potts primal-dual iteration on one MI355X.

Workload (default): K = 3 Potts, lambda = 30, tau = tau' = 1/sqrt(12), FP32,
on the 1024 x 1024 x 1426 synthetic phantom (cylinder radius 460 about
(511.5, 511.5), spheres U[4, 40], 947,999,096 in-mask voxels).  A step is one
primal-dual iteration over the whole volume; its kernels run on the library's
stream and are timed there with HIP events, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--storage fp32|bf16|fp16]

Prints one JSON line.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

T0 = time.time()

# Algorithmic bytes per in-mask voxel-iteration: RW Xi 3(K-1), RW V (K-1),
# RW V-bar (K-1), R u; BF16/FP16 storage halves Xi and V-bar.
def algo_bytes(K: int, storage: str) -> int:
    if storage == "fp32":
        return 4 * (10 * (K - 1) + 1)
    return 24 * (K - 1) + 4


HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench {time.time() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ph, c, tau, iters: int, size: int) -> dict:
    """FP32 host transcription (the oracle's FP32 mirror, OpenMP over lines) on a
    size^3 crop of the same phantom."""
    import oracle as O
    from segmentctscans_amd.phantom import crop_box

    m = ph.field.mapping
    n0 = (m.n_images - size) // 2
    r0 = c0 = (m.height - size) // 2
    crop = crop_box(ph, n0, r0, c0, size)
    lay = O.Layout.from_mapping(crop.mapping)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    t = time.perf_counter()
    O.potts(lay, crop.data, c, 30.0, tau, tau, iters, precision="f32", nthreads=cores)
    dt = time.perf_counter() - t
    vox = crop.mapping.data_size
    return {"value": vox * iters / dt, "unit": "voxel-iters/s", "cores": cores, "kind": "port",
            "sample": f"{size}^3 central crop of the same phantom, K=3, {iters} iterations, "
                      f"FP32 host mirror of the reference update ({dt:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--storage", default="fp32", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--kernel", default="auto", choices=["auto", "twopass", "fused"])
    ap.add_argument("--n", type=int, default=1426)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--radius", type=float, default=460.0)
    ap.add_argument("--rmax", type=float, default=40.0)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path in this build "
              "(run with --gpus 1).", file=sys.stderr)
        return 2

    import segmentctscans_amd as S
    from segmentctscans_amd.hiptime import EventTimeline
    from segmentctscans_amd.phantom import make_phantom

    K, lam = 3, 30.0
    tau = float(np.float32(1.0 / np.sqrt(12.0)))
    log(f"phantom {args.side}^2 x {args.n}, r={args.radius}, spheres U[4,{args.rmax}]")
    ph = make_phantom(args.n, args.side, 4.0, args.rmax, seed=args.seed, r=args.radius)
    c = ph.normalized_labels()
    vox = ph.field.mapping.data_size
    log(f"phantom done: {vox} in-mask voxels, {ph.n_spheres} spheres, porosity {ph.porosity:.3f}")

    ctx = S.make_context(0)
    g = S.upload(ctx, ph.field)
    solver = S.PottsSolver(ctx, g, c, lam, tau, tau, storage=args.storage, kernel=args.kernel)
    log("uploaded; warmup")
    solver.iterate(args.warmup)
    ctx.synchronize()

    tl = EventTimeline(ctx.stream, args.steps + 1)
    ctx.synchronize()
    w0 = time.perf_counter()
    tl.mark()
    for _ in range(args.steps):
        solver.iterate(1)
        tl.mark()
    ctx.synchronize()
    wall = time.perf_counter() - w0
    durs = np.array(tl.durations())
    tl.close()
    total_ms = float(durs.sum())
    ms_step = total_ms / args.steps
    value = vox * args.steps / (total_ms / 1e3)
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step (events), wall {wall:.3f} s; "
        f"min {durs.min():.3f} max {durs.max():.3f} ms")

    bpv = algo_bytes(K, args.storage)
    achieved = bpv * vox / (ms_step / 1e3) / 1e9
    result = {
        "metric": "Potts voxel-iters/s + % HBM peak; BF16 vs FP32 speedup; adjoint/forward ratio",
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "potts_k3_fp32" if args.storage == "fp32" else f"potts_k3_{args.storage}_storage",
            "volume_n_rows_cols": [args.n, args.side, args.side],
            "cylinder_radius": args.radius,
            "in_mask_voxels": vox,
            "K": K, "lambda": lam, "tau": tau,
            "storage": args.storage,
            "kernel": args.kernel,
            "phantom_seed": args.seed,
        },
        "hbm_peak_frac": achieved / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_voxel_iter": bpv,
        },
    }
    solver.close()
    g.close()
    if not args.no_cpu_baseline:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(ph, c, tau, args.cpu_iters, args.cpu_size)
        log(f"CPU baseline {result['cpu_baseline']['value']:.3e} voxel-iters/s")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
