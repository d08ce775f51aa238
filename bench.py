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

# Algorithmic HBM bytes per in-mask voxel of one launch of each kernel: every
# array the kernel must read or write, once (neighbour re-reads excluded).
# sb = bytes per stored Xi / V-bar element (4 FP32, 2 BF16/FP16); V and u FP32.
#   Xi kernel  R V-bar (K-1), R+W Xi 3(K-1)                     7(K-1) sb
#   V kernel   R Xi 3(K-1), R+W V (K-1), W V-bar (K-1), R u     4(K-1) sb + 8(K-1) + 4
#   fused      R+W Xi 3(K-1), R+W V-bar (K-1), R+W V (K-1), R u  8(K-1) sb + 8(K-1) + 4
def kernel_bytes(K: int, storage: str) -> dict:
    """Algorithmic bytes per in-mask voxel of one launch.  The fused kernel
    runs the V step of one iteration and the Xi step of the next: RW Xi at the
    storage size, RW V in FP32, R u (V-bar never leaves the chip)."""
    sb = 4 if storage == "fp32" else 2
    return {"potts_xi_kernel": 7 * (K - 1) * sb,
            "potts_v_kernel": 4 * (K - 1) * sb + 8 * (K - 1) + 4,
            "potts_fused_kernel": 6 * (K - 1) * sb + 8 * (K - 1) + 4}


def single_pass_bytes(K: int, storage: str) -> int:
    """Compulsory bytes of one iteration done in a single pass with V-bar in
    memory: RW Xi and RW V-bar at the storage size, RW V in FP32, R u (84 B
    for K = 3 in FP32)."""
    sb = 4 if storage == "fp32" else 2
    return 8 * (K - 1) * sb + 8 * (K - 1) + 4


KERNEL_NAMES = {"fused": ["potts_fused_kernel"], "twopass": ["potts_xi_kernel", "potts_v_kernel"]}
KERNEL_NAMES["auto"] = KERNEL_NAMES["twopass"]


def committed_traffic(kernel: str, workload: str):
    """HBM bytes per launch of kernel on workload, from the committed rocprofv3
    --pmc summary profiles/<kernel>_<workload>_traffic.md (a line
    'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{kernel}_{workload}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip()), os.path.relpath(path, os.path.dirname(path) + "/..")
    except OSError:
        pass
    return None, None


HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench {time.time() - T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def host_cores() -> dict:
    """The host cores the CPU baseline runs on: every core this process may
    use (sched_getaffinity), at most OMP_NUM_THREADS when that is set (the
    GPU's share of a shared host); with nproc and the CPU model."""
    import subprocess
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    use = min(aff, int(share)) if share.isdigit() and int(share) > 0 else aff
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc", "--all"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count() or 0
    return {"threads": use, "affinity_cores": aff, "nproc_all": nproc, "omp_num_threads": share or None,
            "cpu_model": model}


def _median_time(fn, reps: int) -> float:
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def cpu_baseline(ph, c, tau, iters: int, size: int, tv_ph, tv_out, hist_range, fp64, reps: int = 3) -> dict:
    """The reference CPU path: the oracle's FP32 host transcription of the
    reference kernels (OpenMP over lines), timed with a monotonic clock
    around the iteration loop, median of reps runs:
      - the headline: the 256^3 central crop of the 1426 x 1024^2 phantom,
        K = 3, iters iterations (voxel-iters/s);
      - the 64^3 phantom (K = 3, lambda 30, 500 iterations), a full run;
      - Chambolle TV on the 256^3 central crop of the 512^3 phantom, 20 iterations;
      - calc_hist at 65536 bins over the GPU's 512^3 TV result (one thread);
    and the FP64 oracle's time on the agreement box (not a baseline)."""
    import oracle as O
    from segmentctscans_amd.phantom import crop_box, make_phantom

    hc = host_cores()
    nt = hc["threads"]
    m = ph.field.mapping
    crop = crop_box(ph, (m.n_images - size) // 2, (m.height - size) // 2, (m.width - size) // 2, size)
    lay = O.Layout.from_mapping(crop.mapping)
    vox = crop.mapping.data_size
    dt = _median_time(lambda: O.potts(lay, crop.data, c, 30.0, tau, tau, iters, precision="f32", nthreads=nt), reps)
    out = {"value": vox * iters / dt, "unit": "voxel-iters/s", "cores": nt, "kind": "port",
           "sample": f"{size}^3 central crop of the same phantom ({vox} voxels), K=3, {iters} iterations, "
                     f"FP32 host mirror of the reference update, {nt} threads, median of {reps} "
                     f"({dt:.2f} s each)",
           "achieved_GBs_at_84B": vox * iters * 84 / dt / 1e9, **hc}
    c1 = make_phantom(64, 64, 4.0, 12.0, seed=1)
    l1 = O.Layout.from_mapping(c1.field.mapping)
    v1 = c1.field.mapping.data_size
    c1c = c1.normalized_labels()
    d1 = _median_time(lambda: O.potts(l1, c1.field.data, c1c, 30.0, tau, tau, 500, precision="f32", nthreads=nt),
                      reps)
    out["c1_full_run"] = {"workload": f"64^3 phantom ({v1} voxels), K=3, lambda 30, 500 iterations",
                          "seconds": d1, "voxel_iters_per_s": v1 * 500 / d1, "threads": nt}
    tm = tv_ph.field.mapping
    tcrop = crop_box(tv_ph, (tm.n_images - size) // 2, (tm.height - size) // 2, (tm.width - size) // 2, size)
    tl = O.Layout.from_mapping(tcrop.mapping)
    tv_tau = float(np.float32(1.0 / 6.0) * np.float32(0.99))
    dtv = _median_time(lambda: O.chambolle(tl, tcrop.data, 0.08, tv_tau, 20, precision="f32", nthreads=nt), reps)
    tvox = tcrop.mapping.data_size
    out["tv_crop"] = {"workload": f"{size}^3 central crop of the 512^3 TV phantom, lambda 0.08, 20 iterations",
                      "seconds": dtv, "voxel_iters_per_s": tvox * 20 / dtv, "threads": nt,
                      "achieved_GBs_at_28B": tvox * 20 * 28 / dtv / 1e9}
    if tv_out is not None:
        hl = O.Layout.from_mapping(tm)
        dh = _median_time(lambda: O.calc_hist(hl, tv_out, hist_range[0], hist_range[1], 65536), reps)
        out["calc_hist_65536"] = {"workload": "65536 bins over the 512^3 TV result (lambda 0.12), "
                                              f"{tm.data_size} voxels", "seconds": dh,
                                  "voxels_per_s": tm.data_size / dh, "threads": 1}
    if fp64 is not None:
        out["fp64_oracle"] = fp64
    return out


def adjoint_ratio(S) -> dict:
    """Reverse pass / forward pass of the adjoint on the 128^3 workload (K = 3,
    200 iterations; one warm-up, then the median of 5 runs), with checkpoints
    every 16 (the states inside each segment recomputed) and with every state
    stored (no recompute).  The every-state ratio is priced against the
    checkpointed run's forward pass, the plain forward: its own forward also
    writes 200 fresh state buffers and has varied between runs."""
    from segmentctscans_amd import Field
    from segmentctscans_amd.phantom import make_phantom
    tau = float(np.float32(1.0 / np.sqrt(12.0)))
    ph = make_phantom(128, 128, 4.0, 12.0, seed=5)
    ctx = S.make_context(0)
    g = S.upload(ctx, ph.field)
    t = S.upload(ctx, Field(ph.field.mapping, ph.labels.astype(np.float32)))
    c = ph.normalized_labels()

    def timed(ck):
        runs = [S.potts_adjoint(ctx, g, t, c, 30.0, tau, tau, 200, checkpoint_every=ck, want_grad_u=False)
                for _ in range(6)][1:]
        return [float(r["forward_ms"]) for r in runs], [float(r["backward_ms"]) for r in runs]
    fw, rv = timed(16)
    fa, ra = timed(1)
    g.close()
    t.close()
    f16, r16, fall, rall = (float(np.median(x)) for x in (fw, rv, fa, ra))
    return {"workload": "128^3 cylinder (1,334,784 voxels), K=3, 200 iterations, checkpoints every 16",
            "forward_ms": f16, "reverse_ms": r16, "ratio": r16 / f16, "ratio_total": (f16 + r16) / f16,
            "forward_ms_runs": fw, "reverse_ms_runs": rv,
            "every_state_stored": {"reverse_ms": rall, "ratio": rall / f16,
                                   "own_forward_ms": fall, "ratio_own_forward": rall / fall,
                                   "own_forward_ms_runs": fa, "reverse_ms_runs": ra, "states": 201}}


TV_BYTES = 28        # R xi 3, R g, W xi 3 (FP32) per voxel and iteration
TV_UPDATE_BYTES = 20  # final g - lambda div xi: R xi 3, R+W g


def tv_workload(S, iters: int, keep: dict | None = None) -> dict:
    """Chambolle TV then the 65536-bin histogram on the 512^3 phantom (cylinder
    radius 230.4, spheres U[4, 20]): for lambda = 0.04, 0.05, ..., 0.12, TV
    from the same input (tau = 0.99 / 6, iters iterations), then calc_hist of
    the result over the input's [min, max] with 65536 bins and no dither.
    Each call is timed with HIP events on the library's stream; a call
    includes zeroing both xi buffers (24 B per stored element) and the final
    update, so the time of one iteration is (call - call with 0 iterations)
    / iters, and its achieved bandwidth is 28 B per voxel over that time.
    voxel_iters_per_s is over the whole call.  keep (optional) receives the
    phantom ("ph"), the input's range ("range") and the result of the last
    weight ("out", host array) for the CPU baseline."""
    from segmentctscans_amd.hiptime import EventTimeline
    from segmentctscans_amd.phantom import make_phantom

    tau = float(np.float32(1.0 / 6.0) * np.float32(0.99))
    ph = make_phantom(512, 512, 4.0, 20.0, seed=7, r=230.4)
    m = ph.field.mapping
    vox = m.data_size
    ctx = S.make_context(0)
    g0 = S.upload(ctx, ph.field)
    mn, mx = S.min_max(ctx, g0)
    g0.close()
    g = S.upload(ctx, ph.field)
    S.chambolle(ctx, g, 0.04, tau, 5)   # warm-up
    g.close()
    tv_ms, hist_ms, in_range = [], [], []
    lams = [round(0.04 + 0.01 * k, 2) for k in range(9)]
    for lam in lams:
        g = S.upload(ctx, ph.field)
        ctx.synchronize()
        tl = EventTimeline(ctx.stream, 3)
        tl.mark()
        S.chambolle(ctx, g, lam, tau, iters)
        tl.mark()
        cnt = S.calc_hist(ctx, g, mn, mx, 65536, 0.0, counts=True)
        tl.mark()
        d = tl.durations()
        tl.close()
        if keep is not None and lam == lams[-1]:
            keep["out"] = g.to_host()
        g.close()
        tv_ms.append(float(d[0]))
        hist_ms.append(float(d[1]))
        in_range.append(int(cnt.sum()))
    # the same call with no iteration: set-up, zeroing and the final update
    base_ms = []
    for _ in range(3):
        g = S.upload(ctx, ph.field)
        ctx.synchronize()
        tl = EventTimeline(ctx.stream, 2)
        tl.mark()
        S.chambolle(ctx, g, 0.08, tau, 0)
        tl.mark()
        base_ms.append(float(tl.durations()[0]))
        tl.close()
        g.close()
    t = float(np.median(tv_ms))
    t0 = float(np.median(base_ms))
    it_ms = (t - t0) / iters
    achieved = vox * TV_BYTES / (it_ms / 1e3) / 1e9
    out = {"workload": f"512^3 cylinder r=230.4 ({vox} voxels), spheres U[4,20], lambda 0.04..0.12 x "
                       f"{iters} iterations, tau=0.99/6; then 65536-bin calc_hist over [min, max]",
           "voxel_iters_per_s": vox * iters / (t / 1e3),
           "ms_per_call_median": t, "ms_per_iteration": it_ms, "ms_call_without_iterations": t0,
           "calls_ms": tv_ms, "hist_ms": hist_ms, "hist_in_range": in_range,
           "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS,
           "bytes_per_voxel_iter": TV_BYTES}
    traffic, src = committed_traffic("tv_step_kernel", "tv_512")
    out["traffic"], out["traffic_source"] = traffic, src
    if keep is not None:
        keep["ph"] = ph
        keep["range"] = (mn, mx)
    return out


def long_run(S, ctx, g, c, lam, tau, iters: int, gap_every: int, storage: str = "fp32") -> dict:
    """The full solve of the 1426 x 1024^2 workload: iters iterations with
    the FP64 duality gap every gap_every iterations (potts_min_partition with
    a gap trace), timed with HIP events; the gap evaluations are inside."""
    from segmentctscans_amd.hiptime import EventTimeline
    tl = EventTimeline(ctx.stream, 2)
    ctx.synchronize()
    tl.mark()
    res, gaps = S.potts_min_partition(ctx, g, c, lam, tau, tau, iters, gap_every=gap_every, storage=storage)
    tl.mark()
    ms = float(tl.durations()[0])
    tl.close()
    res.close()
    return {"storage": storage, "iterations": iters, "gap_every": gap_every, "ms_total": ms,
            "ms_per_iteration": ms / iters, "relative_gap_trace": [float(x) for x in gaps]}


class Fp64Box:
    """The oracle's FP64 solve of a size^3 box cut from the centre of the
    workload (box domain, Neumann edges), run on a host thread while the GPU
    legs run (the oracle releases the GIL)."""

    def __init__(self, ph, c, lam, tau, iters: int, size: int, threads: int):
        import threading
        from segmentctscans_amd.phantom import crop_box
        m = ph.field.mapping
        self.box = crop_box(ph, (m.n_images - size) // 2, (m.height - size) // 2, (m.width - size) // 2, size)
        self.args = (c, lam, tau, iters)
        self.size, self.threads, self.iters = size, threads, iters
        self.result, self.seconds, self.error = None, None, None
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def _run(self):
        import oracle as O
        try:
            c, lam, tau, iters = self.args
            lay = O.Layout.from_mapping(self.box.mapping)
            t = time.perf_counter()
            self.result = O.potts(lay, self.box.data, c, lam, tau, tau, iters, precision="f64", want_state=True,
                                  nthreads=self.threads)
            self.seconds = time.perf_counter() - t
            log(f"FP64 oracle on the {self.size}^3 box: {self.seconds:.1f} s")
        except Exception as e:   # reported in the JSON line
            self.error = repr(e)

    def join(self):
        self.thread.join()
        if self.error:
            raise RuntimeError(self.error)
        return self.result

    def timing(self) -> dict:
        vox = self.box.mapping.data_size
        return {"workload": f"{self.size}^3 box, K=3, {self.iters} iterations, FP64", "seconds": self.seconds,
                "voxel_iters_per_s": vox * self.iters / self.seconds, "threads": self.threads}


def agreement_gpu(S, job: Fp64Box, c, lam, tau) -> dict:
    """V and labels of the agreement box after job.iters iterations on the GPU
    with FP32, BF16 and FP16 storage (in-mask voxels only)."""
    box, iters = job.box, job.iters
    valid = box.mapping.valid_mask()
    ctx = S.make_context(0)
    got = {}
    for st in ("fp32", "bf16", "fp16"):
        gb = S.upload(ctx, box)
        sol = S.PottsSolver(ctx, gb, c, lam, tau, tau, storage=st)
        sol.iterate(iters)
        V, _ = sol.state()
        got[st] = (V[:, valid], sol.labels().to_host()[valid])
        sol.close()
        gb.close()
    log(f"agreement box solved on the GPU ({', '.join(got)})")
    return got


def crop_agreement(job: Fp64Box, got: dict) -> dict:
    """Label agreement and ||V - V_fp64||_inf of each storage type against the
    oracle's FP64 solve of the same box."""
    valid = job.box.mapping.valid_mask()
    r64 = job.join()
    out = {"box": job.size, "iterations": job.iters, "voxels": int(valid.sum()), "fp64_seconds": job.seconds}
    for st, (V, res) in got.items():
        out[st] = {"label_agreement": float((res == r64["res"][valid]).mean()),
                   "v_inf": float(np.abs(V - r64["V"][:, valid]).max())}
        log(f"{job.size}^3 box, {st}: agreement {out[st]['label_agreement']:.6f}, "
            f"||V - V64||inf {out[st]['v_inf']:.3e}")
    return out


def reductions_workload(S, reps: int = 5) -> dict:
    """The reductions on the 512^3 phantom (85,370,880 voxels): min_max,
    calc_hist with 1024 and 65536 bins over [min, max], grad_magn,
    grad_min_max and the 512 x 512 uv_hist; and the ingest kernels on the
    same voxels as u16 (phantom x 30000 + 12000): find_intensity_range (its
    65536-bin histogram and tail cut) and convert_to_fp to a float field.
    Each call is timed with HIP events on the library's stream (median of
    reps); achieved bandwidth is the algorithmic bytes per voxel (4 for one
    float read, 8 for two arrays, 2 for a u16 read, 6 for u16 in and float
    out) over that time.  Calls include their set-up and, where the result
    goes to the host, the copy back."""
    from segmentctscans_amd import Field
    from segmentctscans_amd.hiptime import EventTimeline
    from segmentctscans_amd.phantom import make_phantom
    ph = make_phantom(512, 512, 4.0, 20.0, seed=7)
    vox = ph.field.mapping.data_size
    ctx = S.make_context(0)
    u = S.upload(ctx, ph.field)
    v = S.grad_magn(ctx, u)
    mn, mx = S.min_max(ctx, u)
    v0, v1 = S.grad_min_max(ctx, v)
    valid = ph.field.mapping.valid_mask()
    raw = np.zeros(ph.field.mapping.span, np.uint16)
    raw[valid] = np.clip(np.round(ph.field.data[valid] * 30000 + 12000), 0, 65535).astype(np.uint16)
    r16 = S.upload(ctx, Field(ph.field.mapping, raw))
    ops = {
        "min_max": (4, lambda: S.min_max(ctx, u)),
        "calc_hist_1024": (4, lambda: S.calc_hist(ctx, u, mn, mx, 1024, 0.0, counts=True)),
        "calc_hist_65536": (4, lambda: S.calc_hist(ctx, u, mn, mx, 65536, 0.0, counts=True)),
        "grad_magn": (8, lambda: S.grad_magn(ctx, u).close()),
        "grad_min_max": (4, lambda: S.grad_min_max(ctx, v)),
        "uv_hist_512": (8, lambda: S.uv_hist(ctx, u, v, mn, mx, v0, v1, 512)),
        "find_intensity_range_u16": (2, lambda: S.find_intensity_range(ctx, r16, 0.001)),
        "convert_to_fp_u16": (6, lambda: S.convert_to_fp(ctx, r16, 14000.0, 40000.0).close()),
    }
    out = {"workload": f"512^3 cylinder r=230.4 ({vox} voxels)"}
    for name, (bpv, fn) in ops.items():
        fn()
        ms = []
        for _ in range(reps):
            ctx.synchronize()
            tl = EventTimeline(ctx.stream, 2)
            tl.mark()
            fn()
            tl.mark()
            ms.append(float(tl.durations()[0]))
            tl.close()
        t = float(np.median(ms))
        out[name] = {"ms": t, "bytes_per_voxel": bpv, "achieved_GBs": bpv * vox / (t / 1e3) / 1e9,
                     "frac": bpv * vox / (t / 1e3) / 1e9 / HBM_PEAK_GBS}
    r16.close()
    v.close()
    u.close()
    return out


def copy_calibration(mib: int) -> int:
    """A known-size copy: FETCH_SIZE and WRITE_SIZE of its kernel should both read mib MiB."""
    import ctypes as C
    from segmentctscans_amd.hiptime import hip
    h = hip()
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(a), n) == 0 and h.hipMalloc(C.byref(b), n) == 0
    for _ in range(3):
        assert h.hipMemcpy(b, a, n, 3) == 0   # hipMemcpyDeviceToDevice
    h.hipFree(a)
    h.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))
    return 0


def batch_workload(S, seed: int, iters: int, autotune: bool = True) -> dict:
    """576 solves: 64 sub-volumes with edges U[48, 192] per axis (cylinder or
    box, K ~ U{3, 4, 5}, extra phases evenly spaced in [0.030, 0.060] raw)
    x lambda in {20, 25, ..., 60}, iters iterations each: one batched call
    (one captured hipGraph launch for the iterations) against the same solves
    looped through potts_min_partition.  The batched call runs three times:
    with the default Xi tile, with the autotuned one (its trial iterations
    lie outside the event-timed iterations), and as one persistent launch per
    K group (autotuned tiles and workgroups per CU)."""
    from segmentctscans_amd.phantom import C4_LAMBDAS, c4_volumes

    tau = float(np.float32(1.0 / np.sqrt(12.0)))
    ctx = S.make_context(0)
    vols = c4_volumes(seed=seed)
    dev = [S.upload(ctx, v.field) for v in vols]
    probs = [(g, v.c, lam) for v, g in zip(vols, dev) for lam in C4_LAMBDAS]
    vox = sum(g.mapping.data_size for (g, _, _) in probs)
    log(f"batch: {len(probs)} solves, {vox} voxels in all, {iters} iterations")
    S.potts_batch(ctx, probs[:4], tau, tau, 2)   # warm-up (code objects)
    t = time.perf_counter()
    outs0, ms_default = S.potts_batch(ctx, probs, tau, tau, iters, graph=True)
    wall_default = time.perf_counter() - t
    for o in outs0:
        o.close()
    log(f"batched, default tile: iterations {ms_default:.1f} ms (events), call {wall_default:.2f} s")
    t = time.perf_counter()
    outs, ms = S.potts_batch(ctx, probs, tau, tau, iters, graph=True, autotune=autotune)
    wall_batch = time.perf_counter() - t
    tiles = S.potts_batch_tiles(ctx)
    log(f"batched, {'autotuned' if autotune else 'default'} tiles {tiles}: iterations {ms:.1f} ms (events), "
        f"call {wall_batch:.2f} s")
    # one persistent launch per K group (tiles from one ordered ticket queue)
    t = time.perf_counter()
    outs_p, ms_p = S.potts_batch(ctx, probs, tau, tau, iters, graph=True, autotune=autotune, persistent=True)
    wall_p = time.perf_counter() - t
    tiles_p, occ_p = S.potts_batch_tiles(ctx), S.potts_batch_occupancy(ctx)
    same_p = 0
    for (g, _, _), a_, b_ in zip(probs, outs, outs_p):
        m = g.mapping.valid_mask()
        same_p += int(np.array_equal(a_.to_host()[m], b_.to_host()[m]))
        b_.close()
    log(f"persistent, tiles {tiles_p}, workgroups per CU {occ_p}: iterations {ms_p:.1f} ms (events), "
        f"call {wall_p:.2f} s; labels equal to the graph's {same_p}/{len(probs)}")
    t = time.perf_counter()
    refs = [S.potts_min_partition(ctx, g, c, lam, tau, tau, iters) for (g, c, lam) in probs]
    ctx.synchronize()
    wall_loop = time.perf_counter() - t
    agree = 0
    for (g, _, _), o, ref in zip(probs, outs, refs):
        m = g.mapping.valid_mask()
        agree += int(np.array_equal(ref.to_host()[m], o.to_host()[m]))
        ref.close()
        o.close()
    for g in dev:
        g.close()
    log(f"looped single solves: {wall_loop:.2f} s; identical labels {agree}/{len(probs)}")
    return {"workload": "64 sub-volumes (edges U[48,192], box or cylinder, K~U{3,4,5}) x 9 lambdas, "
                        f"{iters} iterations, FP32",
            "solves": len(probs), "voxels": vox, "iterations": iters,
            "voxel_iters_per_s": vox * iters / (ms / 1e3), "iterations_ms": ms,
            "iteration_ms": ms / iters, "iterations_ms_default_tile": ms_default,
            "xi_tile_waves": {str(k): v for k, v in tiles.items()},
            "batched_call_s": wall_batch, "looped_calls_s": wall_loop,
            "speedup_vs_looped": wall_loop / wall_batch,
            "identical_labels": agree,
            "persistent": {"iterations_ms": ms_p, "voxel_iters_per_s": vox * iters / (ms_p / 1e3),
                           "call_s": wall_p, "speedup_vs_looped": wall_loop / wall_p,
                           "xi_tile_waves": {str(k): v for k, v in tiles_p.items()},
                           "workgroups_per_cu": {str(k): v for k, v in occ_p.items()},
                           "labels_equal_to_graph": same_p},
            "shapes": [[v.n, v.height, v.width, "box" if v.box else "cylinder", int(v.c.size)] for v in vols]}


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
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-storage comparison run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing run")
    ap.add_argument("--bf16-steps", type=int, default=30)
    ap.add_argument("--no-tv", action="store_true", help="skip the Chambolle TV + histogram workload")
    ap.add_argument("--no-long", action="store_true", help="skip the 1000-iteration solve with the FP64 gap")
    ap.add_argument("--long-iters", type=int, default=1000)
    ap.add_argument("--gap-every", type=int, default=50)
    ap.add_argument("--no-agreement", action="store_true",
                    help="skip solving a 256^3 central box for --long-iters iterations in FP32 / BF16 / FP16 "
                         "and comparing with the oracle's FP64 solve (minutes of host time, overlapped with "
                         "the GPU legs)")
    ap.add_argument("--tv-iters", type=int, default=200)
    ap.add_argument("--tv-only", action="store_true",
                    help="run only the Chambolle TV + histogram workload (512^3) and print its JSON")
    ap.add_argument("--no-reductions", action="store_true", help="skip the reductions workload")
    ap.add_argument("--reductions-only", action="store_true",
                    help="run only the reductions on the 512^3 phantom and print their JSON")
    ap.add_argument("--batch", action="store_true",
                    help="run only the batched workload (576 solves) and print its JSON")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--batch-iters", type=int, default=300)
    ap.add_argument("--no-autotune", action="store_true", help="batched workload without the tile autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB mebibytes (hipMemcpy's copy "
                         "kernel), to calibrate the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.batch:
        import segmentctscans_amd as S
        r = batch_workload(S, args.seed, args.batch_iters, not args.no_autotune)
        print(json.dumps({"metric": "batched Potts voxel-iters/s", "value": r["voxel_iters_per_s"],
                          "unit": "voxel-iters/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
                          "data": "synthetic", "config": {"workload": "potts_batch_576"}, **r}), flush=True)
        return 0
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.reductions_only:
        import segmentctscans_amd as S
        r = reductions_workload(S)
        log(", ".join(f"{k} {v['ms']:.3f} ms ({v['frac']:.2f})" for k, v in r.items() if isinstance(v, dict)))
        print(json.dumps(r), flush=True)
        return 0
    if args.tv_only:
        import segmentctscans_amd as S
        r = tv_workload(S, args.tv_iters)
        log(f"TV: {r['ms_per_iteration']:.3f} ms/iteration, {r['frac']:.3f} of HBM peak")
        print(json.dumps(r, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
        return 0

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
    # the FP64 oracle's solve of the agreement box runs on host threads while the GPU legs run
    fp64 = None if args.no_agreement else Fp64Box(ph, c, lam, tau, args.long_iters, 256, host_cores()["threads"])
    log(f"phantom done: {vox} in-mask voxels, {ph.n_spheres} spheres, porosity {ph.porosity:.3f}")

    ctx = S.make_context(0)
    g = S.upload(ctx, ph.field)
    solver = S.PottsSolver(ctx, g, c, lam, tau, tau, storage=args.storage, kernel=args.kernel)
    log("uploaded; warmup")
    solver.iterate(args.warmup)
    ctx.synchronize()

    tl = EventTimeline(ctx.stream, 2)
    ctx.synchronize()
    w0 = time.perf_counter()
    tl.mark()
    kms = solver.iterate_timed(args.steps)   # events between the kernel launches
    tl.mark()
    ctx.synchronize()
    wall = time.perf_counter() - w0
    total_ms = float(tl.durations()[0])
    tl.close()
    ms_step = total_ms / args.steps
    value = vox * args.steps / (total_ms / 1e3)
    names = KERNEL_NAMES[args.kernel]
    assert len(names) == kms.shape[1]
    # fused: a call of n iterations is the two-kernel Xi step, n - 1 fused
    # launches and the two-kernel V step; the last entry times that V step
    kl = kms[:-1] if args.kernel == "fused" and args.steps > 1 else kms
    kavg = [float(x) for x in kl.astype(np.float64).mean(axis=0)]
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step (events), wall {wall:.3f} s; per kernel "
        + ", ".join(f"{n} {a:.3f} ms" for n, a in zip(names, kavg)))

    kb = kernel_bytes(K, args.storage)
    dom = int(np.argmax(kavg))
    dom_name = names[dom]
    achieved = kb[dom_name] * vox / (kavg[dom] / 1e3) / 1e9
    step_bytes = sum(kb[n] for n in names)
    step_gbs = step_bytes * vox / (ms_step / 1e3) / 1e9
    workload = "potts_k3_fp32" if args.storage == "fp32" else f"potts_k3_{args.storage}_storage"
    traffic, traffic_src = committed_traffic(dom_name, workload) if (
        args.n, args.side, args.radius, args.rmax, args.seed) == (1426, 1024, 460.0, 40.0, 3) else (None, None)
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
            "workload": workload,
            "volume_n_rows_cols": [args.n, args.side, args.side],
            "cylinder_radius": args.radius,
            "in_mask_voxels": vox,
            "K": K, "lambda": lam, "tau": tau,
            "storage": args.storage,
            "kernel": args.kernel,
            "phantom_seed": args.seed,
        },
        "kernel_ms": {n: float(a) for n, a in zip(names, kavg)},
        "hbm_peak_frac": step_gbs / HBM_PEAK_GBS,
        "step_bytes_per_voxel_iter": step_bytes,
        # the same rate priced at the single-pass compulsory bytes (RW Xi, RW V,
        # RW V-bar, R u): what a one-pass iteration would have to move
        "single_pass_bytes_per_voxel_iter": single_pass_bytes(K, args.storage),
        "hbm_peak_frac_single_pass": single_pass_bytes(K, args.storage) * vox / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": kb[dom_name] * vox,
            "bytes_per_voxel": kb[dom_name],
        },
    }
    solver.close()
    if args.storage == "fp32" and not args.no_bf16:
        sb = S.PottsSolver(ctx, g, c, lam, tau, tau, storage="bf16", kernel=args.kernel)
        sb.iterate(2)
        ctx.synchronize()
        tb = EventTimeline(ctx.stream, 2)
        tb.mark()
        sb.iterate(args.bf16_steps)
        tb.mark()
        ctx.synchronize()
        ms_bf16 = float(tb.durations()[0]) / args.bf16_steps
        tb.close()
        sb.close()
        result["bf16_storage"] = {"ms_per_step": ms_bf16, "speedup_vs_fp32": ms_step / ms_bf16,
                                  "steps": args.bf16_steps}
        log(f"BF16 storage: {ms_bf16:.3f} ms/step, speedup {ms_step / ms_bf16:.3f}")
    if not args.no_long:
        for st in ("fp32", "bf16", "fp16"):
            lr = long_run(S, ctx, g, c, lam, tau, args.long_iters, args.gap_every, st)
            result["long_run" if st == "fp32" else f"long_run_{st}"] = lr
            log(f"{st}: {lr['iterations']} iterations with the FP64 gap every {lr['gap_every']}: "
                f"{lr['ms_total'] / 1e3:.2f} s, relative gap {lr['relative_gap_trace'][0]:.3e} -> "
                f"{lr['relative_gap_trace'][-1]:.3e}")
    g.close()
    box_states = agreement_gpu(S, fp64, c, lam, tau) if fp64 is not None else None
    if not args.no_adjoint:
        result["adjoint"] = adjoint_ratio(S)
        log(f"adjoint: forward {result['adjoint']['forward_ms']:.2f} ms, reverse "
            f"{result['adjoint']['reverse_ms']:.2f} ms, ratio {result['adjoint']['ratio']:.2f}; every state "
            f"stored: ratio {result['adjoint']['every_state_stored']['ratio']:.2f}")
    tv_keep = {}
    if not args.no_tv:
        result["tv"] = tv_workload(S, args.tv_iters, tv_keep)
        log(f"TV: {result['tv']['ms_per_iteration']:.3f} ms/iteration, "
            f"{result['tv']['voxel_iters_per_s']:.3e} voxel-iters/s, {result['tv']['frac']:.3f} of HBM peak; "
            f"65536-bin histogram {np.median(result['tv']['hist_ms']):.3f} ms")
    if not args.no_reductions:
        result["reductions"] = reductions_workload(S)
        log("reductions: " + ", ".join(f"{k} {v['ms']:.3f} ms ({v['frac']:.2f} of peak)"
                                       for k, v in result["reductions"].items() if isinstance(v, dict)))
    if not args.no_batch:
        result["batch"] = batch_workload(S, args.seed, args.batch_iters, not args.no_autotune)
    if fp64 is not None:
        result["agreement"] = crop_agreement(fp64, box_states)   # waits for the FP64 solve
        box_states = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        if "ph" not in tv_keep:
            tv_keep["ph"] = make_phantom(512, 512, 4.0, 20.0, seed=7, r=230.4)
        result["cpu_baseline"] = cpu_baseline(ph, c, tau, args.cpu_iters, args.cpu_size, tv_keep["ph"],
                                              tv_keep.get("out"), tv_keep.get("range"),
                                              fp64.timing() if fp64 is not None else None)
        log(f"CPU baseline {result['cpu_baseline']['value']:.3e} voxel-iters/s on "
            f"{result['cpu_baseline']['cores']} threads ({result['cpu_baseline']['cpu_model']})")
    print(json.dumps(result, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
