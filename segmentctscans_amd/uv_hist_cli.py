"""The uv-hist driver (uv-hist/uv-hist.cpp:28-127) over the C ABI.

    python -m segmentctscans_amd.uv_hist_cli <in.bin> -o out.png

Reads a Field<float> archive, computes the gradient magnitude v of u
(grad_magn), the 512 x 512 joint histogram of (u, v) over u's range and the
range of v > 0 (uv_hist), prints the ranges, and writes the histogram as an
8-bit RGB PNG: log(count + 1) scaled so the largest is 255 (rounded half to
even), v increasing upwards, coloured with the viridis colour map (:107-127).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import archive, grad_magn, make_context, upload, uv_hist
from .potts_model import png_bytes_rgb

NBINS = 512


def viridis_lut() -> np.ndarray:
    """The 256-entry viridis map as uint8 RGB, each channel round(255 v)."""
    from matplotlib import colormaps
    lut = colormaps["viridis"](np.arange(256))[:, :3]
    return np.rint(lut * 255.0).astype(np.uint8)


def histogram_image(hist: np.ndarray, nbins: int = NBINS) -> np.ndarray:
    """uv-hist.cpp:107-127: counts (rows = v bins, columns = u bins) to the
    colour-mapped image, row order flipped so v grows upwards."""
    counts = np.log(hist[:, :nbins].astype(np.float64) + 1.0)[::-1]
    mx = counts.max() if counts.size else 0.0
    scale = 255.0 / mx if mx > 0.0 else 0.0
    idx = np.clip(np.rint(counts * scale), 0, 255).astype(np.uint8)
    return viridis_lut()[idx]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input")
    ap.add_argument("-o", "--output", required=True)
    args = ap.parse_args(argv)
    if not os.path.exists(args.input):
        print(f"Input file {args.input} doesn't exist", file=sys.stderr)
        return 1
    t = time.perf_counter()
    img = archive.load_bin(args.input, np.float32)
    print(f"Read time: {time.perf_counter() - t:.3f} s")
    ctx = make_context(0)
    t = time.perf_counter()
    d = upload(ctx, img)
    print(f"Upload to GPU time: {time.perf_counter() - t:.3f} s")
    t = time.perf_counter()
    v = grad_magn(ctx, d)
    print(f"Gradient magnitude time: {time.perf_counter() - t:.3f} s")
    t = time.perf_counter()
    r = uv_hist(ctx, d, v, NBINS)
    print(f"Histogram time: {time.perf_counter() - t:.3f} s")
    print(f"u ∈ [{r.u_start:g}, {r.u_end:g}], v ∈ [{r.v_start:g}, {r.v_end:g}]")
    with open(args.output, "wb") as f:
        f.write(png_bytes_rgb(histogram_image(r.hist)))
    v.close()
    d.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
