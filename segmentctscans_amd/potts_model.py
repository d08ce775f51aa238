"""The potts-model driver (potts-model/potts-model.cpp:80-244) over the C ABI.

    python -m segmentctscans_amd.potts_model <in.bin> --labels "{a,b,c}" --lambda 30 [--rawout out.vti]

Reads a Field<float> archive, normalises the labels with the field's lo / hi
(:141-149), runs potts_min_partition with tau = 0.99 sqrt(1/12) for 100
iterations (:141, :162-170), writes the PNG slice of column 400 before and
after (:34-49, :137, :182), and with --rawout the labels as VTK ImageData with
raw appended UInt8 data (:184-241).
"""
from __future__ import annotations

import argparse
import re
import struct
import sys
import time
import zlib

import numpy as np

from . import Field, archive, make_context, max_phases, potts_min_partition, upload


def parse_labels(s: str) -> list[float]:
    """'{a, b, c}' with 2..max_phases floats (the boost.parser rule of :56-76)."""
    m = re.fullmatch(r"\s*\{\s*(.*?)\s*\}\s*", s)
    if not m:
        raise ValueError("labels must look like {a, b, c}")
    vals = [float(x) for x in m.group(1).split(",")]
    if not 2 <= len(vals) <= max_phases:
        raise ValueError(f"between 2 and {max_phases} labels")
    return vals


def png_bytes(img: np.ndarray) -> bytes:
    """8-bit greyscale PNG (filter type 0 on every row, zlib stream)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def png_bytes_rgb(img: np.ndarray) -> bytes:
    """8-bit RGB PNG of an (h, w, 3) array (filter type 0 on every row)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w, _ = img.shape

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def slice_image(field: Field, col: int) -> np.ndarray:
    """write_png (:34-49): rows x n of column col, v^(1/2.2) * 255 rounded to
    nearest (ties to even) and saturated to [0, 255].  A picture for the eye:
    the pow of the reference's image library may differ in the last bit."""
    m = field.mapping
    r0, r1 = m.row_begin(col), m.row_end(col)
    line0 = int(m.base[col]) // m.nstride
    sl = field.data[line0 * m.nstride:(line0 + r1 - r0) * m.nstride].reshape(r1 - r0, m.nstride)[:, :m.n_images]
    g = np.power(sl.astype(np.float32), np.float32(1.0 / 2.2)) * np.float32(255.0)
    return np.clip(np.rint(g), 0, 255).astype(np.uint8)


def write_png(field: Field, col: int, lam: float, directory: str = ".") -> str | None:
    m = field.mapping
    if not (m.col_begin <= col < m.col_end) or m.row_end(col) <= m.row_begin(col):
        return None
    name = f"{directory}/slice-{col} (\U0001d706 = {lam:07.3f}).png"
    with open(name, "wb") as f:
        f.write(png_bytes(slice_image(field, col)))
    return name


def vtk_bytes(field: Field) -> bytes:
    """The --rawout file (:184-241): in-mask voxels as clamp(round(255 v)),
    other voxels of the (rows x cols) bounding box 255, n fastest."""
    m = field.mapping
    rb = min(m.row_begin(c) for c in range(m.col_begin, m.col_end))
    re_ = max(m.row_end(c) for c in range(m.col_begin, m.col_end))
    rows, cols, n = re_ - rb, m.col_end - m.col_begin, m.n_images
    out = np.full((cols, rows, n), 255, np.uint8)
    for c in range(m.col_begin, m.col_end):
        r0, r1 = m.row_begin(c), m.row_end(c)
        line0 = int(m.base[c]) // m.nstride
        v = field.data[line0 * m.nstride:(line0 + r1 - r0) * m.nstride].reshape(r1 - r0, m.nstride)[:, :n]
        # 255 * v in float, std::round (half away from zero; exact in double), clamp to [0, 255]
        x = (np.float32(255.0) * v.astype(np.float32)).astype(np.float64)
        q = np.sign(x) * np.floor(np.abs(x) + 0.5)
        out[c - m.col_begin, r0 - rb:r1 - rb, :] = np.clip(q, 0, 255).astype(np.uint8)
    head = ("<?xml version=\"1.0\"?>\n"
            "<VTKFile type=\"ImageData\" byte_order=\"LittleEndian\">\n"
            f"  <ImageData WholeExtent=\"0 {n - 1} 0 {rows - 1} 0 {cols - 1}\">\n"
            f"    <Piece Extent=\"0 {n - 1} 0 {rows - 1} 0 {cols - 1}\">\n"
            "      <PointData Scalars=\"labels\">\n"
            "        <DataArray type=\"UInt8\" Name=\"labels\" format=\"appended\" offset=\"0\"/>\n"
            "      </PointData>\n"
            "    </Piece>\n"
            "  </ImageData>\n"
            "  <AppendedData encoding=\"raw\">\n_")
    data = out.tobytes()
    return head.encode() + struct.pack("<I", len(data)) + data + b"\n  </AppendedData>\n</VTKFile>\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input")
    ap.add_argument("--rawout")
    ap.add_argument("--labels", required=True)
    ap.add_argument("--lambda", dest="lam", type=float, required=True)
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--png-dir", default=".")
    args = ap.parse_args(argv)
    try:
        labels = parse_labels(args.labels)
    except ValueError as x:
        print(f"Command line error: {x}", file=sys.stderr)
        return 1
    t = time.perf_counter()
    img = archive.load_bin(args.input, np.float32)
    print(f"Read time: {time.perf_counter() - t:.3f} s")
    write_png(img, 400, 0.0, args.png_dir)
    lo, hi = np.float32(img.lo), np.float32(img.hi)
    ci = np.array([(np.float32(x) - lo) / (hi - lo) for x in labels], np.float32)
    print("labels: " + ", ".join(f"{round(255 * float(c)):g}" for c in ci))
    ctx = make_context(0)
    g = upload(ctx, img)
    tau = float(np.float32(0.99) * np.float32(np.sqrt(np.float32(1.0 / 12.0))))
    t = time.perf_counter()
    res = potts_min_partition(ctx, g, ci, args.lam, tau, tau, args.iters)
    if res is None:
        print("Not enough memory. Reduce the mask radius or the number of iterations.", file=sys.stderr)
        return 1
    out = Field(img.mapping, res.to_host(), lo=img.lo, hi=img.hi)
    print(f"Potts time: {time.perf_counter() - t:.3f} s")
    write_png(out, 400, args.lam, args.png_dir)
    if args.rawout:
        with open(args.rawout, "wb") as f:
            f.write(vtk_bytes(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
