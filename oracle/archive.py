"""Bit-level restatement of the reference's .bin archive (test infrastructure).

Follows io/outbitstream.h:10-55 (MSB-first bit stream, last byte zero-padded),
io/writeprimitives.h:55-66 / io/readprimitives.h:53-75 (order-0 Exp-Golomb ue,
signed se with k > 0 -> 2k - 1, k <= 0 -> -2k), layout_cylinder.h:233-318
(mapping header) and field.h:92-153 (lo / hi for float fields, voxels in
(col, row, n) order).  Pure Python, one bit at a time: for small fields only.
"""
from __future__ import annotations

import numpy as np


class BitWriter:
    def __init__(self):
        self.bits: list[int] = []

    def put(self, v: int, n: int) -> None:
        for i in range(n - 1, -1, -1):
            self.bits.append((v >> i) & 1)

    def ue(self, v: int) -> None:
        nb = (v + 1).bit_length()
        self.put(0, nb - 1)
        self.put(v + 1, nb)

    def se(self, v: int) -> None:
        a = 2 * abs(v)
        if v > 0:
            a -= 1
        self.ue(a)

    def tobytes(self) -> bytes:
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[i:i + 8])), 2) for i in range(0, len(b), 8))


class BitReader:
    def __init__(self, data: bytes):
        self.data, self.bit = data, 0

    def get(self, n: int) -> int:
        v = 0
        for _ in range(n):
            byte = self.data[self.bit >> 3]
            v = (v << 1) | ((byte >> (7 - (self.bit & 7))) & 1)
            self.bit += 1
        return v

    def ue(self) -> int:
        z = 0
        while self.get(1) == 0:
            z += 1
        return (1 << z) + self.get(z) - 1 if z else 0

    def se(self) -> int:
        u = self.ue()
        return (u + 1) >> 1 if u & 1 else -(u >> 1)


def save(m, data: np.ndarray, lo: float = 0.0, hi: float = 0.0) -> bytes:
    """Field<T>::save for T = float32 or uint16 (m: a cylinder mapping with
    n_images, height, width, nstride, col_begin, col_end, start_row, base)."""
    w = BitWriter()
    for v in (m.n_images, m.height, m.width, m.nstride, m.col_begin, m.col_end):
        w.ue(int(v))
    sr, base = [int(x) for x in m.start_row], [int(x) for x in m.base]
    w.ue(len(sr))
    if sr:
        w.ue(sr[0])
        for a, b in zip(sr, sr[1:]):
            w.se(b - a)
    w.ue(len(base))
    if base:
        w.ue(base[0])
        for a, b in zip(base, base[1:]):
            w.se(b - a)
    nb = 16 if data.dtype == np.uint16 else 32
    if nb == 32:
        w.put(int(np.float32(lo).view(np.uint32)), 32)
        w.put(int(np.float32(hi).view(np.uint32)), 32)
    raw = data.view(np.uint16 if nb == 16 else np.uint32)
    for col in range(m.col_begin, m.col_end):
        rows = (base[col + 1] - base[col]) // m.nstride
        for r in range(rows):
            o = base[col] + r * m.nstride
            for n in range(m.n_images):
                w.put(int(raw[o + n]), nb)
    return w.tobytes()


def load(buf: bytes, dtype) -> dict:
    """Field<T>::load: dict(n_images, height, width, nstride, col_begin,
    col_end, start_row, base, lo, hi, data)."""
    r = BitReader(buf)
    n, h, wd, ns, cb, ce = (r.ue() for _ in range(6))
    sr = []
    k = r.ue()
    if k:
        sr.append(r.ue())
        for _ in range(k - 1):
            sr.append((sr[-1] + r.se()) & 0xFFFFFFFF)
    base = []
    k = r.ue()
    if k:
        base.append(r.ue())
        for _ in range(k - 1):
            base.append((base[-1] + r.se()) & 0xFFFFFFFF)
    out = dict(n_images=n, height=h, width=wd, nstride=ns, col_begin=cb, col_end=ce,
               start_row=np.array(sr, np.uint32), base=np.array(base, np.uint32), lo=0.0, hi=0.0)
    nb = 16 if np.dtype(dtype) == np.uint16 else 32
    if nb == 32:
        out["lo"] = float(np.uint32(r.get(32)).view(np.float32))
        out["hi"] = float(np.uint32(r.get(32)).view(np.float32))
    raw = np.zeros(base[-1] if base else 0, np.uint16 if nb == 16 else np.uint32)
    for col in range(cb, ce):
        for rr in range((base[col + 1] - base[col]) // ns):
            o = base[col] + rr * ns
            for i in range(n):
                raw[o + i] = r.get(nb)
    out["data"] = raw.view(np.dtype(dtype))
    return out
