"""The reference's .bin volume archive (Field<T>::save / load, field.h:92-153)
through the C ABI: load_bin(path_or_bytes, dtype) -> Field, save_bin(field,
path) / to_bytes(field)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import Field, Mapping, SctError
from . import _native as N


def _dt(dtype) -> int:
    return N.SCT_U16 if np.dtype(dtype) == np.uint16 else N.SCT_F32


def parse(buf: bytes | np.ndarray, dtype=np.float32):
    """Header of an archive: (Mapping, lo, hi, info)."""
    arr = np.frombuffer(buf, np.uint8) if isinstance(buf, (bytes, bytearray, memoryview)) else buf
    info = N.SctBinInfo()
    rc = N.lib().sct_bin_parse(arr.ctypes.data, arr.size, _dt(dtype), C.byref(info), None, None)
    if rc:
        raise SctError(rc, "not a readable .bin archive")
    sr = np.zeros(info.width, np.uint32)
    base = np.zeros(info.width + 1, np.uint32)
    rc = N.lib().sct_bin_parse(arr.ctypes.data, arr.size, _dt(dtype), C.byref(info), sr.ctypes.data_as(N.u32p),
                               base.ctypes.data_as(N.u32p))
    if rc:
        raise SctError(rc, "not a readable .bin archive")
    m = Mapping(info.n_images, info.height, info.width, info.nstride, info.col_begin, info.col_end, sr, base,
                int(info.data_size))
    return m, info.lo, info.hi, info


def load_bin(src, dtype=np.float32) -> Field:
    """Field<T> from an archive file (path) or bytes."""
    if isinstance(src, (str, os.PathLike)):
        arr = np.fromfile(src, np.uint8)
    else:
        arr = np.frombuffer(src, np.uint8)
    m, lo, hi, info = parse(arr, dtype)
    data = np.zeros(m.span, np.dtype(dtype))
    rc = N.lib().sct_bin_read(arr.ctypes.data, arr.size, _dt(dtype), C.byref(info), data.ctypes.data)
    if rc:
        raise SctError(rc, "archive truncated")
    return Field(m, data, lo=lo, hi=hi)


def to_bytes(field: Field) -> bytes:
    data = np.ascontiguousarray(field.data)
    lay = field.mapping.c_layout()
    cap = N.lib().sct_bin_bound(C.byref(lay), _dt(data.dtype))
    buf = np.zeros(cap, np.uint8)
    n = C.c_size_t()
    rc = N.lib().sct_bin_write(C.byref(lay), _dt(data.dtype), float(field.lo), float(field.hi), data.ctypes.data,
                               buf.ctypes.data, cap, C.byref(n))
    if rc:
        raise SctError(rc, "archive write failed")
    return buf[: n.value].tobytes()


def save_bin(field: Field, path) -> None:
    with open(path, "wb") as f:
        f.write(to_bytes(field))
