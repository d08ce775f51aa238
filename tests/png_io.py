"""A minimal PNG reader for the tests: checks the signature and every chunk's
CRC, reads IHDR, inflates the IDAT stream and undoes filter type 0 (the only
filter the writers use).  Returns (header fields, pixel array)."""
import struct
import zlib

import numpy as np


def read_png(buf: bytes):
    assert buf[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(buf):
        (n,) = struct.unpack(">I", buf[pos:pos + 4])
        tag, data = buf[pos + 4:pos + 8], buf[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", buf[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(tag + data) & 0xFFFFFFFF
        if tag == b"IHDR":
            hdr = dict(zip(("w", "h", "depth", "color", "comp", "filter", "interlace"), struct.unpack(">IIBBBBB", data)))
        elif tag == b"IDAT":
            idat += data
        pos += 12 + n
    ch = {0: 1, 2: 3}[hdr["color"]]
    raw = zlib.decompress(idat)
    row = 1 + hdr["w"] * ch
    assert len(raw) == hdr["h"] * row
    rows = np.frombuffer(raw, np.uint8).reshape(hdr["h"], row)
    assert np.all(rows[:, 0] == 0)
    px = rows[:, 1:].reshape(hdr["h"], hdr["w"], ch)
    return hdr, (px[..., 0] if ch == 1 else px)
