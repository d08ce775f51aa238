"""The .bin archive (Field<T>::save / load, field.h:92-153) through the C ABI,
host only.  The oracle's bit-level restatement (oracle/archive.py) is pinned
by the textbook Exp-Golomb codewords and a hand-assembled stream; the library
writer must produce the oracle's bytes and the reader must invert both."""
import ctypes as C

import numpy as np
import pytest

from oracle import archive as OA
from segmentctscans_amd import Field, Mapping, SctError
from segmentctscans_amd import _native as N
from segmentctscans_amd import archive as A
from segmentctscans_amd.phantom import make_phantom


def _code(fn, v):
    w = OA.BitWriter()
    getattr(w, fn)(v)
    return "".join(map(str, w.bits))


def test_exp_golomb_codewords():
    ue = ["1", "010", "011", "00100", "00101", "00110", "00111", "0001000", "0001001"]
    assert [_code("ue", v) for v in range(9)] == ue
    se = {0: "1", 1: "010", -1: "011", 2: "00100", -2: "00101", 3: "00110", -3: "00111"}
    assert {v: _code("se", v) for v in se} == se
    for v in list(range(300)) + [2**31, 2**32 + 5]:
        w = OA.BitWriter()
        w.ue(v)
        w.se(-v)
        r = OA.BitReader(w.tobytes())
        assert r.ue() == v and r.se() == -v


def test_hand_assembled_stream():
    """A 2 x 1 x 1 field (n = 2, one row, one column, nstride 32) of u16."""
    m = Mapping(2, 1, 1, 32, 0, 1, np.array([0], np.uint32), np.array([0, 32], np.uint32), 2)
    data = np.zeros(32, np.uint16)
    data[:2] = [0xABCD, 0x0001]
    bits = ("011" + "010" + "010"                            # ue(2) n, ue(1) height, ue(1) width
            + "00000100001" + "1" + "010"                      # ue(32) nstride, ue(0), ue(1) columns
            + "010" + "1"                                      # #start_row = 1, start_row[0] = 0
            + "011" + "1" + "0000001000000"                    # #base = 2, base[0] = 0, se(32) = ue(63)
            + format(0xABCD, "016b") + format(1, "016b"))
    bits += "0" * (-len(bits) % 8)
    want = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
    assert OA.save(m, data) == want
    assert A.to_bytes(Field(m, data)) == want
    f = A.load_bin(want, np.uint16)
    assert np.array_equal(f.data[:2], data[:2]) and f.mapping.nstride == 32


@pytest.mark.parametrize("dtype", [np.float32, np.uint16])
@pytest.mark.parametrize("shape", [(7, 9, 11, 5.0), (12, 10, 10, 40.0), (5, 1, 1, 3.0), (30, 16, 13, 6.3)])
def test_writer_matches_oracle_and_reader_inverts(dtype, shape):
    n, h, w, r = shape
    m = Mapping.cylinder(n, h, w, (w - 1) / 2, (h - 1) / 2, r)
    rng = np.random.default_rng(n + h)
    data = np.zeros(m.span, dtype)
    valid = m.valid_mask()
    if dtype == np.uint16:
        data[valid] = rng.integers(0, 65536, valid.sum())
    else:
        data[valid] = rng.standard_normal(valid.sum()).astype(np.float32)
    f = Field(m, data, lo=0.125, hi=1.75)
    b = A.to_bytes(f)
    assert b == OA.save(m, data, 0.125, 1.75)
    g = A.load_bin(b, dtype)
    assert np.array_equal(g.data.view(np.uint8), data.view(np.uint8))
    assert np.array_equal(g.mapping.start_row, m.start_row) and np.array_equal(g.mapping.base, m.base)
    assert g.mapping.data_size == m.data_size
    if dtype == np.float32:
        assert (g.lo, g.hi) == (0.125, 1.75)
    o = OA.load(b, dtype)
    assert np.array_equal(o["data"].view(np.uint8), data.view(np.uint8))


def test_large_round_trip(tmp_path):
    ph = make_phantom(256, 256, 4.0, 20.0, seed=11)
    p = tmp_path / "vol.bin"
    A.save_bin(ph.field, p)
    g = A.load_bin(str(p))
    assert np.array_equal(g.data, ph.field.data)
    assert (np.float32(g.lo), np.float32(g.hi)) == (np.float32(ph.field.lo), np.float32(ph.field.hi))


def test_truncated_and_invalid():
    m = Mapping.cylinder(8, 8, 8, 3.5, 3.5, 3.6)
    b = A.to_bytes(Field(m, np.ones(m.span, np.float32)))
    with pytest.raises(SctError):
        A.load_bin(b[:-3])
    with pytest.raises(SctError):
        A.load_bin(b"\x00\x00\x00\x00")


def test_header_width_beyond_the_stream_is_rejected():
    """A few header bytes can claim a width of about 2^32 columns; the parser
    rejects it from the stream length before sizing any table by it."""
    w = OA.BitWriter()
    for v in (10, 10, 0xFFFFFFF0, 16, 0, 0xFFFFFFF0, 0xFFFFFFF0):   # extents, nstride, col range, |start_row|
        w.ue(v)
    buf = np.frombuffer(w.tobytes() + b"\xff" * 16, np.uint8).copy()
    rc = N.lib().sct_bin_parse(buf.ctypes.data, buf.size, N.SCT_F32, C.byref(N.SctBinInfo()), None, None)
    assert rc == N.SCT_ERR_SIZE
    with pytest.raises(SctError):
        A.load_bin(buf.tobytes())


def test_cpp_save_load_bin_matches_oracle_bytes(tmp_path):
    """save_bin / load_bin of the C++ header (tests/cpp/archive_example.cpp):
    the files it writes are the oracle's bytes for the same fields."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "cpp" / "_build" / "archive_example"
    assert exe.exists(), "build first (__graft_entry__.build())"
    p = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "archive_example ok" in p.stdout, p.stderr
    for dtype, name, lo, hi in ((np.float32, "f32.bin", -0.25, 1.5), (np.uint16, "u16.bin", 0.0, 0.0)):
        f = Field.cylinder(37, 29, 23, 14.0, 11.0, 11.2)
        m = f.mapping
        data = np.zeros(m.span, dtype)
        for col in range(m.col_begin, m.col_end):
            for row in range(m.row_begin(col), m.row_end(col)):
                o = int(m.base[col]) + (row - m.row_begin(col)) * m.nstride
                n = np.arange(m.n_images)
                if dtype == np.float32:
                    data[o:o + m.n_images] = (n + 3 * row + 7 * col).astype(np.float32) / np.float32(1000.0)
                else:
                    data[o:o + m.n_images] = (n * 31 + row * 17 + col * 13) & 0xFFFF
        assert (tmp_path / name).read_bytes() == OA.save(m, data, lo, hi)
