"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-side helpers agree with the oracle."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

import segmentctscans_amd as S
from segmentctscans_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    syms = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        syms |= set(re.findall(r"\b(sct_[a-z0-9_]+)\s*\(", text))
    return syms


def test_header_symbols_exported():
    syms = declared_symbols()
    assert len(syms) >= 30
    lib = C.CDLL(str(N.LIB_DIR / "libsct.so"))
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert declared_symbols() == set(N.SIGNATURES)


def test_status_codes_match_header():
    text = (ROOT / "include" / "sct.h").read_text()
    for name in ("SCT_OK", "SCT_ERR_INVALID", "SCT_ERR_HIP", "SCT_ERR_SIZE", "SCT_ERR_NOMEM"):
        val = int(re.search(rf"#define {name}\s+(\d+)", text).group(1))
        assert getattr(N, name) == val


def test_opts_struct_size():
    o = N.SctPottsOpts()
    N.lib().sct_potts_opts_init(C.byref(o))
    assert o.size == C.sizeof(N.SctPottsOpts)
    assert o.mem_available == -1 and o.slab_w == 0 and o.storage == 0


def test_invalid_layout_rejected_without_device():
    # sct_cylinder_layout rejects nstride 0 and null outputs (pure host code)
    lib = N.lib()
    sr = np.zeros(4, np.uint32)
    base = np.zeros(5, np.uint32)
    cb, ce, ds = C.c_uint32(), C.c_uint32(), C.c_uint64()
    rc = lib.sct_cylinder_layout(4, 4, 4, 1.5, 1.5, 2.0, 0, sr.ctypes.data_as(N.u32p),
                                 base.ctypes.data_as(N.u32p), C.byref(cb), C.byref(ce), C.byref(ds))
    assert rc == N.SCT_ERR_INVALID
    # 32-bit span overflow is refused rather than wrapped
    big_w = 70000
    sr = np.zeros(big_w, np.uint32)
    base = np.zeros(big_w + 1, np.uint32)
    rc = lib.sct_cylinder_layout(65536, big_w, big_w, big_w / 2, big_w / 2, big_w / 2, 65536,
                                 sr.ctypes.data_as(N.u32p), base.ctypes.data_as(N.u32p),
                                 C.byref(cb), C.byref(ce), C.byref(ds))
    assert rc == N.SCT_ERR_INVALID


def test_context_fails_loudly_without_gpu():
    import pytest
    try:
        ctx = S.make_context(0)
    except S.SctError as e:
        assert e.code in (N.SCT_ERR_HIP, N.SCT_ERR_INVALID)
        return
    ctx.close()
    pytest.skip("a GPU is visible")
