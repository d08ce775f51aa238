"""The solver's device rounding to 16-bit storage (the stored Xi and V-bar
elements of SCT_STORAGE_BF16 / SCT_STORAGE_F16) equals the oracle's rounding
bit for bit on every class of input: for every 16-bit upper half (sign,
exponent and the bfloat16 mantissa, so every NaN, infinity, subnormal and
normal class) combined with the lower halves that decide round to nearest
even (0, 1, just below, at and above the tie, all ones), and on random bit
patterns."""
import numpy as np
import pytest

import oracle as O
from segmentctscans_amd import _native as N

pytestmark = pytest.mark.gpu


def _inputs():
    hi = np.arange(1 << 16, dtype=np.uint32) << 16
    lows = np.array([0x0000, 0x0001, 0x7FFF, 0x8000, 0x8001, 0xFFFF], np.uint32)
    grid = (hi[:, None] | lows[None, :]).reshape(-1)
    rng = np.random.default_rng(11)
    rand = rng.integers(0, 2**32, 1 << 18, dtype=np.uint64).astype(np.uint32)
    return np.concatenate([grid, rand]).view(np.float32)


def _device(ctx, storage, x):
    out = np.zeros(x.size, np.uint16)
    ctx.check(N.lib().sct_selftest_store_round(ctx.handle, N.STORAGE[storage], x.ctypes.data, x.size,
                                               out.ctypes.data))
    return out


def test_bf16_store_round_equals_oracle(ctx):
    x = _inputs()
    got = _device(ctx, "bf16", x).astype(np.uint32) << 16
    want = O.store_round(x, "bf16").view(np.uint32)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(hex(int(x.view(np.uint32)[i])), hex(int(got[i])), hex(int(want[i]))) for i in bad[:8]]


def test_f16_store_round_equals_oracle(ctx):
    x = _inputs()
    got = _device(ctx, "fp16", x).view(np.float16).astype(np.float32).view(np.uint32)
    want = O.store_round(x, "fp16").view(np.uint32)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(hex(int(x.view(np.uint32)[i])), hex(int(got[i])), hex(int(want[i]))) for i in bad[:8]]
