"""Ingest on the GPU (transpose/): convert_to_fp (u16 and float input) and
find_intensity_range (65536-bin histogram and tail cut) equal the oracle bit
for bit on a scanner-like u16 volume."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd import Field
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def raw():
    ph = make_phantom(96, 96, 3.0, 10.0, seed=77)
    m = ph.field.mapping
    valid = m.valid_mask()
    rng = np.random.default_rng(1)
    u16 = np.zeros(m.span, np.uint16)
    u16[valid] = np.clip(np.round(ph.field.data[valid] * 30000 + 12000 + rng.normal(0, 300, valid.sum())), 0, 65535)
    return m, u16


@pytest.mark.parametrize("tail", [0.0, 0.001, 0.02])
def test_find_intensity_range(ctx, raw, tail):
    m, u16 = raw
    g = S.upload(ctx, Field(m, u16))
    lo, hi, h = S.find_intensity_range(ctx, g, tail, hist=True)
    lay = O.Layout.from_mapping(m)
    assert np.array_equal(h, O.u16_hist(lay, u16))
    assert (lo, hi) == O.find_intensity_range(lay, u16, tail)
    g.close()


@pytest.mark.parametrize("as_float", [False, True])
def test_convert_to_fp(ctx, raw, as_float):
    m, u16 = raw
    x = u16.astype(np.float32) if as_float else u16
    g = S.upload(ctx, Field(m, x))
    out = S.convert_to_fp(ctx, g, 14000.0, 40000.5).to_host()
    want = O.convert_to_fp(O.Layout.from_mapping(m), x, 14000.0, 40000.5)
    valid = m.valid_mask()
    assert np.array_equal(out[valid].view(np.uint32), want[valid].view(np.uint32))
    assert np.all(out[~valid] == 0)
    g.close()
