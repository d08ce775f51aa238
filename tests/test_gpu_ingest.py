"""Ingest on the GPU (transpose/): convert_to_fp (u16 and float input) and
find_intensity_range (65536-bin histogram and tail cut) equal the oracle bit
for bit on scanner-like u16 volumes, over line layouts that take every path
of the line walks: full 4-voxel vectors, an odd tail, padding after the
line, and a stride that is not a multiple of 4."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd import Field

pytestmark = pytest.mark.gpu


# (n_images, nstride): a stride equal to the line length (lines cover the
# span), the default 32-rounded stride (padding after an odd tail of 97 =
# 24 * 4 + 1 voxels), and a stride that is not a multiple of 4 (one voxel per
# lane, padding of 2)
LAYOUTS = [(96, None), (97, None), (97, 99)]


@pytest.fixture(scope="module", params=LAYOUTS, ids=lambda p: f"n{p[0]}_stride{p[1] or 'auto'}")
def raw(request):
    n, nstride = request.param
    m = S.Mapping.cylinder(n, 90, 88, 43.5, 44.5, 40.0, nstride)
    valid = m.valid_mask()
    rng = np.random.default_rng(n + (nstride or 0))
    base = np.clip(rng.normal(27000, 6000, valid.sum()), 0, 65535)
    u16 = np.zeros(m.span, np.uint16)
    u16[valid] = np.round(base).astype(np.uint16)
    # values at the ends of the range and on the partition edge
    u16[np.flatnonzero(valid)[:4]] = [0, 65535, 32767, 32768]
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
    if (~valid).any():
        assert np.all(out[~valid] == 0)
    g.close()


def test_layouts_take_every_path(raw):
    """The fixture's layouts cover what the kernels branch on."""
    m, _ = raw
    assert m.n_images % 4 in (0, 1)
    assert (m.nstride == m.n_images) or (m.nstride > m.n_images)
