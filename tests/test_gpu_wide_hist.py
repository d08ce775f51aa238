"""The wide histograms on the GPU (16-bit LDS counters, 65536 bins per
partition, spill lists) through the C ABI: calc_hist, uv_hist and
find_intensity_range equal the oracle, or plain counting, bit for bit on
volumes whose hot bins collect far more than 32768 voxels per slice of lines
-- so counters spill many times, in both halves of a counter word at once --
and over bin counts that leave a half word, a short last partition and a
number of slices that is not a multiple of the 8 XCDs."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd import Field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hot():
    """256 x 256 lines (the radius covers the square) of 300 voxels, 19.7 M
    voxels: 60% at 0.25, 30% at 0.75 and 10% uniform on [0, 1), with 0.0 and
    1.0 present so that [0, 1] is the value range."""
    m = S.Mapping.cylinder(300, 256, 256, 127.5, 127.5, 400.0)
    valid = m.valid_mask()
    nv = int(valid.sum())
    assert nv == 256 * 256 * 300
    rng = np.random.default_rng(11)
    r = rng.random(nv)
    x = np.where(r < 0.6, np.float32(0.25), np.where(r < 0.9, np.float32(0.75), rng.random(nv).astype(np.float32)))
    x[:2] = [0.0, 1.0]
    data = np.zeros(m.span, np.float32)
    data[valid] = x
    return m, data


@pytest.mark.parametrize("nbins", [8193, 65535, 65536, 65537, 131072, 300001])
def test_calc_hist_hot_bins(ctx, hot, nbins):
    m, data = hot
    lay = O.Layout.from_mapping(m)
    g = S.upload(ctx, Field(m, data))
    cnt = S.calc_hist(ctx, g, 0.0, 1.0, nbins, 0.0, counts=True)
    want = O.calc_hist(lay, data, 0.0, 1.0, nbins)
    assert np.array_equal(cnt, want)
    assert want.max() > 0.5 * m.data_size   # the 0.25 bin: spills in every slice
    g.close()


def test_calc_hist_adjacent_hot_bins_share_a_word(ctx, hot):
    """Two hot bins 2k and 2k + 1 (one 32-bit LDS word) spill concurrently."""
    m, _ = hot
    valid = m.valid_mask()
    nv = int(valid.sum())
    nbins = 65536
    # bin b holds t in [b / nbins, (b + 1) / nbins); take the centres of bins 1000 and 1001
    x = np.where(np.arange(nv) % 2 == 0, np.float32(1000.5 / nbins), np.float32(1001.5 / nbins))
    x[:2] = [0.0, 1.0]
    data = np.zeros(m.span, np.float32)
    data[valid] = x
    g = S.upload(ctx, Field(m, data))
    cnt = S.calc_hist(ctx, g, 0.0, 1.0, nbins, 0.0, counts=True)
    want = O.calc_hist(O.Layout.from_mapping(m), data, 0.0, 1.0, nbins)
    assert np.array_equal(cnt, want)
    assert want[1000] > 1e6 and want[1001] > 1e6
    g.close()


def test_calc_hist_constant_field(ctx, hot):
    """Every voxel in one bin: the count is the voxel count (plain counting)."""
    m, _ = hot
    data = np.where(m.valid_mask(), np.float32(0.5), np.float32(0)).astype(np.float32)
    g = S.upload(ctx, Field(m, data))
    cnt = S.calc_hist(ctx, g, 0.0, 1.0, 65536, 0.0, counts=True)
    want = np.zeros(65536, np.uint32)
    want[32768] = m.data_size   # 65536 * 0.5 - 0.5 = 32767.5 rounds away from zero
    assert np.array_equal(cnt, want)
    g.close()


@pytest.mark.parametrize("nbins", [100, 512, 600, 800])
def test_uv_hist_hot_bins(ctx, hot, nbins):
    """512 and 600 bins: 4 and 6 wide partitions (the joint bins computed
    once); 100: one partition; 800: more than 8 partitions, device atomics."""
    m, data = hot
    lay = O.Layout.from_mapping(m)
    u = S.upload(ctx, Field(m, data))
    vdata = np.where(m.valid_mask(), np.float32(1.0) - data, np.float32(0)).astype(np.float32)
    v = S.upload(ctx, Field(m, vdata))
    h, stride = S.uv_hist(ctx, u, v, 0.0, 1.0, 0.0, 1.0, nbins)
    want, wstride = O.uv_hist(lay, data, vdata, 0.0, 1.0, 0.0, 1.0, nbins)
    assert stride == wstride
    assert np.array_equal(h.reshape(-1)[: want.size], want)
    assert want.max() > 0.5 * m.data_size
    u.close()
    v.close()


@pytest.mark.parametrize("n_images,nstride", [(300, None), (301, 304), (301, 301)])
def test_find_intensity_range_hot_values(ctx, n_images, nstride):
    """u16 volumes where one value holds 70% of the voxels: strides that are
    a multiple of 8 (16-byte loads) or not (one value per lane)."""
    m = S.Mapping.cylinder(n_images, 256, 256, 127.5, 127.5, 400.0, nstride)
    valid = m.valid_mask()
    nv = int(valid.sum())
    rng = np.random.default_rng(n_images)
    r = rng.random(nv)
    x = np.where(r < 0.7, 30001, np.where(r < 0.85, 30000, rng.integers(0, 65536, nv))).astype(np.uint16)
    u16 = np.zeros(m.span, np.uint16)
    u16[valid] = x
    g = S.upload(ctx, Field(m, u16))
    lo, hi, h = S.find_intensity_range(ctx, g, 0.001, hist=True)
    lay = O.Layout.from_mapping(m)
    want = np.bincount(x, minlength=65536).astype(np.uint64)
    assert np.array_equal(h, want)
    assert np.array_equal(h, O.u16_hist(lay, u16))
    assert (lo, hi) == O.find_intensity_range(lay, u16, 0.001)
    g.close()
