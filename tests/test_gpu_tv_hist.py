"""Chambolle TV and the reductions on the GPU through the C ABI, against the
oracle: Chambolle bitwise vs the FP32 mirror, histograms / extrema / gradient
magnitude bit-exact."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu

TAU_TV = float(np.float32(1.0 / 6.0) * np.float32(0.99))


@pytest.fixture(scope="module")
def ph24():
    return make_phantom(24, 24, 2.0, 5.0, seed=31)


@pytest.fixture(scope="module")
def ph128():
    """Config-2-style phantom at 128^3 (radius 0.45 N, spheres U[4, 20])."""
    return make_phantom(128, 128, 4.0, 20.0, seed=32)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("lam", [0.04, 0.12])
def test_chambolle_bitwise_vs_fp32_mirror(ctx, ph24, lam):
    g = S.upload(ctx, ph24.field)
    dq = S.chambolle(ctx, g, lam, TAU_TV, 50, dq_trace=True)
    got = g.to_host()
    lay = O.Layout.from_mapping(ph24.field.mapping)
    want, _, dq32 = O.chambolle(lay, ph24.field.data, lam, TAU_TV, 50, precision="f32")
    m = ph24.field.mapping.valid_mask()
    assert np.array_equal(_bits(got[m]), _bits(want[m]))
    assert np.array_equal(_bits(dq), _bits(dq32))


@pytest.mark.parametrize("n_images,nstride", [(200, None), (129, None), (37, 37), (130, 131), (256, None),
                                              (33, 34)])
def test_chambolle_chunk_edges_bitwise(ctx, n_images, nstride):
    """Lines longer than one chunk, partial chunks, odd lengths and odd
    strides (one voxel per lane; nstride 34: two), bitwise vs the FP32 mirror,
    dq trace too."""
    mp = S.Mapping.cylinder(n_images, 21, 19, 9.3, 8.7, 8.9, nstride)
    rng = np.random.default_rng(n_images)
    data = np.zeros(mp.span, np.float32)
    valid = mp.valid_mask()
    data[valid] = rng.random(int(valid.sum()), dtype=np.float32)
    f = S.Field(mp, data)
    g = S.upload(ctx, f)
    dq = S.chambolle(ctx, g, 0.08, TAU_TV, 12, dq_trace=True)
    lay = O.Layout.from_mapping(mp)
    want, _, dq32 = O.chambolle(lay, data, 0.08, TAU_TV, 12, precision="f32")
    assert np.array_equal(_bits(g.to_host()[valid]), _bits(want[valid]))
    assert np.array_equal(_bits(dq), _bits(dq32))


def test_tv_then_65536_bin_histogram_bit_exact(ctx, ph128):
    """Config 2 pipeline on a 128^3 phantom: TV at the nine weights
    0.04..0.12 (200 iterations each, from the same input), then a 65536-bin
    histogram over the input's [min, max]; bit-exact against the oracle."""
    lay = O.Layout.from_mapping(ph128.field.mapping)
    m = ph128.field.mapping.valid_mask()
    g0 = S.upload(ctx, ph128.field)
    mn, mx = S.min_max(ctx, g0)
    assert (mn, mx) == O.min_max(lay, ph128.field.data)
    for k, lam in enumerate(np.round(np.arange(0.04, 0.1201, 0.01), 2)):
        g = S.upload(ctx, ph128.field)
        S.chambolle(ctx, g, float(lam), TAU_TV, 200)
        got = g.to_host()
        cnt = S.calc_hist(ctx, g, mn, mx, 65536, 0.0, counts=True)
        if k in (0, 8):   # the FP32 mirror is slow on the host: check both ends
            want, _, _ = O.chambolle(lay, ph128.field.data, float(lam), TAU_TV, 200, precision="f32")
            assert np.array_equal(_bits(got[m]), _bits(want[m]))
        assert np.array_equal(cnt, O.calc_hist(lay, got, mn, mx, 65536))
        assert cnt.sum() == np.count_nonzero((got[m] >= mn) & (got[m] <= mx))


@pytest.mark.parametrize("nbins", [1, 256, 1000, 1024, 8192, 8193, 32768, 65536, 100000, 300000, 600000])
def test_calc_hist_modes_bit_exact(ctx, ph128, nbins):
    lay = O.Layout.from_mapping(ph128.field.mapping)
    g = S.upload(ctx, ph128.field)
    cnt = S.calc_hist(ctx, g, 0.1, 0.7, nbins, 0.0, counts=True)
    assert np.array_equal(cnt, O.calc_hist(lay, ph128.field.data, 0.1, 0.7, nbins))
    den, mn, mx = S.calc_hist(ctx, g, nbins, 0.0)
    want = O.calc_hist(lay, ph128.field.data, mn, mx, nbins)
    bw = (np.float64(mx) - np.float64(mn)) / nbins
    assert np.array_equal(den, want * (1.0 / (ph128.field.mapping.data_size * bw)))


def test_calc_hist_dither_deterministic(ctx, ph128):
    g = S.upload(ctx, ph128.field)
    a = S.calc_hist(ctx, g, 0.0, 1.0, 1024, 1e-3, seed=5, counts=True)
    b = S.calc_hist(ctx, g, 0.0, 1.0, 1024, 1e-3, seed=5, counts=True)
    c = S.calc_hist(ctx, g, 0.0, 1.0, 1024, 1e-3, seed=6, counts=True)
    plain = S.calc_hist(ctx, g, 0.0, 1.0, 1024, 0.0, counts=True)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(int(a.sum()) - int(plain.sum())) < 0.01 * plain.sum()
    # the dither of a voxel does not depend on the bins: every layout of the
    # counters (LDS, partitioned LDS, device atomics) accepts the same voxels
    for nb in (8193, 65536, 300000):
        big = S.calc_hist(ctx, g, 0.0, 1.0, nb, 1e-3, seed=5, counts=True)
        assert int(big.sum()) == int(a.sum())
        assert np.array_equal(big, S.calc_hist(ctx, g, 0.0, 1.0, nb, 1e-3, seed=5, counts=True))


def test_histogram_known_answers_on_device(ctx):
    nbins = 256
    k = np.arange(nbins + 1)
    vals = np.r_[(k / nbins), (k[:-1] + 0.5) / nbins, [-0.25, 1.5, np.nan]].astype(np.float32)
    m = S.Mapping.cylinder(len(vals), 1, 1, 0.0, 0.0, 5.0)
    f = S.Field(m)
    f.data[: len(vals)] = vals
    g = S.upload(ctx, f)
    cnt = S.calc_hist(ctx, g, 0.0, 1.0, nbins, 0.0, counts=True)
    want = np.zeros(nbins, np.uint32)
    np.add.at(want, np.minimum(k, nbins - 1), 1)
    np.add.at(want, k[:-1], 1)
    assert np.array_equal(cnt, want)


def test_grad_magn_uv_hist_bit_exact(ctx, ph128):
    lay = O.Layout.from_mapping(ph128.field.mapping)
    m = ph128.field.mapping.valid_mask()
    u = S.upload(ctx, ph128.field)
    v = S.grad_magn(ctx, u)
    vh = v.to_host()
    assert np.array_equal(_bits(vh[m]), _bits(O.grad_magn(lay, ph128.field.data)[m]))
    assert S.grad_min_max(ctx, v) == O.min_max(lay, vh, positive=True)
    r = S.uv_hist(ctx, u, v, 512)
    u0, u1 = O.min_max(lay, ph128.field.data)
    v0, v1 = O.min_max(lay, vh, positive=True)
    assert (r.u_start, r.u_end, r.v_start, r.v_end) == (u0, u1, v0, v1)
    want, stride = O.uv_hist(lay, ph128.field.data, vh, u0, u1, v0, v1, 512)
    assert r.stride == stride
    assert np.array_equal(r.hist.reshape(-1)[: want.size], want)
    # voxels without all six neighbours carry -1 and are dropped by v >= v0 > 0
    assert want.sum() == np.count_nonzero(vh[m] > 0)
    h2, s2 = S.uv_hist(ctx, u, v, 0.2, 0.6, 0.0, 0.5, 100)
    w2, _ = O.uv_hist(lay, ph128.field.data, vh, 0.2, 0.6, 0.0, 0.5, 100)
    assert np.array_equal(h2.reshape(-1)[: w2.size], w2)
    # 6 wide partitions; then more bins than 8 partitions hold: device
    # atomics, wave-aggregated
    for nb in (600, 800):
        h3, _ = S.uv_hist(ctx, u, v, u0, u1, v0, v1, nb)
        w3, _ = O.uv_hist(lay, ph128.field.data, vh, u0, u1, v0, v1, nb)
        assert np.array_equal(h3.reshape(-1)[: w3.size], w3)


def test_download_size_mismatch_raises(ctx, ph24):
    g = S.upload(ctx, ph24.field)
    bad = S.Field(S.Mapping.cylinder(8, 8, 8, 3.5, 3.5, 3.0))
    with pytest.raises(S.SctError) as e:
        S.download(ctx, g, bad)
    assert e.value.code == 3


def test_device_memory(ctx):
    avail, total = S.device_memory(ctx)
    assert 0 < avail <= total
    assert total > 200 * 2**30   # MI355X: 288 GB HBM3E


def test_dykstra_cr_sqrt_and_division_match_ieee(ctx):
    """The Dykstra sweep's square root and quotient skip the compiler's range
    scaling; on their ranges (x, b in [1, 2^64), a in [2^-23, b)) they must be
    IEEE correctly rounded: 2^26 random operands each, bit for bit.  The same
    quotient bins the histograms' values (b in [2^-32, 2^32], |a| in
    [b 2^-40, b], both signs), checked the same way."""
    import ctypes as C
    from segmentctscans_amd import _native as N
    bs, bd = C.c_uint64(), C.c_uint64()
    ctx.check(N.lib().sct_selftest_cr_math(ctx.handle, 1 << 26, 12345, C.byref(bs), C.byref(bd)))
    assert bs.value == 0 and bd.value == 0, (bs.value, bd.value)


@pytest.mark.parametrize("n_images,nstride", [(40, None), (40, 45), (64, 64), (42, 44), (301, 304)])
def test_output_fields_padding(ctx, n_images, nstride):
    """grad_magn and convert_to_fp write every voxel of every line; where the
    lines do not cover the span (nstride > n_images) the rest of the new field
    is zero.  Strides that are a multiple of 4 take 16-byte accesses (42 in 44:
    a partial last vector; 301 in 304: two trips per line, the second partial);
    45 takes one voxel per lane.  The second round runs after the first round's fields were
    freed, so memory handed out again would show any stale values."""
    mp = S.Mapping.cylinder(n_images, 17, 15, 7.1, 8.2, 6.9, nstride)
    rng = np.random.default_rng(n_images + (nstride or 0))
    data = np.zeros(mp.span, np.float32)
    valid = mp.valid_mask()
    data[valid] = rng.random(int(valid.sum()), dtype=np.float32)
    lay = O.Layout.from_mapping(mp)
    u = S.upload(ctx, S.Field(mp, data))
    for _ in range(2):
        fv = S.grad_magn(ctx, u)
        v = fv.to_host()
        fv.close()
        assert np.array_equal(_bits(v[valid]), _bits(O.grad_magn(lay, data)[valid]))
        assert np.all(v[~valid] == 0)
        fc = S.convert_to_fp(ctx, u, 0.25, 0.75)
        c = fc.to_host()
        fc.close()
        assert np.array_equal(_bits(c[valid]), _bits(O.convert_to_fp(lay, data, 0.25, 0.75)[valid]))
        assert np.all(c[~valid] == 0)
    u.close()
