"""sct_calc_hist_noalloc (no counterpart in the reference; for callers that
cannot allocate) against sct_calc_hist through the C ABI: identical counts
and densities for every counting mode (LDS counters <= 8192 bins, bin
partitions <= 262144 bins, with and without dither), two calls in a row with
the second on a smaller input, and the size limit rejected."""
import ctypes as C

import numpy as np
import pytest

import segmentctscans_amd as S
from segmentctscans_amd import _native as N
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu


def _hist(fn, ctx, f, rs, re, nbins, delta, seed):
    cnt = np.full(nbins, 0xDEADBEEF, np.uint32)
    den = np.full(nbins, np.nan)
    rc = fn(ctx.handle, f.handle, rs, re, nbins, delta, seed, cnt.ctypes.data_as(N.u32p), den.ctypes.data_as(N.f64p))
    ctx.check(rc)
    return cnt, den


@pytest.fixture(scope="module")
def fields(ctx):
    big = make_phantom(160, 128, 4.0, 20.0, seed=41)
    small = make_phantom(40, 48, 2.0, 6.0, seed=42)
    out = [S.upload(ctx, big.field), S.upload(ctx, small.field)]
    yield out
    for f in out:
        f.close()


@pytest.mark.parametrize("nbins,delta", [(64, 0.0), (8192, 0.0), (1000, 0.01), (65536, 0.0), (65536, 0.002),
                                         (262144, 0.0), (100000, 0.0)])
def test_noalloc_equals_calc_hist_twice(ctx, fields, nbins, delta):
    L = N.lib()
    for f in fields:   # the second call on a smaller input
        mn, mx = S.min_max(ctx, f)
        want = _hist(L.sct_calc_hist, ctx, f, mn, mx, nbins, delta, 7)
        got = _hist(L.sct_calc_hist_noalloc, ctx, f, mn, mx, nbins, delta, 7)
        assert np.array_equal(got[0], want[0]), (f.mapping.n_images, nbins, delta)
        assert np.array_equal(got[1], want[1])
        assert int(got[0].sum()) <= f.mapping.data_size
        if delta == 0.0:
            assert int(got[0].sum()) == f.mapping.data_size


def test_noalloc_rejects_too_many_bins(ctx, fields):
    L = N.lib()
    f = fields[1]
    nb = 262145
    cnt = np.zeros(nb, np.uint32)
    rc = L.sct_calc_hist_noalloc(ctx.handle, f.handle, 0.0, 1.0, nb, 0.0, 0, cnt.ctypes.data_as(N.u32p), None)
    assert rc == N.SCT_ERR_INVALID
    assert b"SCT_NOALLOC_MAX_BINS" in L.sct_ctx_last_error(ctx.handle)
