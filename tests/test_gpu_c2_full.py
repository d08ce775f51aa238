"""The TV workload at its full size: Chambolle TV (ROF) on the 512^3 phantom
(cylinder radius 0.45 * 512 = 230.4, spheres U[4, 20], seed 7), lambda =
0.12, tau = 0.99 / 6, 200 iterations, then the 65536-bin histogram of the
result over the input's [min, max] (delta = 0).

Bit-exact against the oracle: min_max (min_max.cpp:18-81), the TV result
against the FP32 mirror of chambolle.cpp:33-276, and the 65536-bin counts
(calc_hist.cpp:23-138) against the oracle's histogram of that result."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd.phantom import make_phantom

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

TAU_TV = float(np.float32(1.0 / 6.0) * np.float32(0.99))


def test_tv_512_then_65536_bin_histogram_bit_exact(ctx):
    ph = make_phantom(512, 512, 4.0, 20.0, seed=7, r=230.4)
    m = ph.field.mapping
    lay = O.Layout.from_mapping(m)
    valid = m.valid_mask()
    g = S.upload(ctx, ph.field)
    mn, mx = S.min_max(ctx, g)
    assert (mn, mx) == O.min_max(lay, ph.field.data)
    S.chambolle(ctx, g, 0.12, TAU_TV, 200)
    got = g.to_host()
    cnt = S.calc_hist(ctx, g, mn, mx, 65536, 0.0, counts=True)
    g.close()
    print(f"512^3 ({m.data_size} voxels): TV and histogram done on the GPU", flush=True)
    want, _, _ = O.chambolle(lay, ph.field.data, 0.12, TAU_TV, 200, precision="f32")
    diff = int(np.count_nonzero(got[valid].view(np.uint32) != want[valid].view(np.uint32)))
    assert diff == 0, f"{diff} voxels differ from the FP32 mirror"
    ref = O.calc_hist(lay, want, mn, mx, 65536)
    assert int(ref.sum()) == int(cnt.sum())
    assert np.array_equal(cnt, ref)
    print(f"TV bitwise == FP32 mirror; 65536-bin counts == oracle ({int(cnt.sum())} in range, "
          f"{int(np.count_nonzero(cnt))} non-empty bins)")
