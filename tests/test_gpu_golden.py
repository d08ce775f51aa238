"""The HIP library reproduces the committed golden fixtures bit for bit: Potts
(two-pass and fused kernels, K = 2..6), Chambolle and the 64 / 65536-bin
histograms.  No oracle call: the expected bits come from the fixtures."""
import numpy as np
import pytest

import segmentctscans_amd as S
from golden_io import TAU, TAU_REF, TAU_TV, field, load

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("kernel", ["twopass", "fused"])
def test_potts_small(ctx, kernel):
    z = load("potts_small.npz")
    f = field(z)
    valid = f.mapping.valid_mask()
    g = S.upload(ctx, f)
    for K in range(2, 7):
        s = S.PottsSolver(ctx, g, z[f"c{K}"], 30.0, TAU_REF, TAU_REF, dq_trace=True, kernel=kernel)
        s.iterate(10)
        V, xi = s.state()
        res = s.labels().to_host()
        assert np.array_equal(_bits(V[:, valid]), _bits(z[f"V{K}"][:, valid])), K
        assert np.array_equal(_bits(xi[:, valid]), _bits(z[f"xi{K}"][:, valid])), K
        assert np.array_equal(res[valid], z[f"res{K}"][valid]), K
        assert np.array_equal(_bits(s.dq()), _bits(z[f"dq{K}"])), K
        s.close()
    g.close()


@pytest.mark.parametrize("kernel", ["twopass", "fused"])
def test_potts_24_cube(ctx, kernel):
    z = load("potts_24.npz")
    f = field(z)
    valid = f.mapping.valid_mask()
    g = S.upload(ctx, f)
    out, dq = S.potts_min_partition(ctx, g, z["c"], 30.0, TAU, TAU, 50, kernel=kernel, dq_trace=True)
    assert np.array_equal(out.to_host()[valid], z["res"][valid])
    assert np.array_equal(_bits(dq), _bits(z["dq"]))
    g.close()


def test_chambolle_24_cube(ctx):
    z = load("chambolle_24.npz")
    f = field(z, "g")
    valid = f.mapping.valid_mask()
    for tag, lam in (("004", 0.04), ("012", 0.12)):
        g = S.upload(ctx, f)
        dq = S.chambolle(ctx, g, lam, TAU_TV, 50, dq_trace=True)
        assert np.array_equal(_bits(g.to_host()[valid]), _bits(z[f"out_{tag}"][valid]))
        assert np.array_equal(_bits(dq), _bits(z[f"dq_{tag}"]))
        g.close()


def test_hist_edges(ctx):
    z = load("hist_edges.npz")
    g = S.upload(ctx, field(z))
    for nb in (64, 65536):
        got = S.calc_hist(ctx, g, float(z["rs"]), float(z["re"]), nb, 0.0, counts=True)
        assert np.array_equal(got, z[f"counts_{nb}"])
    g.close()
