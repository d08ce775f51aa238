"""The oracle reproduces the committed golden fixtures bit for bit (FP32 mirror
and FP64), and the fixtures' FP32 labels agree with their FP64 labels."""
import numpy as np

import oracle as O
from golden_io import TAU, TAU_REF, TAU_TV, load, mapping


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def test_potts_small_k2_to_k6():
    z = load("potts_small.npz")
    m = mapping(z)
    lay = O.Layout.from_mapping(m)
    valid = m.valid_mask()
    for K in range(2, 7):
        c = z[f"c{K}"]
        r = O.potts(lay, z["u"], c, 30.0, TAU_REF, TAU_REF, 10, precision="f32", want_state=True)
        for k in ("V", "xi", "res", "dq"):
            assert np.array_equal(_bits(r[k]), _bits(z[f"{k}{K}"])), (K, k)
        r64 = O.potts(lay, z["u"], c, 30.0, TAU_REF, TAU_REF, 10, precision="f64", want_state=True)
        assert np.array_equal(_bits(r64["V"]), _bits(z[f"V64_{K}"]))
        assert np.array_equal(r64["res"], z[f"res64_{K}"])
        assert (z[f"res{K}"][valid] == z[f"res64_{K}"][valid]).mean() >= 0.999


def test_potts_24_cube():
    z = load("potts_24.npz")
    m = mapping(z)
    lay = O.Layout.from_mapping(m)
    r = O.potts(lay, z["u"], z["c"], 30.0, TAU, TAU, 50, precision="f32", want_state=True)
    for k in ("V", "res", "dq"):
        assert np.array_equal(_bits(r[k]), _bits(z[k])), k
    valid = m.valid_mask()
    assert (z["res"][valid] == z["res64"][valid]).mean() >= 0.999
    assert np.abs(z["V"][:, valid] - z["V64"][:, valid]).max() < 1e-3


def test_chambolle_24_cube():
    z = load("chambolle_24.npz")
    lay = O.Layout.from_mapping(mapping(z))
    for tag, lam in (("004", 0.04), ("012", 0.12)):
        got, _, dq = O.chambolle(lay, z["g"], lam, TAU_TV, 50, precision="f32")
        assert np.array_equal(_bits(got), _bits(z[f"out_{tag}"]))
        assert np.array_equal(_bits(dq), _bits(z[f"dq_{tag}"]))


def test_hist_edges():
    z = load("hist_edges.npz")
    lay = O.Layout.from_mapping(mapping(z))
    for nb in (64, 65536):
        assert np.array_equal(O.calc_hist(lay, z["u"], float(z["rs"]), float(z["re"]), nb), z[f"counts_{nb}"])
    # every voxel but NaN and the out-of-range ones is counted
    assert z["counts_64"].sum() == z["counts_65536"].sum()
