"""The uv-hist driver end to end on the GPU: a phantom written as a .bin
archive; the PNG equals the image built from the oracle's gradient magnitude
and 512 x 512 joint histogram of the same input."""
import numpy as np
import pytest

import oracle as O
from png_io import read_png
from segmentctscans_amd.archive import save_bin
from segmentctscans_amd.phantom import make_phantom
from segmentctscans_amd.uv_hist_cli import NBINS, histogram_image, main

pytestmark = pytest.mark.gpu


def test_uv_hist_driver_end_to_end(tmp_path, capsys):
    ph = make_phantom(48, 52, 3.0, 9.0, seed=321)
    src = tmp_path / "in.bin"
    save_bin(ph.field, src)
    out = tmp_path / "uv.png"
    assert main([str(src), "-o", str(out)]) == 0
    lay = O.Layout.from_mapping(ph.field.mapping)
    u = ph.field.data
    v = O.grad_magn(lay, u)
    u0, u1 = O.min_max(lay, u)
    v0, v1 = O.min_max(lay, v, positive=True)
    flat, stride = O.uv_hist(lay, u, v, u0, u1, v0, v1, NBINS)
    full = np.zeros(NBINS * stride, np.uint32)
    full[: flat.size] = flat
    hdr, px = read_png(out.read_bytes())
    assert (hdr["w"], hdr["h"], hdr["color"]) == (NBINS, NBINS, 2)
    assert np.array_equal(px, histogram_image(full.reshape(NBINS, stride)))
    printed = capsys.readouterr().out
    assert f"u ∈ [{u0:g}, {u1:g}], v ∈ [{v0:g}, {v1:g}]" in printed
