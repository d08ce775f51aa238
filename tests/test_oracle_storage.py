"""Reduced storage of Xi and V-bar in the oracle, pinned against independent
conversions: torch's float32 -> bfloat16 cast and numpy's float32 -> float16
cast (both IEEE round to nearest even), on random values and on the hard
cases (ties, subnormals, overflow to infinity, signed zeros)."""
import numpy as np
import pytest
import torch

import oracle as O
from segmentctscans_amd.phantom import make_phantom


def _cases():
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    ties16 = (rng.integers(0, 2**16, 2000).astype(np.uint32) << 16 | 0x8000).view(np.float32)
    ties16 = ties16[np.isfinite(ties16)]
    half_ties = (np.arange(1, 2000, dtype=np.float64) * 2.0**-11 + 2.0**-12).astype(np.float32)
    special = np.array([0.0, -0.0, 1e-45, -1e-45, 6e-8, 3e-8, 65504.0, 65520.0, 65519.99, 1e30, -1e30,
                        3.4e38, 1.0, -1.0, 0.5, 2.0**-14, 2.0**-24, 2.0**-25], np.float32)
    return np.concatenate([x, ties16, half_ties, special, rng.standard_normal(5000).astype(np.float32)])


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_bf16_round_matches_torch():
    x = _cases()
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = O.store_round(x, "bf16")
    assert np.array_equal(_bits(got), _bits(want))


def test_f16_round_matches_numpy():
    x = _cases()
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    got = O.store_round(x, "fp16")
    assert np.array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_reduced_storage_state_is_representable_and_agrees(storage):
    """Xi leaves every iteration rounded to the storage type; V stays FP32 and
    on the simplex; labels agree with the FP64 oracle on a small phantom."""
    ph = make_phantom(24, 24, 2.0, 5.0, seed=7)
    c = ph.normalized_labels()
    lay = O.Layout.from_mapping(ph.field.mapping)
    tau = float(np.float32(1 / np.sqrt(12)))
    r = O.potts(lay, ph.field.data, c, 30.0, tau, tau, 40, precision="f32", want_state=True, storage=storage)
    r64 = O.potts(lay, ph.field.data, c, 30.0, tau, tau, 40, precision="f64", want_state=True)
    m = ph.field.mapping.valid_mask()
    xi = r["xi"][:, m]
    assert np.array_equal(_bits(O.store_round(xi, storage)), _bits(xi))
    V = r["V"][:, m]
    assert V.min() >= 0 and (1 - V.sum(0)).min() >= -1e-6
    assert (r["res"][m] == r64["res"][m]).mean() >= 0.995
