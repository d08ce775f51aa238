"""Potts parity on the GPU: HIP FP32 vs the FP32 host mirror (bitwise) and the
FP64 oracle (label agreement, ||V - V_fp64||_inf), through the C ABI."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1.0 / np.sqrt(12.0)))
TAU_REF = float(np.float32(0.99) * np.sqrt(np.float32(1.0 / 12.0)))


def _run_gpu(ctx, ph, c, lam, tau, iters, **kw):
    g = S.upload(ctx, ph.field)
    s = S.PottsSolver(ctx, g, c, lam, tau, tau, dq_trace=True, **kw)
    s.iterate(iters)
    V, xi = s.state()
    res = s.labels().to_host()
    dq = s.dq()
    s.close()
    return res, V, xi, dq


@pytest.mark.parametrize("kernel", ["twopass"])
def test_potts_c1_bitwise_vs_fp32_mirror(ctx, c1_phantom, kernel):
    ph = c1_phantom
    c = ph.normalized_labels()
    iters = 60
    res, V, xi, dq = _run_gpu(ctx, ph, c, 30.0, TAU, iters, kernel=kernel)
    lay = O.Layout.from_mapping(ph.field.mapping)
    r32 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, iters, precision="f32", want_state=True)
    m = ph.field.mapping.valid_mask()
    assert np.array_equal(V[:, m].view(np.uint32), r32["V"][:, m].view(np.uint32))
    assert np.array_equal(xi[:, m].view(np.uint32), r32["xi"][:, m].view(np.uint32))
    assert np.array_equal(res[m], r32["res"][m])
    assert np.array_equal(dq.view(np.uint32), r32["dq"].view(np.uint32))


def test_potts_c1_500_vs_fp64(ctx, c1_phantom):
    """Config 1 (64^3, K=3, lambda=30, tau=1/sqrt(12), 500 iterations)."""
    ph = c1_phantom
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    out = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 500)
    res = out.to_host()
    lay = O.Layout.from_mapping(ph.field.mapping)
    r64 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 500, precision="f64")
    m = ph.field.mapping.valid_mask()
    agree = (res[m] == r64["res"][m]).mean()
    print(f"C1 label agreement vs FP64: {agree:.6f}")
    assert agree >= 0.999
