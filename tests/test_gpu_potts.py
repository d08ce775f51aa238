"""Potts parity on the GPU through the C ABI.

HIP FP32 (two-pass and fused kernels) against the oracle's FP32 mirror: bitwise
equal V, Xi, labels and dq trace.  Against the FP64 oracle: label agreement
>= 99.9% and the max-norm difference of V, stated per case.
"""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1.0 / np.sqrt(12.0)))
TAU_REF = float(np.float32(0.99) * np.float32(np.sqrt(np.float32(1.0 / 12.0))))
KERNELS = ["twopass", "fused"]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _run_solver(ctx, field, c, lam, tau, iters, kernel):
    g = S.upload(ctx, field)
    s = S.PottsSolver(ctx, g, c, lam, tau, tau, dq_trace=True, kernel=kernel)
    s.iterate(iters)
    V, xi = s.state()
    res = s.labels().to_host()
    dq = s.dq()
    s.close()
    g.close()
    return res, V, xi, dq


def _check_bitwise(ctx, field, c, lam, tau, iters, kernel):
    res, V, xi, dq = _run_solver(ctx, field, c, lam, tau, iters, kernel)
    lay = O.Layout.from_mapping(field.mapping)
    r32 = O.potts(lay, field.data, c, lam, tau, tau, iters, precision="f32", want_state=True)
    m = field.mapping.valid_mask()
    assert np.array_equal(_bits(V[:, m]), _bits(r32["V"][:, m])), "V differs"
    assert np.array_equal(_bits(xi[:, m]), _bits(r32["xi"][:, m])), "Xi differs"
    assert np.array_equal(res[m], r32["res"][m]), "labels differ"
    assert np.array_equal(_bits(dq), _bits(r32["dq"])), "dq trace differs"


@pytest.mark.parametrize("kernel", KERNELS)
def test_c1_bitwise_vs_fp32_mirror(ctx, c1_phantom, kernel):
    _check_bitwise(ctx, c1_phantom.field, c1_phantom.normalized_labels(), 30.0, TAU, 60, kernel)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("shape", [(10, 12, 16), (130, 40, 37), (24, 24, 24)])
def test_shapes_and_labels_bitwise(ctx, kernel, K, shape):
    """Odd sizes: n not a multiple of the tile, several row bands, columns
    with one row, K = 2..6 at the reference's tau = 0.99/sqrt(12)."""
    n, h, w = shape
    ph = make_phantom(n, max(h, w), 2.0, 5.0, seed=K + n, height=h, width=w,
                      xc=(w - 1) / 2, yc=(h - 1) / 2, r=0.47 * min(h, w))
    c = np.linspace(0.2, 0.75, K).astype(np.float32)
    _check_bitwise(ctx, ph.field, c, 30.0, TAU_REF, 12, kernel)


@pytest.mark.parametrize("kernel", KERNELS)
def test_c1_500_iterations_vs_fp64(ctx, c1_phantom, kernel):
    """Config 1 (64^3, K=3, lambda=30, tau=1/sqrt(12), 500 iterations)."""
    ph = c1_phantom
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    out, dq = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 500, kernel=kernel, dq_trace=True)
    res = out.to_host()
    s = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, kernel=kernel)
    s.iterate(500)
    V, _ = s.state()
    lay = O.Layout.from_mapping(ph.field.mapping)
    r64 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 500, precision="f64", want_state=True)
    m = ph.field.mapping.valid_mask()
    agree = (res[m] == r64["res"][m]).mean()
    vinf = np.abs(V[:, m] - r64["V"][:, m]).max()
    print(f"C1 {kernel}: label agreement vs FP64 {agree:.6f}, ||V - V_fp64||_inf {vinf:.3e}")
    assert agree >= 0.999
    assert vinf < 1e-3
    assert np.allclose(dq, r64["dq"], rtol=1e-3, atol=1e-9)


@pytest.mark.parametrize("kernel", KERNELS)
def test_multi_slab_equals_single_slab(ctx, kernel):
    """n-slabs with halo = iters (potts-min-partition.cpp:627-675) give the
    one-slab labels bitwise; the forced slab width also exercises the
    out-of-memory return."""
    ph = make_phantom(150, 24, 2.0, 5.0, seed=21)
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    one = S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 8, kernel=kernel).to_host()
    m = ph.field.mapping.valid_mask()
    for slab_w in (32, 64, 96):
        many = S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 8, kernel=kernel, slab_w=slab_w)
        assert np.array_equal(many.to_host()[m], one[m])
    assert S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 20, kernel=kernel, slab_w=32) is None


def test_memory_driven_slabbing(ctx):
    """With device_memory() reporting little memory the solver slabs by itself
    (potts-min-partition.cpp:560-577) and still matches the one-slab result."""
    ph = make_phantom(200, 20, 2.0, 4.0, seed=22)
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    m = ph.field.mapping.valid_mask()
    one = S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 5).to_host()
    lines = ph.field.mapping.lines
    per_slice = lines * 4 * 5 * 2          # two-pass kernels, K = 3: Xi, V-bar, V in FP32
    few = S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 5, mem_available=per_slice * 70)
    assert np.array_equal(few.to_host()[m], one[m])
    assert S.potts_min_partition(ctx, g, c, 30.0, TAU_REF, TAU_REF, 5, mem_available=per_slice * 8) is None


def test_constant_image_fixed_point(ctx):
    ph = make_phantom(40, 20, 2.0, 4.0, seed=23)
    m = ph.field.mapping.valid_mask()
    f = S.Field(ph.field.mapping, np.where(m, np.float32(0.5), np.float32(0)).astype(np.float32))
    g = S.upload(ctx, f)
    c = np.array([0.2, 0.5, 0.9], np.float32)
    res = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 25).to_host()
    assert np.all(res[m] == np.float32(0.5))


def test_invalid_arguments(ctx, c1_phantom):
    g = S.upload(ctx, c1_phantom.field)
    with pytest.raises(ValueError):
        S.potts_min_partition(ctx, g, [0.5], 30.0, TAU, TAU, 5)
    with pytest.raises(S.SctError):
        S.potts_min_partition(ctx, g, [0.2, 0.5], 30.0, TAU, TAU, -1)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("storage", ["bf16", "fp16"])
@pytest.mark.parametrize("K", [2, 3, 5])
def test_reduced_storage_bitwise_vs_mirror(ctx, storage, K, kernel):
    """Xi and V-bar stored as bfloat16 / half, arithmetic FP32: bitwise equal to
    the FP32 mirror with the same storage rounding."""
    ph = make_phantom(70, 30, 2.0, 5.0, seed=40 + K)
    c = np.linspace(0.2, 0.7, K).astype(np.float32)
    g = S.upload(ctx, ph.field)
    s = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, dq_trace=True, storage=storage, kernel=kernel)
    s.iterate(15)
    V, xi = s.state()
    res = s.labels().to_host()
    dq = s.dq()
    lay = O.Layout.from_mapping(ph.field.mapping)
    r = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 15, precision="f32", want_state=True, storage=storage)
    m = ph.field.mapping.valid_mask()
    assert np.array_equal(_bits(V[:, m]), _bits(r["V"][:, m]))
    assert np.array_equal(_bits(xi[:, m]), _bits(r["xi"][:, m]))
    assert np.array_equal(res[m], r["res"][m])
    assert np.array_equal(_bits(dq), _bits(r["dq"]))


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_reduced_storage_c1_agreement_vs_fp64(ctx, c1_phantom, storage):
    ph = c1_phantom
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    s = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, storage=storage)
    s.iterate(500)
    V, _ = s.state()
    res = s.labels().to_host()
    lay = O.Layout.from_mapping(ph.field.mapping)
    r64 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 500, precision="f64", want_state=True)
    m = ph.field.mapping.valid_mask()
    agree = (res[m] == r64["res"][m]).mean()
    vinf = np.abs(V[:, m] - r64["V"][:, m]).max()
    print(f"C1 {storage} storage: label agreement vs FP64 {agree:.6f}, ||V - V_fp64||_inf {vinf:.3e}")
    assert agree >= 0.999


@pytest.mark.parametrize("kernel", KERNELS)
def test_iterate_timed_matches_iterate(ctx, kernel):
    """The timed stepping call runs the same iterations: same state bitwise,
    one positive duration per kernel launch."""
    ph = make_phantom(40, 24, 2.0, 5.0, seed=31)
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    a = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, kernel=kernel)
    b = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, kernel=kernel)
    a.iterate(7)
    t = b.iterate_timed(7)
    assert t.shape == (7, 1 if kernel == "fused" else 2)
    assert np.all(t > 0)
    Va, xa = a.state()
    Vb, xb = b.state()
    assert np.array_equal(_bits(Va), _bits(Vb)) and np.array_equal(_bits(xa), _bits(xb))
    a.close()
    b.close()
    g.close()
