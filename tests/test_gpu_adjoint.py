"""Adjoint of the unrolled Potts iteration on the GPU (FP32, no counterpart in
the reference) against the FP64 host adjoint of the same loss: dL/dc,
dL/dlambda to 1e-3 relative, dL/du to 1e-3 in relative L2 norm, K = 2..6; the
checkpoint interval changes nothing; and the 128^3 case (200 iterations,
checkpoints every 16) against the FP64 host adjoint and FP64 central
differences of the loss with the start labels held fixed."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd import Field
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1.0 / np.sqrt(12.0)))


def _setup(ctx, ph, K):
    c = ph.normalized_labels() if K == 3 else np.linspace(0.1, 0.8, K).astype(np.float32)
    tgt = np.minimum(ph.labels.astype(np.float32), K - 1)
    g = S.upload(ctx, ph.field)
    t = S.upload(ctx, Field(ph.field.mapping, tgt))
    lay = O.Layout.from_mapping(ph.field.mapping)
    return c, tgt, g, t, lay


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
def test_adjoint_matches_fp64_host(ctx, K):
    ph = make_phantom(20, 20, 2.0, 4.0, seed=95 + K)
    c, tgt, g, t, lay = _setup(ctx, ph, K)
    iters = 15
    r = S.potts_adjoint(ctx, g, t, c, 30.0, TAU, TAU, iters, checkpoint_every=4)
    loss, gc, gl, gu = O.adj_grad(lay, ph.field.data, tgt, c, 30.0, TAU, TAU, iters, ck=4)
    m = ph.field.mapping.valid_mask()
    print(f"K={K}: loss {r['loss']:.8e} vs {loss:.8e}; dc {r['grad_c']} vs {gc}; dlambda {r['grad_lambda']:.6e} vs "
          f"{gl:.6e}; |du - du64|/|du64| {_rel_l2(r['grad_u'][m], gu[m]):.2e}")
    assert r["loss"] == pytest.approx(loss, rel=1e-5)
    assert np.allclose(r["grad_c"], gc, rtol=1e-3, atol=1e-3 * np.abs(gc).max())
    assert r["grad_lambda"] == pytest.approx(gl, rel=1e-3)
    assert _rel_l2(r["grad_u"][m], gu[m]) < 1e-3
    assert np.all(r["grad_u"][~m] == 0)


@pytest.mark.parametrize("iters,cks", [(13, (0, 3, 13, 20)), (12, (2, 4, 6, 12))])
def test_checkpoint_interval_does_not_change_result(ctx, iters, cks):
    """Every state is stored with checkpoint_every=1 (no recompute); other
    intervals recompute the states inside each segment and take its last state
    from the next checkpoint or the final state, including intervals that
    divide iters and intervals longer than the run."""
    ph = make_phantom(18, 18, 2.0, 4.0, seed=99)
    c, tgt, g, t, _ = _setup(ctx, ph, 3)
    ref = S.potts_adjoint(ctx, g, t, c, 30.0, TAU, TAU, iters, checkpoint_every=1)
    for ck in cks:
        r = S.potts_adjoint(ctx, g, t, c, 30.0, TAU, TAU, iters, checkpoint_every=ck)
        assert r["loss"] == pytest.approx(ref["loss"], rel=1e-12)
        assert np.allclose(r["grad_c"], ref["grad_c"], rtol=1e-9)
        assert r["grad_lambda"] == pytest.approx(ref["grad_lambda"], rel=1e-9)
        assert np.array_equal(r["grad_u"], ref["grad_u"])


@pytest.mark.timeout(900)
def test_adjoint_128_cube_200_iterations(ctx):
    """128^3 cylinder (r = 57.6, 1,334,784 voxels), K = 3, lambda = 30,
    200 iterations, checkpoints every 16."""
    ph = make_phantom(128, 128, 4.0, 12.0, seed=5)
    assert ph.field.mapping.data_size == 1334784
    c, tgt, g, t, lay = _setup(ctx, ph, 3)
    iters = 200
    r = S.potts_adjoint(ctx, g, t, c, 30.0, TAU, TAU, iters, checkpoint_every=16)
    loss, gc, gl, gu = O.adj_grad(lay, ph.field.data, tgt, c, 30.0, TAU, TAU, iters, ck=16)
    m = ph.field.mapping.valid_mask()
    eu = _rel_l2(r["grad_u"][m], gu[m])
    print(f"128^3: loss {r['loss']:.8e} vs FP64 {loss:.8e}; dc {r['grad_c']} vs {gc}; dlambda {r['grad_lambda']:.6e} "
          f"vs {gl:.6e}; |du - du64|/|du64| {eu:.2e}; forward {r['forward_ms']:.1f} ms, reverse "
          f"{r['backward_ms']:.1f} ms (ratio {r['backward_ms'] / r['forward_ms']:.2f})")
    assert np.allclose(r["grad_c"], gc, rtol=1e-3)
    assert r["grad_lambda"] == pytest.approx(gl, rel=1e-3)
    assert eu < 1e-3
    # FP64 central differences of the same loss (start labels fixed).  The loss
    # is piecewise smooth (simplex active sets, Dykstra ball tests); h = 1e-6
    # stays clear of the kinks, h = 1e-4 straddles some (measured 1.1e-3 and
    # 1.25e-3 on dc_0 and dc_1 here).
    init = O.nearest_labels(ph.field.data, c)
    u64 = ph.field.data.astype(np.float64)
    c64 = c.astype(np.float64)

    def L(cc=c64, ll=30.0):
        return O.adj_loss(lay, u64, tgt, cc, ll, TAU, TAU, iters, init=init)

    for h, rel in ((1e-6, 1e-3), (1e-4, 2.5e-3)):
        fd_c = []
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            fd_c.append((L(cc=c64 + e) - L(cc=c64 - e)) / (2 * h))
        fd_l = (L(ll=30.0 + h) - L(ll=30.0 - h)) / (2 * h)
        print(f"128^3 central differences h={h:g}: dc {np.array(fd_c)}, dlambda {fd_l:.6e}")
        assert np.allclose(gc, fd_c, rtol=rel)
        assert gl == pytest.approx(fd_l, rel=rel)


def test_tape_forward_backward(ctx):
    """The split forward / backward (PottsTape): V of the forward pass is bitwise
    the solver's V after the same iterations; backward of the MSE loss's
    dL/dV equals potts_adjoint (which the FP64 host adjoint pins); backward is
    repeatable and exactly linear in dV (2 dV -> exactly twice every gradient)."""
    K, iters, lam = 3, 12, 30.0
    ph = make_phantom(18, 18, 2.0, 4.0, seed=101)
    c, tgt, g, t, _ = _setup(ctx, ph, K)
    mp = ph.field.mapping
    m = mp.valid_mask()
    tape = S.PottsTape(ctx, g, c, lam, TAU, TAU, iters, checkpoint_every=5)
    V = [f.to_host() for f in tape.V]
    sol = S.PottsSolver(ctx, g, c, lam, TAU, TAU)
    sol.iterate(iters)
    Vs, _ = sol.state()
    sol.close()
    last = np.ones(mp.span, np.float32)
    for i in range(K - 1):
        assert np.array_equal(V[i][m].view(np.uint32), Vs[i][m].view(np.uint32))
        last = last - Vs[i]
    assert np.array_equal(V[K - 1][m].view(np.uint32), last[m].view(np.uint32))
    # dL/dV of L = 1/(N K) sum (V_i - [target = i])^2
    wgt = np.float32(1.0 / (m.sum() * K))
    dv = []
    for i in range(K):
        d = np.zeros(mp.span, np.float32)
        d[m] = np.float32(2) * wgt * (V[i][m] - (tgt[m] == i).astype(np.float32))
        dv.append(d)
    dev = [S.upload(ctx, Field(mp, d)) for d in dv]
    r = tape.backward(dev)
    ref = S.potts_adjoint(ctx, g, t, c, lam, TAU, TAU, iters, checkpoint_every=5)
    print(f"tape dc {r['grad_c']} vs {ref['grad_c']}; dlambda {r['grad_lambda']:.8e} vs {ref['grad_lambda']:.8e}")
    assert np.allclose(r["grad_c"], ref["grad_c"], rtol=1e-5, atol=1e-6 * np.abs(ref["grad_c"]).max())
    assert r["grad_lambda"] == pytest.approx(ref["grad_lambda"], rel=1e-5)
    assert _rel_l2(r["grad_u"][m], ref["grad_u"][m]) < 1e-5
    again = tape.backward(dev)
    assert np.array_equal(again["grad_c"], r["grad_c"]) and again["grad_lambda"] == r["grad_lambda"]
    assert np.array_equal(again["grad_u"], r["grad_u"])
    dev2 = [S.upload(ctx, Field(mp, np.float32(2) * d)) for d in dv]
    r2 = tape.backward(dev2)
    assert np.array_equal(r2["grad_c"], 2 * r["grad_c"]) and r2["grad_lambda"] == 2 * r["grad_lambda"]
    assert np.array_equal(r2["grad_u"], np.float32(2) * r["grad_u"])
    for f in dev + dev2:
        f.close()
    tape.close()


def test_tape_rejects_mismatched_gradients(ctx):
    ph = make_phantom(12, 12, 2.0, 3.0, seed=102)
    c, _, g, _, _ = _setup(ctx, ph, 3)
    tape = S.PottsTape(ctx, g, c, 30.0, TAU, TAU, 4, want_v=False)
    for n, side in ((14, 12), (12, 14)):   # other n_images (same span), other cylinder
        other = S.upload(ctx, make_phantom(n, side, 2.0, 3.0, seed=103).field)
        with pytest.raises(S.SctError):
            tape.backward([g, g, other])
        other.close()
    with pytest.raises(ValueError):
        tape.backward([g, g])
    tape.close()
