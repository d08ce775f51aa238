"""FP64 primal-dual gap on the GPU (no counterpart in the reference).

The GPU's E(V) and D(Xi) on a state equal the oracle's FP64 energies of the
same downloaded state (independent implementations of the same definition,
differing only in summation order); the gap recorded every 50 iterations of
the 64^3 phantom solve is non-negative and shrinks."""
import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1.0 / np.sqrt(12.0)))


@pytest.mark.parametrize("kernel,storage", [("twopass", "fp32"), ("fused", "fp32"), ("twopass", "bf16")])
@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
def test_gap_matches_oracle_energies(ctx, K, kernel, storage):
    ph = make_phantom(30, 22, 2.0, 4.0, seed=60 + K)
    c = np.linspace(0.2, 0.75, K).astype(np.float32)
    g = S.upload(ctx, ph.field)
    s = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, kernel=kernel, storage=storage)
    s.iterate(25)
    E, D = s.gap()
    V, xi = s.state()
    lay = O.Layout.from_mapping(ph.field.mapping)
    Eo, Do = O.potts_gap(lay, ph.field.data, c, 30.0, V, xi)
    print(f"K={K} {kernel} {storage}: E {E:.12e} D {D:.12e} rel gap {(E - D) / E:.3e}")
    tol = 1e-10 if K <= 3 else 1e-7     # K >= 4: IRLS support values, rounding-sensitive
    assert E == pytest.approx(Eo, rel=tol)
    assert D == pytest.approx(Do, rel=1e-10, abs=1e-9 * abs(Eo))
    assert E >= D
    s.close()
    g.close()


def test_gap_trace_c1(ctx, c1_phantom):
    ph = c1_phantom
    c = ph.normalized_labels()
    g = S.upload(ctx, ph.field)
    out, gaps = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 500, gap_every=50)
    assert out is not None and gaps.shape == (10,)
    print("C1 relative gap every 50 iterations:", " ".join(f"{x:.3e}" for x in gaps))
    assert np.all(gaps >= 0)
    assert gaps[-1] < gaps[0] and gaps[-1] < 0.02
    # the stepping interface records the same values
    s = S.PottsSolver(ctx, g, c, 30.0, TAU, TAU, gap_every=50, gap_capacity=10)
    s.iterate(230)
    s.iterate(270)
    assert np.array_equal(s.gap_trace(), gaps)
    s.close()
    g.close()


def test_gap_stopping_test_on_c1(ctx, c1_phantom):
    """The FP64 stopping test: with gap_tol the solve of the 64^3 phantom ends
    at the first gap evaluation (every 10 iterations) whose relative gap is
    below the tolerance, and its labels equal a fixed-iteration run of that
    many iterations."""
    c = c1_phantom.normalized_labels()
    g = S.upload(ctx, c1_phantom.field)
    res, gaps, done = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 500, gap_every=10, gap_tol=1e-4)
    print(f"stopped after {done} iterations, relative gap {gaps[-1]:.3e}")
    assert 0 < done < 500 and done % 10 == 0
    assert gaps.size == done // 10
    assert gaps[-1] < 1e-4 and np.all(gaps[:-1] >= 1e-4)
    fixed = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, done)
    m = c1_phantom.field.mapping.valid_mask()
    assert np.array_equal(res.to_host()[m], fixed.to_host()[m])
    # never below the tolerance: all the iterations run
    _, _, full = S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 30, gap_every=10, gap_tol=1e-12)
    assert full == 30
    with pytest.raises(S.SctError):
        S.potts_min_partition(ctx, g, c, 30.0, TAU, TAU, 30, gap_tol=1e-4)   # no gap_every
    res.close()
    fixed.close()
    g.close()
