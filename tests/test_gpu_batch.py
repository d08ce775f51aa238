"""Batched Potts solves (no counterpart in the reference): every problem of a
batch gets exactly the labels its own potts_min_partition call returns --
mixed shapes (cylinders and boxes), K = 2..6, two lambdas per volume, FP32 and
BF16 storage, iterations launched directly or as one captured hipGraph."""
import numpy as np
import pytest

import segmentctscans_amd as S
from segmentctscans_amd.phantom import make_phantom

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1.0 / np.sqrt(12.0)))
SHAPES = [  # n, height, width, box?, K
    (20, 18, 22, False, 3), (33, 25, 25, True, 4), (48, 30, 20, False, 5),
    (17, 40, 31, True, 3), (26, 26, 26, False, 2), (70, 12, 9, True, 4), (24, 21, 19, False, 6)]


@pytest.fixture(scope="module")
def volumes(ctx):
    out = []
    for i, (n, h, w, box, K) in enumerate(SHAPES):
        r = 4.0 * max(h, w) if box else 0.45 * min(h, w)
        ph = make_phantom(n, max(h, w), 2.0, 5.0, seed=70 + i, height=h, width=w,
                          xc=(w - 1) / 2, yc=(h - 1) / 2, r=r)
        c = np.linspace(0.2, 0.7, K).astype(np.float32)
        out.append((ph, S.upload(ctx, ph.field), c))
    return out


def _problems(volumes):
    return [(g, c, lam) for (_, g, c) in volumes for lam in (20.0, 45.0)]


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_batch_equals_single_solves(ctx, volumes, storage):
    probs = _problems(volumes)
    outs, ms = S.potts_batch(ctx, probs, TAU, TAU, 30, storage=storage, graph=True)
    assert len(outs) == len(probs) and ms > 0
    for (g, c, lam), o in zip(probs, outs):
        want = S.potts_min_partition(ctx, g, c, lam, TAU, TAU, 30, storage=storage).to_host()
        m = g.mapping.valid_mask()
        assert np.array_equal(o.to_host()[m], want[m]), (g.mapping.n_images, c.size, lam)
        o.close()


def test_batch_graph_equals_direct_launches(ctx, volumes):
    probs = _problems(volumes)
    a, _ = S.potts_batch(ctx, probs, TAU, TAU, 12, graph=True)
    b, _ = S.potts_batch(ctx, probs, TAU, TAU, 12, graph=False)
    for x, y, (g, _, _) in zip(a, b, probs):
        m = g.mapping.valid_mask()
        assert np.array_equal(x.to_host()[m], y.to_host()[m])
        x.close()
        y.close()


def test_batch_rejects_bad_problems(ctx, volumes):
    g = volumes[0][1]
    with pytest.raises(ValueError):
        S.potts_batch(ctx, [(g, [0.5], 30.0)], TAU, TAU, 5)
    outs, _ = S.potts_batch(ctx, [], TAU, TAU, 5)
    assert outs == []


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_every_xi_tile_equals_single_solves(ctx, volumes, waves):
    probs = _problems(volumes)
    S.potts_batch_set_tiles(ctx, {K: waves for K in range(2, 7)})
    try:
        outs, _ = S.potts_batch(ctx, probs, TAU, TAU, 20, graph=True)
        assert set(S.potts_batch_tiles(ctx).values()) == {waves}
    finally:
        S.potts_batch_set_tiles(ctx, {})
    for (g, c, lam), o in zip(probs, outs):
        want = S.potts_min_partition(ctx, g, c, lam, TAU, TAU, 20).to_host()
        m = g.mapping.valid_mask()
        assert np.array_equal(o.to_host()[m], want[m]), (waves, g.mapping.n_images, c.size, lam)
        o.close()


def test_autotuned_batch_equals_untuned(ctx, volumes):
    """The autotuner's trial iterations leave no trace: the groups restart
    from their initial state, so the labels are the untuned ones; the choice
    is cached per context."""
    probs = _problems(volumes)
    a, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True, autotune=True)
    tiles = S.potts_batch_tiles(ctx)
    assert tiles and set(tiles.values()) <= {1, 2, 4}
    assert set(tiles) == {c.size for (_, c, _) in probs}
    b, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True)
    for x, y, (g, _, _) in zip(a, b, probs):
        m = g.mapping.valid_mask()
        assert np.array_equal(x.to_host()[m], y.to_host()[m])
        x.close()
        y.close()
    c2, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True, autotune=True)
    assert S.potts_batch_tiles(ctx) == tiles
    for x in c2:
        x.close()
    with pytest.raises(S.SctError):
        S.potts_batch_set_tiles(ctx, {3: 3})


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
@pytest.mark.parametrize("graph", [False, True])
def test_persistent_batch_equals_single_solves(ctx, volumes, storage, graph):
    """One persistent launch per K group (tiles from one ordered ticket
    queue, each half-step waiting for the one before) gives every problem its
    own potts_min_partition labels."""
    probs = _problems(volumes)
    outs, ms = S.potts_batch(ctx, probs, TAU, TAU, 30, storage=storage, graph=graph, persistent=True)
    assert len(outs) == len(probs) and ms > 0
    for (g, c, lam), o in zip(probs, outs):
        want = S.potts_min_partition(ctx, g, c, lam, TAU, TAU, 30, storage=storage).to_host()
        m = g.mapping.valid_mask()
        assert np.array_equal(o.to_host()[m], want[m]), (g.mapping.n_images, c.size, lam)
        o.close()


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_persistent_every_tile_equals_graph(ctx, volumes, waves):
    probs = _problems(volumes)
    S.potts_batch_set_tiles(ctx, {K: waves for K in range(2, 7)})
    try:
        a, _ = S.potts_batch(ctx, probs, TAU, TAU, 20, graph=True)
        b, _ = S.potts_batch(ctx, probs, TAU, TAU, 20, graph=False, persistent=True)
        assert set(S.potts_batch_tiles(ctx).values()) == {waves}
    finally:
        S.potts_batch_set_tiles(ctx, {})
    for x, y, (g, _, _) in zip(a, b, probs):
        m = g.mapping.valid_mask()
        assert np.array_equal(x.to_host()[m], y.to_host()[m]), waves
        x.close()
        y.close()


def test_persistent_autotune(ctx, volumes):
    """The persistent autotuner searches tiles x workgroups per CU; its trial
    launches leave no trace and its choice is cached per context."""
    probs = _problems(volumes)
    a, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True, autotune=True, persistent=True)
    tiles, occ = S.potts_batch_tiles(ctx), S.potts_batch_occupancy(ctx)
    assert set(tiles.values()) <= {1, 2, 4} and all(1 <= v <= 32 for v in occ.values())
    assert set(occ) == {c.size for (_, c, _) in probs}
    b, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True)
    for x, y, (g, _, _) in zip(a, b, probs):
        m = g.mapping.valid_mask()
        assert np.array_equal(x.to_host()[m], y.to_host()[m])
        x.close()
        y.close()
    c2, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=False, autotune=True, persistent=True)
    assert S.potts_batch_tiles(ctx) == tiles and S.potts_batch_occupancy(ctx) == occ
    for x in c2:
        x.close()


def test_persistent_grid_larger_than_resident(ctx, volumes):
    """A persistent launch of 32 workgroups per CU, several times what a CU
    holds (at most 8 of these four-wave workgroups), so most of the grid starts
    only as resident workgroups leave.  A workgroup waits only for tickets that
    started workgroups hold, so the launch completes, with the graph launch's
    labels.  Counts outside 0..64 are rejected."""
    probs = _problems(volumes)
    ks = sorted({c.size for (_, c, _) in probs})
    with pytest.raises(S.SctError):
        S.potts_batch_set_occupancy(ctx, {ks[0]: 65})
    ref, _ = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=True)
    S.potts_batch_set_tiles(ctx, {K: 4 for K in ks})
    S.potts_batch_set_occupancy(ctx, {K: 32 for K in ks})
    try:
        out, ms = S.potts_batch(ctx, probs, TAU, TAU, 15, graph=False, persistent=True)
        assert S.potts_batch_occupancy(ctx) == {K: 32 for K in ks} and ms > 0
    finally:
        S.potts_batch_set_occupancy(ctx, {})
        S.potts_batch_set_tiles(ctx, {})
    for x, y, (g, _, _) in zip(ref, out, probs):
        m = g.mapping.valid_mask()
        assert np.array_equal(x.to_host()[m], y.to_host()[m])
        x.close()
        y.close()
