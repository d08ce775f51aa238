"""Chambolle ROF and the reductions of the oracle, pinned by closed forms and
exact known answers."""
import numpy as np
import pytest

import oracle as O
from segmentctscans_amd.phantom import make_phantom


def _valid(lay):
    m = np.zeros((lay.span // lay.nstride, lay.nstride), bool)
    m[:, : lay.n_images] = True
    return m.reshape(-1)


def _line(values):
    """A single-line layout (height = width = 1) holding `values` along n."""
    lay = O.Layout(len(values), 1, 1, 0.0, 0.0, 5.0)
    u = np.zeros(lay.span, np.float32)
    u[: len(values)] = values
    return lay, u


# ---- Chambolle ---------------------------------------------------------------
def test_rof_step_edge_exact_1d():
    """1-D ROF of a step: the plateaus move by lambda / length (Neumann ends)."""
    N, m, lam = 20, 8, 0.5
    g = np.r_[np.zeros(m), np.ones(N - m)]
    lay, g32 = _line(g)
    u, xi, dq = O.chambolle(lay, g32, lam, 0.25, 4000, precision="f64")
    want = np.r_[np.full(m, lam / m), np.full(N - m, 1 - lam / (N - m))]
    assert np.abs(u[:N] - want).max() < 1e-6
    assert np.abs(xi[0][:N]).max() <= 1 + 1e-12


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_chambolle_invariants(precision):
    ph = make_phantom(24, 24, 2.0, 5.0, seed=3)
    lay = O.Layout.from_mapping(ph.field.mapping)
    m = _valid(lay)
    u, xi, dq = O.chambolle(lay, ph.field.data, 0.06, 0.99 / 6, 60, precision=precision)
    norm = np.sqrt((xi[:, m].astype(np.float64) ** 2).sum(0))
    assert norm.max() <= 1 + 1e-6
    # TV denoising reduces total variation and keeps the mean (div xi sums to 0)
    assert abs(u[m].mean() - ph.field.data[m].astype(np.float64).mean()) < 1e-5
    # constant image: fixed point
    c = np.where(m, np.float32(0.4), np.float32(0))
    uc, xic, dqc = O.chambolle(lay, c, 0.06, 0.99 / 6, 10, precision=precision)
    assert np.array_equal(uc[m], c[m].astype(uc.dtype)) and np.all(xic == 0)
    # lambda -> 0: u -> g (|div xi| <= 6)
    us, _, _ = O.chambolle(lay, ph.field.data, 1e-6, 0.99 / 6, 10, precision=precision)
    assert np.abs(us[m] - ph.field.data[m]).max() <= 7e-6


def test_chambolle_fp32_vs_fp64():
    ph = make_phantom(24, 24, 2.0, 5.0, seed=4)
    lay = O.Layout.from_mapping(ph.field.mapping)
    m = _valid(lay)
    u32, _, _ = O.chambolle(lay, ph.field.data, 0.08, 0.99 / 6, 100, precision="f32")
    u64, _, _ = O.chambolle(lay, ph.field.data, 0.08, 0.99 / 6, 100, precision="f64")
    assert np.abs(u32[m] - u64[m]).max() < 2e-5


# ---- calc_hist ------------------------------------------------------------------
@pytest.mark.parametrize("nbins", [256, 1024, 65536])
def test_calc_hist_bin_edges_power_of_two(nbins):
    k = np.arange(nbins + 1)
    edges = (k / nbins).astype(np.float32)           # exact in float for nbins <= 2^24
    centres = ((k[:-1] + 0.5) / nbins).astype(np.float32)
    outside = np.array([-0.25, 1.5, np.nan, -np.inf, np.inf], np.float32)
    lay, u = _line(np.r_[edges, centres, outside])
    got = O.calc_hist(lay, u, 0.0, 1.0, nbins)
    want = np.zeros(nbins, np.uint32)
    np.add.at(want, np.minimum(k, nbins - 1), 1)      # edge k -> bin k, 1.0 -> last bin
    np.add.at(want, k[:-1], 1)                        # centre of bin k -> bin k
    assert np.array_equal(got, want)


def test_calc_hist_ramp_closed_form():
    nbins = 1000
    x = ((np.arange(nbins * 7) + 0.5) / (nbins * 7)).astype(np.float32)   # 7 per bin, away from edges
    lay, u = _line(x)
    got = O.calc_hist(lay, u, 0.0, 1.0, nbins)
    assert np.array_equal(got, np.full(nbins, 7, np.uint32))


def test_min_max_and_padding_ignored():
    ph = make_phantom(24, 24, 2.0, 5.0, seed=9)
    lay = O.Layout.from_mapping(ph.field.mapping)
    m = _valid(lay)
    u = ph.field.data.copy()
    u[~m] = 1e9                                       # padding must be ignored
    assert O.min_max(lay, u) == (float(u[m].min()), float(u[m].max()))
    pos = u[m][u[m] > 0]
    assert O.min_max(lay, u, positive=True) == (float(pos.min()), float(pos.max()))


def test_grad_magn_linear_ramp():
    lay = O.Layout(9, 10, 10, 4.5, 4.5, 4.2)
    u = np.zeros(lay.span, np.float32)
    for col in range(lay.col_begin, lay.col_end):
        for row in range(int(lay.start_row[col]), int(lay.start_row[col]) + (int(lay.base[col + 1] - lay.base[col])) // lay.nstride):
            L = int(lay.base[col]) // lay.nstride + row - int(lay.start_row[col])
            u[L * lay.nstride: L * lay.nstride + 9] = np.arange(9) + 2 * row + 3 * col
    v = O.grad_magn(lay, u)
    m = _valid(lay)
    vals = v[m]
    inner = vals[vals != -1]
    assert inner.size > 0 and np.all(inner == np.float32(np.sqrt(np.float32(14))))
    # a voxel on the n ends always lacks a neighbour
    assert np.all(v.reshape(-1, lay.nstride)[:, 0] == -1) and np.all(v.reshape(-1, lay.nstride)[:, 8] == -1)


def test_uv_hist_known_answer():
    nbins = 8
    # (u, v) pairs at bin centres of a [0, 1] x [0, 2] grid, plus rejected points
    us = (np.array([0, 1, 7, 7, 3]) + 0.5) / 8
    vs = 2 * (np.array([0, 0, 7, 7, 4]) + 0.5) / 8
    u = np.r_[us, 2.0, 0.5].astype(np.float32)
    v = np.r_[vs, 1.0, -1.0].astype(np.float32)
    lay, uu = _line(u)
    _, vv = _line(v)
    cnt, stride = O.uv_hist(lay, uu, vv, 0.0, 1.0, 0.0, 2.0, nbins)
    assert stride == 32
    want = np.zeros_like(cnt)
    for x, y in [(0, 0), (1, 0), (7, 7), (7, 7), (3, 4)]:
        want[y * stride + x] += 1
    assert np.array_equal(cnt, want)
