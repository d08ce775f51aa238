"""Oracle operators pinned by mathematics, not by re-typing their formulas.

* forward gradient / backward divergence: adjointness <grad v, p> = -<v, div p>
  on the masked domain with the reference's neighbour predicates (exact with
  integer data), and constants are annihilated;
* simplex projection: KKT conditions and an independent bisection solver;
* Dykstra projection onto {|q_i - q_j| <= 1}: K = 2 closed form, K >= 3
  feasibility + variational inequality of the 10^4-sweep limit;
* the Potts iteration itself: two iterations recomposed in numpy from the
  pinned pieces match the oracle's solver state.
"""
import numpy as np
import pytest

import oracle as O

LAYOUTS = [(10, 12, 16, 7.5, 5.5, 5.0), (7, 9, 11, 5.0, 4.0, 4.0), (6, 10, 10, 4.5, 4.5, 3.2), (5, 1, 1, 0, 0, 3)]


@pytest.mark.parametrize("args", LAYOUTS)
def test_grad_div_adjoint_exact(args):
    lay = O.Layout(*args)
    rng = np.random.default_rng(0)
    valid = np.zeros(lay.span, bool)
    v = np.zeros(lay.span)
    v_all = rng.integers(-50, 50, lay.span).astype(np.float64)
    # only in-mask voxels carry data
    m = _valid(lay)
    v[m] = v_all[m]
    p = [np.where(m, rng.integers(-50, 50, lay.span), 0).astype(np.float64) for _ in range(3)]
    g = O.grad(lay, v)
    d = O.div(lay, p)
    lhs = sum(float((gi * pi)[m].sum()) for gi, pi in zip(g, p))
    rhs = -float((v * d)[m].sum())
    assert lhs == rhs
    # constants are annihilated
    g1 = O.grad(lay, np.where(m, 3.0, 0.0))
    assert all(np.all(x[m] == 0) for x in g1)
    del valid


def _valid(lay):
    lines = lay.span // lay.nstride
    m = np.zeros((lines, lay.nstride), bool)
    m[:, : lay.n_images] = True
    return m.reshape(-1)


def _simplex_bisect(q):
    lo, hi = q.min() - 1.0, q.max()
    for _ in range(200):
        th = 0.5 * (lo + hi)
        if np.maximum(q - th, 0).sum() > 1:
            lo = th
        else:
            hi = th
    return np.maximum(q - 0.5 * (lo + hi), 0)


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6])
def test_simplex_projection(K):
    rng = np.random.default_rng(K)
    for _ in range(300):
        q = rng.normal(0, 1.5, K)
        p = O.proj_simplex(q)
        assert np.all(p >= 0) and abs(p.sum() - 1) < 1e-12
        # KKT: q - p is constant on the support and <= that constant off it
        s = p > 0
        t = (q - p)[s]
        assert np.ptp(t) < 1e-12
        assert np.all((q - p)[~s] <= t[0] + 1e-12)
        assert np.allclose(p, _simplex_bisect(q), atol=1e-12)
        pf = O.proj_simplex(q.astype(np.float32), "f32")
        assert np.allclose(pf, p, atol=1e-6)
    # ties
    p = O.proj_simplex(np.full(K, 0.3))
    assert np.allclose(p, 1.0 / K)


def test_dykstra_k2_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(500):
        q = rng.normal(0, 1.0, (2, 3))
        p, sweeps = O.dykstra(q)
        mid = 0.5 * (q[0] + q[1])
        d = q[1] - q[0]
        d = d * min(1.0, 1.0 / np.linalg.norm(d))
        assert np.allclose(p[0], mid - 0.5 * d, atol=1e-12)
        assert np.allclose(p[1], mid + 0.5 * d, atol=1e-12)
        assert sweeps <= 2


def _feasible(p, tol):
    K = p.shape[0]
    return all(np.linalg.norm(p[i] - p[j]) <= 1 + tol for i in range(K) for j in range(i + 1, K))


@pytest.mark.parametrize("K", [3, 4, 5, 6])
def test_dykstra_limit_is_the_projection(K):
    rng = np.random.default_rng(10 + K)
    for _ in range(40):
        q = rng.normal(0, 1.2, (K, 3))
        lim, _ = O.dykstra(q, max_sweeps=10000, tol=0.0)
        assert _feasible(lim, 1e-9)
        # variational inequality <q - P q, y - P q> <= 0 for feasible y
        for _ in range(20):
            y, _ = O.dykstra(rng.normal(0, 1.2, (K, 3)), max_sweeps=10000, tol=0.0)
            assert np.sum((q - lim) * (y - lim)) <= 1e-7
        # the solver's 20-sweep / 1e-6 form stays close to the limit
        p20, sweeps = O.dykstra(q)
        assert 1 <= sweeps <= 20
        assert np.abs(p20 - lim).max() < 0.05
        # already-feasible input: one sweep, unchanged
        p1, s1 = O.dykstra(lim * 0.9)
        assert s1 == 1 and np.array_equal(p1, lim * 0.9)


def _potts_numpy(lay, u, c, lam, tau1, tau2, iters):
    """Two-phase iteration recomposed from the pinned operators."""
    K = len(c)
    m = _valid(lay)
    cc = np.asarray(c, np.float64)
    uu = u.astype(np.float64)
    lab = np.argmin(np.abs(cc[None, :] - uu[:, None]), axis=1)
    V = np.zeros((K - 1, lay.span))
    for i in range(K - 1):
        V[i] = np.where(m & (lab == i), 1.0, 0.0)
    Vb = V.copy()
    xi = np.zeros((K - 1, 3, lay.span))
    for _ in range(iters):
        vfull = np.vstack([Vb, 1.0 - Vb.sum(0)])
        grads = [O.grad(lay, np.where(m, vfull[i], 0.0)) for i in range(K)]
        xk = -xi.sum(0)
        new = np.zeros_like(xi)
        for o in np.nonzero(m)[0]:
            q = np.zeros((K, 3))
            for i in range(K):
                x = xi[i, :, o] if i < K - 1 else xk[:, o]
                q[i] = [tau2 * grads[i][a][o] + x[a] for a in range(3)]
            p, _ = O.dykstra(q)
            new[:, :, o] = p[: K - 1]
        xi = new
        xfull = np.concatenate([xi, -xi.sum(0, keepdims=True)])
        divs = [O.div(lay, [np.where(m, xfull[i, a], 0.0) for a in range(3)]) for i in range(K)]
        vf = np.vstack([V, 1.0 - V.sum(0)])
        nV = np.zeros_like(V)
        for o in np.nonzero(m)[0]:
            q = np.array([vf[i, o] + tau1 * (divs[i][o] - lam * (uu[o] - cc[i]) ** 2) for i in range(K)])
            nV[:, o] = O.proj_simplex(q)[: K - 1]
        Vb = 2 * nV - V
        V = nV
    return V, xi.reshape(3 * (K - 1), lay.span)


@pytest.mark.parametrize("K", [2, 3, 4])
def test_potts_iteration_recomposed(K):
    lay = O.Layout(6, 7, 8, 3.5, 3.0, 3.6)
    rng = np.random.default_rng(K)
    c = np.linspace(0.2, 0.8, K).astype(np.float32)
    u = np.where(_valid(lay), rng.uniform(0, 1, lay.span), 0).astype(np.float32)
    tau = float(np.float32(1 / np.sqrt(12)))
    for iters in (1, 3):
        Vn, xin = _potts_numpy(lay, u, c, 2.0, tau, tau, iters)
        r = O.potts(lay, u, c, 2.0, tau, tau, iters, precision="f64", want_state=True)
        m = _valid(lay)
        assert np.abs(r["V"][:, m] - Vn[:, m]).max() < 1e-12
        assert np.abs(r["xi"][:, m] - xin[:, m]).max() < 1e-12
