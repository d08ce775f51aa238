"""The bimm-peaks oracle (oracle/bimm.py) pinned without running the reference:

  - against the reference's own fitted marginals (bimm-peaks/pdf-0[67]-*.txt,
    kept in tests/golden/bimm_pdfs.npz): each pure phase is a Gaussian in u
    and all phases of a fit share one deviation (puv.h:132-138); each
    interface marginal is reproduced by marginal_p_ij with the phases'
    intensities and deviation and two free numbers, its weight and ds, to the
    files' precision (puv.h:140-180); the full marginal is the sum of its
    parts (puv.h:182-202);
  - by normalisation: every component density integrates to 1 over u and
    v > 0, and full_p_ij integrated over v is marginal_p_ij;
  - by the box masses M_i and M_ij against an independent tensor-product
    quadrature of the densities over [Lu, Hu] x (0, Hv]."""
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import optimize

from oracle import bimm as B

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "bimm_pdfs.npz")


def _gauss_fit(xy):
    """ln pdf = c0 + c1 x + c2 x^2 by least squares: weight, mean, sigma."""
    x, p = xy[:, 0], xy[:, 1]
    keep = p > 1e-6 * p.max()
    c2, c1, c0 = np.polyfit(x[keep], np.log(p[keep]), 2)
    s2 = -1.0 / (2 * c2)
    mu = c1 * s2
    w = math.exp(c0 + mu * mu / (2 * s2)) * math.sqrt(2 * math.pi * s2)
    return w, mu, math.sqrt(s2)


@pytest.mark.parametrize("run", ["06", "07"])
def test_pure_phases_are_gaussians_with_one_deviation(run):
    fits = []
    for k in ("00", "01", "02"):
        xy = GOLD[f"{run}_{k}"]
        w, mu, s = _gauss_fit(xy)
        model = w / (math.sqrt(2 * math.pi) * s) * np.exp(-(xy[:, 0] - mu) ** 2 / (2 * s * s))
        keep = xy[:, 1] > 1e-6 * xy[:, 1].max()
        # the files print x to 8-9 digits: about 5e-10, which moves the
        # density by |x - mu| / s^2 * 5e-10 < 1e-5 where it is above 1e-6 of its peak
        assert np.max(np.abs(model[keep] / xy[keep, 1] - 1)) < 1e-5, (run, k)
        fits.append((w, mu, s))
    sig = [f[2] for f in fits]
    assert max(sig) / min(sig) - 1 < 1e-5, sig


def _interface_fit(run, i, j):
    """marginal_p_ij in raw units: with u = (x - lo) / (hi - lo) the density
    w_ij marginal_p_ij / (hi - lo) is the same function of x with I and sn
    in raw units, so lo and hi drop out.  Returns (w_ij, ds, max rel. error)."""
    pure = [_gauss_fit(GOLD[f"{run}_0{k}"]) for k in range(3)]
    Ii, Ij, sn = pure[i][1], pure[j][1], pure[0][2]
    xy = GOLD[f"{run}_0{i}_0{j}"]
    keep = xy[:, 1] > 1e-3 * xy[:, 1].max()
    x, p = xy[keep, 0], xy[keep, 1]
    idx = np.linspace(0, x.size - 1, 40).astype(int)

    def model(params, xs):
        w, ds = math.exp(params[0]), math.exp(params[1])
        return np.array([w * B.marginal_p_ij(float(u), Ii, Ij, sn, ds) for u in xs])

    def resid(params):
        return model(params, x[idx]) / p[idx] - 1

    sol = optimize.least_squares(resid, [math.log(0.01), math.log(1.0)], xtol=1e-14, ftol=1e-14, gtol=1e-14)
    return math.exp(sol.x[0]), math.exp(sol.x[1]), float(np.max(np.abs(model(sol.x, x) / p - 1)))


@pytest.mark.parametrize("run", ["06", "07"])
def test_interface_marginals_match_reference(run):
    """Each interface file is reproduced to its print precision; the three
    pairs of a fit share one ds (one model parameter), and with the pure
    phases' weights all six weights sum to 1 (a softmax)."""
    fits = [_interface_fit(run, i, j) for i, j in ((0, 1), (0, 2), (1, 2))]
    for (w, ds, err), pair in zip(fits, ((0, 1), (0, 2), (1, 2))):
        assert err < 2e-5, (run, pair, err, w, ds)
    dss = [f[1] for f in fits]
    assert max(dss) / min(dss) - 1 < 1e-5, dss
    total = sum(_gauss_fit(GOLD[f"{run}_0{k}"])[0] for k in range(3)) + sum(f[0] for f in fits)
    assert abs(total - 1) < 1e-4, total


@pytest.mark.parametrize("run", ["06", "07"])
def test_full_marginal_is_the_sum_of_its_parts(run):
    full = GOLD[f"{run}_full"]
    parts = sum(GOLD[f"{run}_{k}"][:, 1] for k in ("00", "01", "02", "00_01", "00_02", "01_02"))
    assert np.allclose(GOLD[f"{run}_00"][:, 0], full[:, 0])
    assert np.max(np.abs(parts / full[:, 1] - 1)) < 1e-12


# A model whose densities are smooth enough for plain tensor quadrature.
I3 = (0.2, 0.5, 0.8)
SN, SB, DS, RHO = 0.05, 1.0, 1.2, 0.3


def _gl(a, b, n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def _full_p_ij_grid(u, v, t, Ii, Ij, sn, sb, ds, rho):
    """full_p_ij on a (u, v, t) grid, summed over t with the given weights
    outside: the integrand of puv.h:70-91 written independently in numpy."""
    a = sn * sn * (1 - rho)
    from scipy.special import erf
    I = Ii + 0.5 * (Ij - Ii) * (1 + erf(t / math.sqrt(2)))
    G = abs(Ij - Ii) / sb * np.exp(-t * t / 2) / math.sqrt(2 * math.pi)
    U, V, T = np.meshgrid(u, v, t, indexing="ij")
    II, GG = I[None, None, :], G[None, None, :]
    f = np.exp(-0.5 * (U - II) ** 2 / (sn * sn) - (V - GG) ** 2 / a) * (-np.expm1(-4 * V * GG / a)) / GG
    return V / (2 * math.sqrt(2) * math.pi * ds * sn * math.sqrt(a)) * f


def test_component_densities_are_normalised():
    for Ii in I3:
        assert abs(B.marginal_p_i(Ii, Ii, SN) * SN * math.sqrt(2 * math.pi) - 1) < 1e-15
    u, wu = _gl(-0.5, 1.5, 400)
    assert abs(sum(B.marginal_p_ij(float(x), 0.2, 0.8, SN, DS) * w for x, w in zip(u, wu)) - 1) < 1e-9
    v, wv = _gl(0.0, 1.0, 300)
    m = sum(B.full_p_i(float(x), float(y), 0.5, math.log(SN), RHO) * a * b
            for x, a in zip(u[::4], wu[::4] * 4) for y, b in zip(v, wv))
    assert abs(m - 1) < 1e-6
    t, wt = _gl(-DS, DS, 64)
    g = _full_p_ij_grid(u[::4], v, t, 0.2, 0.8, SN, SB, DS, RHO)
    total = np.einsum("ijk,i,j,k->", g, wu[::4] * 4, wv, wt)
    assert abs(total - 1) < 1e-5


def test_full_p_ij_over_v_is_its_u_marginal():
    v, wv = _gl(0.0, 1.5, 400)
    for u in (0.25, 0.5, 0.7):
        s = sum(B.full_p_ij(u, float(y), 0.2, 0.8, math.log(SN), math.log(SB), math.log(DS), RHO) * b
                for y, b in zip(v, wv))
        assert abs(s / B.marginal_p_ij(u, 0.2, 0.8, SN, DS) - 1) < 1e-7


def test_box_masses():
    Lu, Hu, Hv = 0.15, 0.75, 0.06
    u, wu = _gl(Lu, Hu, 120)
    v, wv = _gl(0.0, Hv, 120)
    for Ii in I3:
        m = sum(B.full_p_i(float(x), float(y), Ii, math.log(SN), RHO) * a * b
                for x, a in zip(u, wu) for y, b in zip(v, wv))
        assert abs(B.M_i(Lu, Hu, Hv, Ii, SN, RHO) - m) < 1e-9
    t, wt = _gl(-DS, DS, 96)
    g = _full_p_ij_grid(u, v, t, 0.2, 0.8, SN, SB, DS, RHO)
    m = np.einsum("ijk,i,j,k->", g, wu, wv, wt)
    assert abs(B.M_ij(Lu, Hu, Hv, 0.2, 0.8, SN, SB, DS, RHO) / m - 1) < 1e-6


def test_pair_order_is_strict_upper():
    """layout_strict_upper.h:24-26: index (2n - 1 - i) i / 2 + j - i - 1."""
    for K in range(2, 7):
        for k, (i, j) in enumerate(B.pairs(K)):
            assert k == (2 * K - 1 - i) * i // 2 + j - i - 1
