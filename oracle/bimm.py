"""FP64 oracle of bimm-peaks' bivariate intensity / gradient-magnitude mixture
(bimm-peaks/puv.h, bimm-peaks.cpp:395-478).  Test infrastructure only: the
tests compare the GPU likelihood and the host model against it.

A mixture of K pure phases and K(K-1)/2 interfaces over (u, v), u the
normalised intensity and v the gradient magnitude:
  full_p_i   pure phase i: u ~ N(I_i, sn^2), v ~ a Maxwell-like density of
             scale a = sn^2 (1 - rho)                           (puv.h:42-50)
  full_p_ij  interface i|j: u ~ N(I(t), sn^2) with I(t) = I_i + (I_j - I_i)
             Phi(t), v noncentral around G(t) = |I_j - I_i| phi(t) / sb, t
             uniform on [-ds, ds]                                (puv.h:52-110)
  marginal_p_i, marginal_p_ij   the u marginals                (puv.h:132-202)
  M_i, M_ij  the mass of each component in [Lu, Hu] x (0, Hv]  (puv.h:204-285)
The integrals over t use scipy.integrate.quad at tolerances far below the
reference's (epsabs 1e-10, epsrel 1e-6 with GSL QAGS), with breakpoints where
the integrand can peak (I(t) = u, G(t) = v) so that narrow peaks are not
missed.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate
from scipy.special import erf

SQRT2 = math.sqrt(2.0)
SQRTPI = math.sqrt(math.pi)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)
EPSABS, EPSREL = 1e-14, 1e-12


def _quad(f, a, b, points=None):
    pts = sorted(x for x in (points or ()) if a < x < b)
    return integrate.quad(f, a, b, epsabs=EPSABS, epsrel=EPSREL, limit=2000, points=pts or None)[0]


def _peaks(u, v, Ii, Ij, sb):
    """Where the interface integrand can peak in t: I(t) = u and G(t) = v."""
    from scipy.special import erfinv
    out = []
    dI = Ij - Ii
    if dI != 0:
        s = (u - Ii) / dI
        if 0 < s < 1:
            out.append(SQRT2 * float(erfinv(2 * s - 1)))
    G0 = abs(dI) * INV_SQRT_2PI / sb
    if v is not None and 0 < v < G0:
        t = math.sqrt(2 * math.log(G0 / v))
        out += [t, -t]
    return out


def interface_intensity(t, Ii, Ij):
    """I(t) = I_i + (I_j - I_i) (1 + erf(t / sqrt 2)) / 2."""
    return Ii + 0.5 * (Ij - Ii) * (1.0 + erf(t / SQRT2))


def full_p_i(u, v, Ii, log_sn, rho):
    """puv.h:42-50."""
    sn2 = math.exp(2.0 * log_sn)
    return (2.0 * SQRT2 / math.pi) * v * v / (sn2 * sn2 * (1.0 - rho) ** 1.5) * \
        math.exp(-0.5 * (u - Ii) ** 2 / sn2 - v * v / sn2 / (1.0 - rho))


def full_p_ij(u, v, Ii, Ij, log_sn, log_sb, log_ds, rho):
    """puv.h:52-110."""
    sn2, sn, sb, ds = math.exp(2 * log_sn), math.exp(log_sn), math.exp(log_sb), math.exp(log_ds)
    a = sn2 * (1.0 - rho)
    dI = Ij - Ii

    def f(t):
        I = interface_intensity(t, Ii, Ij)
        if abs(t) < 30 and abs(dI) > 1e-9:
            G = INV_SQRT_2PI * abs(dI) / sb * math.exp(-t * t / 2)
            return math.exp(-0.5 * (u - I) ** 2 / sn2 - (v - G) ** 2 / a) * (-math.expm1(-4 * v * G / a)) / G
        return math.exp(-0.5 * (u - I) ** 2 / sn2 - v * v / a) * 4 * v / a

    return v / (2.0 * SQRT2 * math.pi * ds * sn * math.sqrt(a)) * _quad(f, -ds, ds, _peaks(u, v, Ii, Ij, sb))


def pairs(K):
    """(i, j), i < j, in the strict-upper order of layout_strict_upper.h:24-26."""
    return [(i, j) for i in range(K - 1) for j in range(i + 1, K)]


def full_p(u, v, w, I, log_sn, log_sb, log_ds, rho):
    """puv.h:112-130: w[0..K-1] pure phases, w[K..] the pairs in order."""
    K = len(I)
    p = sum(w[i] * full_p_i(u, v, I[i], log_sn, rho) for i in range(K))
    for k, (i, j) in enumerate(pairs(K)):
        p += w[K + k] * full_p_ij(u, v, I[i], I[j], log_sn, log_sb, log_ds, rho)
    return p


def marginal_p_i(u, Ii, sn):
    """puv.h:132-138."""
    return INV_SQRT_2PI / sn * math.exp(-(u - Ii) ** 2 / (2 * sn * sn))


def marginal_p_ij(u, Ii, Ij, sn, ds):
    """puv.h:140-180."""
    return INV_SQRT_2PI / (2 * ds * sn) * _quad(
        lambda t: math.exp(-(u - interface_intensity(t, Ii, Ij)) ** 2 / (2 * sn * sn)), -ds, ds,
        _peaks(u, None, Ii, Ij, 1.0))


def marginal_p(u, w, I, sn, ds):
    """puv.h:182-202."""
    K = len(I)
    p = sum(w[i] * marginal_p_i(u, I[i], sn) for i in range(K))
    for k, (i, j) in enumerate(pairs(K)):
        p += w[K + k] * marginal_p_ij(u, I[i], I[j], sn, ds)
    return p


def M_i(Lu, Hu, Hv, Ii, sn, rho):
    """puv.h:204-218: mass of full_p_i in [Lu, Hu] x (0, Hv]."""
    a = sn * sn * (1.0 - rho)
    sa = math.sqrt(a)
    return (erf((Hu - Ii) / (SQRT2 * sn)) - erf((Lu - Ii) / (SQRT2 * sn))) * \
        (0.5 * erf(Hv / sa) - Hv * math.exp(-Hv * Hv / a) / sa / SQRTPI)


def M_ij(Lu, Hu, Hv, Ii, Ij, sn, sb, ds, rho):
    """puv.h:220-285: mass of full_p_ij in [Lu, Hu] x (0, Hv]."""
    a = sn * sn * (1.0 - rho)
    sa = math.sqrt(a)
    h = Hv / sa

    def f(t):
        I = interface_intensity(t, Ii, Ij)
        G = abs(Ij - Ii) / sb * math.exp(-t * t / 2) / SQRTPI / SQRT2
        g = G / sa
        z = 4 * g * h
        ratio = 1.0 if z == 0 else -math.expm1(-z) / z
        mass2 = erf(g + h) - erf(g - h) - 4 * h / SQRTPI * math.exp(-(g - h) ** 2) * ratio
        return mass2 * (erf((Hu - I) / (SQRT2 * sn)) - erf((Lu - I) / (SQRT2 * sn)))

    return _quad(f, -ds, ds) / (8 * ds)


def nll_sum(u, v, w, I, log_sn, log_sb, log_ds, rho):
    """sum over the samples of -log full_p (bimm-peaks.cpp:458-472)."""
    return float(sum(-math.log(full_p(float(a), float(b), w, I, log_sn, log_sb, log_ds, rho)) for a, b in zip(u, v)))


def normaliser(Lu, Hu, Hv, w, I, sn, sb, ds, rho):
    """Z of bimm-peaks.cpp:427-456."""
    K = len(I)
    Z = sum(w[i] * M_i(Lu, Hu, Hv, I[i], sn, rho) for i in range(K))
    for k, (i, j) in enumerate(pairs(K)):
        Z += w[K + k] * M_ij(Lu, Hu, Hv, I[i], I[j], sn, sb, ds, rho)
    return Z
