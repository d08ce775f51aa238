"""The bivariate intensity / gradient-magnitude mixture fit of bimm-peaks
(bimm-peaks/bimm-peaks.cpp:80-172, :395-560, bimm-peaks/puv.h).

A sample (u, v) is an in-mask voxel's normalised intensity u and gradient
magnitude v.  The model mixes K pure phases (intensity I_i) and the K(K-1)/2
interfaces between them (intensity ramping from I_i to I_j across a blurred
edge), with a common noise deviation sn, blur sb, interface half-width ds and
noise correlation rho.  Its parameters are fitted by minimising

    cost = mean over samples of -log full_p(u, v) + log Z,

Z the model's mass in the sampled box [lo_u, hi_u] x (0, hi_v], with a
Nelder-Mead simplex over unconstrained parameters mapped into their ranges by
sigmoids (weights by a softmax with the first fixed at 0).  The sample sums
run on the GPU (sct_bimm_nll); Z and the simplex are host logic.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
from scipy import integrate
from scipy.special import erf

from . import _native as N
from . import Context

SQRT2, SQRTPI = math.sqrt(2.0), math.sqrt(math.pi)

# parameter ranges of bimm-peaks.cpp:404-409 (weights before the softmax;
# sn, sb and ds as logarithms)
W_RANGE = (-4.6, 4.6)
I_RANGE = (0.0, 1.0)
LOG_SN_RANGE = (math.log(1e-3), math.log(2.0))
LOG_SB_RANGE = (math.log(0.1), math.log(10.0))
RHO_RANGE = (0.0, 0.999)


def pairs(K: int):
    """Interface pairs (i, j), i < j, in strict-upper order."""
    return [(i, j) for i in range(K - 1) for j in range(i + 1, K)]


class UVSamples:
    """(u, v) samples resident on the device."""

    def __init__(self, ctx: Context, u, v):
        self.ctx = ctx
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        if u.shape != v.shape:
            raise ValueError("u and v differ in length")
        h = C.c_void_p()
        ctx.check(N.lib().sct_uv_samples_create(ctx.handle, u.ctypes.data_as(N.f64p), v.ctypes.data_as(N.f64p),
                                                u.size, C.byref(h)))
        self.handle, self.n = h, u.size

    def nll_sum(self, w, I, log_sn, log_sb, log_ds, rho, first: int = 0, count: int | None = None) -> float:
        """sum over samples [first, first + count) of -log full_p."""
        w = np.ascontiguousarray(w, np.float64)
        I = np.ascontiguousarray(I, np.float64)
        count = self.n - first if count is None else count
        out = C.c_double()
        self.ctx.check(N.lib().sct_bimm_nll(self.handle, I.size, w.ctypes.data_as(N.f64p), I.ctypes.data_as(N.f64p),
                                            log_sn, log_sb, log_ds, rho, first, count, C.byref(out)))
        return out.value

    def close(self) -> None:
        if self.handle:
            N.lib().sct_uv_samples_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sigm(x, lo, hi):
    return (hi - lo) / (1.0 + math.exp(-x)) + lo


def _inv_sigm(y, lo, hi):
    return -math.log((hi - lo) / (y - lo) - 1.0)


@dataclass
class Params:
    """Model parameters in their own ranges (w over all K + K(K-1)/2 components)."""
    w: np.ndarray
    I: np.ndarray
    log_sn: float
    log_sb: float
    log_ds: float
    rho: float


class Model:
    """Maps the simplex's unconstrained vector x (bimm-peaks.cpp:397-413):
    x[0 .. nw) weights (nw = K(K+1)/2 - 1, the first weight fixed at 0),
    x[nw .. nw + K) intensities, then log sn, log sb, rho, log ds."""

    def __init__(self, K: int, max_ds: float = 1.5):
        self.K = K
        self.n_weights = K * (K + 1) // 2 - 1
        self.i0 = self.n_weights
        self.n_params = K * (K + 3) // 2 + 3
        self.ds_range = (math.log(0.5), math.log(max_ds))

    def restrict(self, x) -> Params:
        K, nw, i0 = self.K, self.n_weights, self.i0
        a = np.concatenate([[0.0], [_sigm(x[k], *W_RANGE) for k in range(nw)]])
        e = np.exp(a)
        return Params(w=e / e.sum(), I=np.array([_sigm(x[i0 + i], *I_RANGE) for i in range(K)]),
                      log_sn=_sigm(x[i0 + K], *LOG_SN_RANGE), log_sb=_sigm(x[i0 + K + 1], *LOG_SB_RANGE),
                      rho=_sigm(x[i0 + K + 2], *RHO_RANGE), log_ds=_sigm(x[i0 + K + 3], *self.ds_range))

    def initial(self, peaks) -> np.ndarray:
        """bimm-peaks.cpp:519-525: intensities at the given peaks, interface
        weights at the bottom of their range, the rest mid-range."""
        x = np.zeros(self.n_params)
        for i, p in enumerate(peaks):
            x[self.i0 + i] = _inv_sigm(p, *I_RANGE)
        for k in range(self.K - 1, self.n_weights):
            x[k] = _inv_sigm(W_RANGE[0] * 0.99, *W_RANGE)
        return x


def mass_pure(Lu, Hu, Hv, Ii, sn, rho):
    """Mass of pure phase i in [Lu, Hu] x (0, Hv]: the normal's mass in u times
    the Maxwell-like mass in v (puv.h:204-218)."""
    a = sn * sn * (1.0 - rho)
    sa = math.sqrt(a)
    h = Hv / sa
    return (erf((Hu - Ii) / (SQRT2 * sn)) - erf((Lu - Ii) / (SQRT2 * sn))) * \
        (0.5 * erf(h) - h * math.exp(-h * h) / SQRTPI)


def mass_interface(Lu, Hu, Hv, Ii, Ij, sn, sb, ds, rho):
    """Mass of interface i|j in the box (puv.h:220-285): over t in [-ds, ds],
    the noncentral v mass below Hv times the normal u mass in [Lu, Hu]."""
    a = sn * sn * (1.0 - rho)
    sa = math.sqrt(a)
    h = Hv / sa
    dI = Ij - Ii

    def f(t):
        I = Ii + 0.5 * dI * (1.0 + math.erf(t / SQRT2))
        g = abs(dI) / sb * math.exp(-0.5 * t * t) / (SQRTPI * SQRT2) / sa
        z = 4.0 * g * h
        ratio = 1.0 if z == 0 else -math.expm1(-z) / z
        v_mass = math.erf(g + h) - math.erf(g - h) - 4.0 * h / SQRTPI * math.exp(-(g - h) ** 2) * ratio
        return v_mass * (math.erf((Hu - I) / (SQRT2 * sn)) - math.erf((Lu - I) / (SQRT2 * sn)))

    return integrate.quad(f, -ds, ds, epsabs=1e-12, epsrel=1e-10, limit=500)[0] / (8.0 * ds)


def normaliser(box, p: Params) -> float:
    """Z: the model's mass in the sampled box (bimm-peaks.cpp:427-456)."""
    Lu, Hu, Hv = box
    K = p.I.size
    sn, sb, ds = math.exp(p.log_sn), math.exp(p.log_sb), math.exp(p.log_ds)
    Z = sum(p.w[i] * mass_pure(Lu, Hu, Hv, p.I[i], sn, p.rho) for i in range(K))
    for k, (i, j) in enumerate(pairs(K)):
        Z += p.w[K + k] * mass_interface(Lu, Hu, Hv, p.I[i], p.I[j], sn, sb, ds, p.rho)
    return Z


def marginal_pure(u, Ii, sn):
    """u marginal of a pure phase (puv.h:132-138)."""
    return math.exp(-(u - Ii) ** 2 / (2 * sn * sn)) / (math.sqrt(2 * math.pi) * sn)


def marginal_interface(u, Ii, Ij, sn, ds):
    """u marginal of an interface (puv.h:140-180); the integrand peaks where
    I(t) = u, which is passed to the integrator as a breakpoint."""
    from scipy.special import erfinv
    f = lambda t: math.exp(-(u - (Ii + 0.5 * (Ij - Ii) * (1 + math.erf(t / SQRT2)))) ** 2 / (2 * sn * sn))
    pts = None
    if Ij != Ii and 0 < (u - Ii) / (Ij - Ii) < 1:
        t = SQRT2 * float(erfinv(2 * (u - Ii) / (Ij - Ii) - 1))
        pts = [t] if -ds < t < ds else None
    return integrate.quad(f, -ds, ds, epsabs=1e-13, epsrel=1e-10, limit=500, points=pts)[0] / \
        (2 * ds * sn * math.sqrt(2 * math.pi))


def nelder_mead(f, x0, steps, max_iter: int, stop_size: float, progress=None):
    """Minimise f from the simplex x0, x0 + steps[k] e_k (Nelder and Mead,
    Computer Journal 7, 1965: reflection 1, expansion 2, contraction 1/2,
    shrink 1/2).  The size is the mean distance of the vertices from their
    centroid; the iteration ends when it falls below stop_size.  progress(it,
    fbest, size, xbest) returning True restarts the simplex around the best
    vertex with the steps it returns (an array) -- the reference's restarts
    with shrinking steps and a new sample batch (bimm-peaks.cpp:148-165).
    Returns (x, f(x), iterations)."""
    def simplex(x, st):
        pts = [np.array(x, float)]
        for k in range(len(x)):
            y = np.array(x, float)
            y[k] += st[k]
            pts.append(y)
        return pts, [f(p) for p in pts]

    pts, vals = simplex(x0, steps)
    it = 0
    for it in range(1, max_iter + 1):
        order = np.argsort(vals)
        pts = [pts[k] for k in order]
        vals = [vals[k] for k in order]
        cen = np.mean(pts[:-1], axis=0)
        xr = cen + (cen - pts[-1])
        fr = f(xr)
        if fr < vals[0]:
            xe = cen + 2.0 * (cen - pts[-1])
            fe = f(xe)
            pts[-1], vals[-1] = (xe, fe) if fe < fr else (xr, fr)
        elif fr < vals[-2]:
            pts[-1], vals[-1] = xr, fr
        else:
            xc = cen + 0.5 * (pts[-1] - cen) if fr >= vals[-1] else cen + 0.5 * (xr - cen)
            fc = f(xc)
            if fc < min(fr, vals[-1]):
                pts[-1], vals[-1] = xc, fc
            else:
                pts = [pts[0]] + [pts[0] + 0.5 * (p - pts[0]) for p in pts[1:]]
                vals = [vals[0]] + [f(p) for p in pts[1:]]
        best = int(np.argmin(vals))
        allc = np.mean(pts, axis=0)
        size = float(np.mean([np.linalg.norm(p - allc) for p in pts]))
        if progress is not None:
            restart = progress(it, vals[best], size, pts[best])
            if restart is not None and restart is not False:
                pts, vals = simplex(pts[best], restart)
                continue
        if size < stop_size:
            break
    best = int(np.argmin(vals))
    return pts[best], vals[best], it


@dataclass
class FitResult:
    params: Params
    cost: float
    iterations: tuple


def fit(samples: UVSamples, box, peaks, *, max_ds: float = 1.5, batch: int = 10_000, batches: int = 100,
        restart_period: int = 100, full_iters: int = 10_000, log=None) -> FitResult:
    """bimm-peaks.cpp:527-560: a simplex over mini-batches of `batch`
    samples (a new batch and steps x 0.95 every restart_period iterations,
    restart_period x batches iterations in all, stop size 1e-5), then over
    all batch x batches samples (up to full_iters iterations, stop size
    1e-4).  box = (lo_u, hi_u, hi_v); peaks: initial intensities."""
    K = len(peaks)
    model = Model(K, max_ds)
    n_all = batch * batches
    if samples.n < n_all:
        raise ValueError(f"{samples.n} samples, {n_all} needed")
    state = {"batch": 0}

    def cost(x, first, count):
        p = model.restrict(x)
        Z = normaliser(box, p)
        if not Z > 0:
            return math.inf
        return samples.nll_sum(p.w, p.I, p.log_sn, p.log_sb, p.log_ds, p.rho, first, count) / count + math.log(Z)

    steps = np.ones(model.n_params)

    def progress_batches(it, fbest, size, x):
        if log:
            log(it, fbest, size, model.restrict(x))
        if it % restart_period == 0:
            state["batch"] = (state["batch"] + 1) % batches
            steps[:] *= 0.95
            return steps.copy()
        return None

    x, fx, it1 = nelder_mead(lambda x: cost(x, batch * state["batch"], batch), model.initial(peaks), steps.copy(),
                             restart_period * batches, 1e-5, progress_batches)
    x, fx, it2 = nelder_mead(lambda x: cost(x, 0, n_all), x, steps.copy(), full_iters, 1e-4,
                             (lambda it, fb, sz, xb: log(it, fb, sz, model.restrict(xb)) and None) if log else None)
    return FitResult(model.restrict(x), fx, (it1, it2))
