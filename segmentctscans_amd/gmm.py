"""One-dimensional Gaussian mixture fit of gmm-peaks (gmm-peaks/gmm.h:52-285)
with its sample passes on the GPU (sct_gmm1d_pass / sct_gmm1d_assign).

The EM iteration, the rescue of lost or degenerate components and the random
restarts follow gmm.h; the restarts draw from numpy's PCG64 with a caller's
seed where gmm.h seeds std::mt19937 from std::random_device, so runs repeat.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import Context
from . import _native as N

_dp = C.POINTER(C.c_double)


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Samples:
    """Samples resident on the device (doubles)."""

    def __init__(self, ctx: Context, x):
        self.ctx = ctx
        self.x = np.ascontiguousarray(x, np.float64)
        h = C.c_void_p()
        ctx.check(N.lib().sct_samples_create(ctx.handle, _d(self.x), self.x.size, C.byref(h)))
        self.handle = h

    @property
    def n(self) -> int:
        return self.x.size

    def em_pass(self, pi, mu, var):
        """(w, s1, s2, nll) of the fused E+M pass, gmm.h:145-201."""
        K = len(mu)
        pi, mu, var = (np.ascontiguousarray(a, np.float64) for a in (pi, mu, var))
        w, s1, s2 = np.zeros(K), np.zeros(K), np.zeros(K)
        nll = C.c_double()
        self.ctx.check(N.lib().sct_gmm1d_pass(self.handle, K, _d(pi), _d(mu), _d(var), _d(w), _d(s1), _d(s2),
                                              C.byref(nll)))
        return w, s1, s2, nll.value

    def assign(self, pi, mu, var) -> np.ndarray:
        K = len(mu)
        pi, mu, var = (np.ascontiguousarray(a, np.float64) for a in (pi, mu, var))
        cl = np.zeros(self.n, np.int32)
        self.ctx.check(N.lib().sct_gmm1d_assign(self.handle, K, _d(pi), _d(mu), _d(var),
                                                cl.ctypes.data_as(C.POINTER(C.c_int32))))
        return cl

    def close(self) -> None:
        if self.handle:
            N.lib().sct_samples_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gmm(samples: Samples, init_centers, init_clusters, attempts: int = 10, progress=None, *, seed: int = 0,
        n_iters: int = 4000):
    """gmm(x, init_centers, init_clusters, attempts, progress), gmm.h:52-285,
    for one-dimensional samples.  Returns (nll, var[K], mu[K], clusters) of the
    attempt with the smallest negative log-likelihood."""
    x = samples.x
    K = len(init_centers)
    n = x.size
    mu = np.array(init_centers, np.float64)
    cl0 = np.asarray(init_clusters)
    rng = np.random.default_rng(seed)
    span = float(x.max() - x.min()) if n else 0.0

    var = np.zeros(K)
    dets = np.zeros(K)

    def rescue(k):
        # half the bounding box as the deviation (gmm.h:101-106)
        var[k] = (span / 2.0) ** 2
        dets[k] = math.sqrt(abs(var[k]))

    pi = np.full(K, 1.0 / K)
    counts = np.bincount(cl0, minlength=K)
    sq = np.bincount(cl0, weights=(x - mu[cl0]) ** 2, minlength=K)
    for k in range(K):
        if counts[k]:
            var[k] = sq[k] / counts[k]
            dets[k] = math.sqrt(abs(var[k]))
        else:
            rescue(k)
    if progress:
        progress(-1, -1, 0.0, var, mu, cl0)

    best = None
    for attempt in range(attempts):
        nll = prev = 0.0
        for it in range(n_iters):
            w, s1, s2, nll = samples.em_pass(pi, mu, var)
            if progress:
                progress(attempt, it, nll, var, mu, None)
            # M step from the raw moments (gmm.h:240-251)
            for k in range(K):
                if pi.min() * 1e-9 < w[k] and 0 < w[k]:
                    mu[k] = s1[k] / w[k]
                    var[k] = s2[k] / w[k] - mu[k] * mu[k]
                    dets[k] = math.sqrt(abs(var[k]))
                else:
                    rescue(k)
                pi[k] = w[k] / n
            pi /= pi.sum()
            # degenerate components (gmm.h:253-263)
            for k in range(K):
                if dets[k] < 1e-12:
                    rescue(k)
                    pi *= 1.0 - 1.0 / K
                    pi[k] = 1.0 / K
                    pi /= pi.sum()
            if it > 0 and abs(prev - nll) / abs(nll) < 1e-9:
                break
            prev = nll
        if attempt == 0 or nll < best[0]:
            best = (nll, var.copy(), mu.copy(), samples.assign(pi, mu, var))
        # random perturbation of the centres (gmm.h:274-283); a 1 x 1
        # covariance keeps its one eigenvalue
        mu += rng.uniform(-span / 4, span / 4, K)
        dets[:] = np.sqrt(np.abs(var))
    return best
