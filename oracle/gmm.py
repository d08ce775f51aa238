"""The sample passes of gmm-peaks' 1-D Gaussian mixture fit restated in numpy
(FP64) -- test infrastructure only: the checker of sct_gmm1d_pass /
sct_gmm1d_assign and the stand-in pass for testing the host EM loop on the CPU.

em_pass follows gmm-peaks/gmm.h:145-201 (densities :160-163, underflow
:170-175, moments :177-182); assign follows :204-219.  Pinned in
tests/test_oracle_gmm.py by closed forms (one component: the sample moments
and the Gaussian log-likelihood; identical components: equal shares), the
underflow rule and nearest-component assignment.
"""
import numpy as np

PITOD = (1.0 / (2.0 * np.pi)) ** 0.5


def densities(x, pi, mu, var):
    x = np.asarray(x, np.float64)[:, None]
    mu = np.asarray(mu, np.float64)[None, :]
    var = np.asarray(var, np.float64)
    a = x - mu
    return np.asarray(pi, np.float64) * PITOD * np.exp(-0.5 * (a * (1.0 / var) * a)) / np.sqrt(np.abs(var))


def em_pass(x, pi, mu, var):
    """(w, s1, s2, nll)."""
    x = np.asarray(x, np.float64)
    K = len(mu)
    d = densities(x, pi, mu, var)
    s = d.sum(axis=1)
    ok = s > 0
    nll = -np.log(s[ok]).sum() - np.count_nonzero(~ok) * np.log(np.finfo(np.float64).tiny)
    g = np.where(ok[:, None], d / np.where(ok, s, 1.0)[:, None], 1.0 / K)
    return g.sum(axis=0), (g * x[:, None]).sum(axis=0), (g * (x * x)[:, None]).sum(axis=0), nll


def assign(x, pi, mu, var):
    return np.argmax(densities(x, pi, mu, var), axis=1).astype(np.int32)


class HostSamples:
    """The interface of segmentctscans_amd.gmm.Samples over em_pass / assign,
    so the host EM loop can be exercised without a GPU."""

    def __init__(self, x):
        self.x = np.ascontiguousarray(x, np.float64)

    @property
    def n(self):
        return self.x.size

    def em_pass(self, pi, mu, var):
        return em_pass(self.x, pi, mu, var)

    def assign(self, pi, mu, var):
        return assign(self.x, pi, mu, var)
