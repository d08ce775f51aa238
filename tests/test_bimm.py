"""Host side of bimm-peaks (segmentctscans_amd/bimm.py, bimm_peaks.py): the
simplex, the parameter maps, the box masses against the oracle, the tail
cut and the peak syntax."""
import math

import numpy as np
import pytest

from oracle import bimm as OB
from segmentctscans_amd import bimm as B
from segmentctscans_amd.bimm_peaks import parse_peaks, tail_cutoffs


def test_nelder_mead_finds_the_rosenbrock_minimum():
    f = lambda x: (1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2
    x, fx, it = B.nelder_mead(f, [-1.2, 1.0], [0.5, 0.5], 5000, 1e-12)
    assert np.allclose(x, [1.0, 1.0], atol=1e-5) and fx < 1e-10 and it < 5000


def test_nelder_mead_quadratic_and_restarts():
    A = np.diag([1.0, 4.0, 9.0, 0.5])
    c = np.array([0.3, -1.0, 2.0, 0.7])
    f = lambda x: float((x - c) @ A @ (x - c))
    calls = []

    def progress(it, fb, size, xb):
        calls.append(it)
        return np.full(4, 0.1) if it in (20, 40) else None

    x, fx, _ = B.nelder_mead(f, np.zeros(4), np.ones(4), 4000, 1e-10, progress)
    assert np.allclose(x, c, atol=1e-5)
    assert calls[:3] == [1, 2, 3]


def test_parameter_maps():
    m = B.Model(3, max_ds=1.5)
    assert m.n_params == 3 * 6 // 2 + 3 and m.n_weights == 5
    peaks = [0.2, 0.45, 0.8]
    p = m.restrict(m.initial(peaks))
    assert np.allclose(p.I, peaks, atol=1e-12)
    assert abs(p.w.sum() - 1) < 1e-15 and p.w.size == 6
    assert np.all(p.w[3:] < 0.02 * p.w[0])            # interfaces start at e^-4.55 of the phases
    assert B.LOG_SN_RANGE[0] < p.log_sn < B.LOG_SN_RANGE[1]
    assert math.log(0.5) < p.log_ds < math.log(1.5)
    x = np.random.default_rng(0).normal(0, 30, m.n_params)   # any x maps into the ranges
    q = m.restrict(x)
    assert np.all((q.I >= 0) & (q.I <= 1)) and 0 <= q.rho <= 0.999


def test_masses_and_marginals_match_the_oracle():
    Lu, Hu, Hv = 0.1, 0.9, 0.05
    I, sn, sb, ds, rho = (0.2, 0.5, 0.85), 0.03, 0.7, 1.1, 0.4
    for Ii in I:
        assert abs(B.mass_pure(Lu, Hu, Hv, Ii, sn, rho) - OB.M_i(Lu, Hu, Hv, Ii, sn, rho)) < 1e-14
        assert abs(B.marginal_pure(0.3, Ii, sn) - OB.marginal_p_i(0.3, Ii, sn)) < 1e-12
    for i, j in B.pairs(3):
        a = B.mass_interface(Lu, Hu, Hv, I[i], I[j], sn, sb, ds, rho)
        b = OB.M_ij(Lu, Hu, Hv, I[i], I[j], sn, sb, ds, rho)
        assert abs(a / b - 1) < 1e-8
        for u in (0.25, 0.5, 0.7):
            assert abs(B.marginal_interface(u, I[i], I[j], sn, ds) / OB.marginal_p_ij(u, I[i], I[j], sn, ds) - 1) < 1e-8
    w = np.array([0.3, 0.3, 0.2, 0.1, 0.05, 0.05])
    p = B.Params(w, np.array(I), math.log(sn), math.log(sb), math.log(ds), rho)
    assert abs(B.normaliser((Lu, Hu, Hv), p) / OB.normaliser(Lu, Hu, Hv, w, I, sn, sb, ds, rho) - 1) < 1e-8


def test_tail_cutoffs_known_answers():
    h = np.array([0, 10, 10, 0, 20, 0], float)   # 40 in all over [0, 6)
    assert tail_cutoffs(h, 0.0, 0.0, 0.0, 6.0) == (0.0, 6.0)
    lo, hi = tail_cutoffs(h, 0.125, 0.25, 0.0, 6.0)
    assert lo == pytest.approx(1.5)               # 5 of 40 at half of bin 1
    assert hi == pytest.approx(4.5, rel=1e-6)      # 30 of 40 at half of bin 4


def test_peak_syntax():
    assert parse_peaks("{0.0339, 0.0409,0.0527}") == [0.0339, 0.0409, 0.0527]
    with pytest.raises(Exception):
        parse_peaks("0.1,0.2")
    with pytest.raises(Exception):
        parse_peaks("{0.1}")
