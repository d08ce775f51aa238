"""bimm-peaks on the GPU: the sample likelihood sct_bimm_nll against the FP64
oracle (oracle/bimm.py, its integrals by QUADPACK at 1e-12) for smooth and
sharp parameter sets; the fit recovering a known mixture from samples drawn
from it; and the driver end to end on a phantom volume."""
import math

import numpy as np
import pytest

from oracle import bimm as OB
from segmentctscans_amd import archive, make_context
from segmentctscans_amd import bimm as B
from segmentctscans_amd.phantom import make_phantom

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def draw(rng, n, w, I, sn, sb, ds, rho):
    """Samples of the mixture: a pure phase's v is sqrt(a / 2) times a chi
    variate with 3 degrees of freedom, an interface's v the length of
    (G(t), 0, 0) plus that noise (noncentral chi, 3 degrees), t ~ U(-ds, ds)."""
    from scipy.special import erf
    K = len(I)
    a = sn * sn * (1 - rho)
    comp = rng.choice(len(w), n, p=np.asarray(w) / np.sum(w))
    u, v = np.empty(n), np.empty(n)
    noise = rng.normal(0, math.sqrt(a / 2), (n, 3))
    for c in range(len(w)):
        sel = comp == c
        m = int(sel.sum())
        if c < K:
            u[sel] = rng.normal(I[c], sn, m)
            v[sel] = np.linalg.norm(noise[sel], axis=1)
        else:
            i, j = B.pairs(K)[c - K]
            t = rng.uniform(-ds, ds, m)
            It = I[i] + 0.5 * (I[j] - I[i]) * (1 + erf(t / math.sqrt(2)))
            G = abs(I[j] - I[i]) / sb * np.exp(-t * t / 2) / math.sqrt(2 * math.pi)
            u[sel] = rng.normal(It, sn)
            vec = noise[sel].copy()
            vec[:, 0] += G
            v[sel] = np.linalg.norm(vec, axis=1)
    return u, v


CASES = {
    "smooth": (np.array([0.35, 0.3, 0.2, 0.05, 0.05, 0.05]), (0.2, 0.5, 0.8), 0.04, 1.0, 1.2, 0.3),
    "sharp": (np.array([0.3, 0.3, 0.3, 0.04, 0.03, 0.03]), (0.3, 0.35, 0.6), 2e-3, 0.15, 1.45, 0.95),
}


@pytest.fixture(scope="module")
def ctx():
    return make_context(0)


@pytest.mark.parametrize("case", sorted(CASES))
def test_nll_matches_oracle(ctx, case):
    w, I, sn, sb, ds, rho = CASES[case]
    rng = np.random.default_rng(3)
    u, v = draw(rng, 40, w, I, sn, sb, ds, rho)
    s = B.UVSamples(ctx, u, v)
    got = s.nll_sum(w, I, math.log(sn), math.log(sb), math.log(ds), rho)
    want = OB.nll_sum(u, v, w, I, math.log(sn), math.log(sb), math.log(ds), rho)
    print(f"{case}: GPU {got:.15g}, oracle {want:.15g}")
    assert abs(got - want) <= 1e-9 * abs(want) + 1e-9
    # per-sample terms through the range arguments
    for k in (0, 7, 39):
        one = s.nll_sum(w, I, math.log(sn), math.log(sb), math.log(ds), rho, k, 1)
        ref = -math.log(OB.full_p(u[k], v[k], w, I, math.log(sn), math.log(sb), math.log(ds), rho))
        assert abs(one - ref) <= 1e-9 * max(1.0, abs(ref))
    a = s.nll_sum(w, I, math.log(sn), math.log(sb), math.log(ds), rho, 0, 17)
    b = s.nll_sum(w, I, math.log(sn), math.log(sb), math.log(ds), rho, 17, 23)
    assert abs(a + b - got) <= 1e-12 * abs(got)
    s.close()


def test_fit_recovers_a_known_mixture(ctx):
    w, I, sn, sb, ds, rho = (np.array([0.45, 0.3, 0.15, 0.04, 0.03, 0.03]), (0.25, 0.5, 0.75), 0.03, 1.0, 1.2, 0.2)
    rng = np.random.default_rng(11)
    u, v = draw(rng, 60000, w, I, sn, sb, ds, rho)
    Lu, Hu, Hv = 0.12, 0.88, float(np.quantile(v, 0.99))
    keep = (u >= Lu) & (u <= Hu) & (v > 0) & (v <= Hv)
    u, v = u[keep][:40000], v[keep][:40000]
    s = B.UVSamples(ctx, u, v)
    res = B.fit(s, (Lu, Hu, Hv), [0.22, 0.53, 0.72], batch=10000, batches=4, restart_period=60, full_iters=1500)
    s.close()
    p = res.params
    print(f"fitted I {p.I}, sn {math.exp(p.log_sn):.4f}, w {p.w}, iterations {res.iterations}")
    assert np.allclose(p.I, I, atol=5e-3)
    assert abs(math.exp(p.log_sn) / sn - 1) < 0.05
    assert np.allclose(p.w[:3], w[:3], atol=0.03)


def test_driver_end_to_end(tmp_path, capsys):
    from segmentctscans_amd import bimm_peaks
    # the default noise (sd 0.008 raw, 0.14 in normalised units) leaves the
    # matrix and brine peaks 0.9 sd apart, which no fit separates; at sd 0.002
    # (0.035 normalised) the three phases are resolved
    ph = make_phantom(64, 96, 4.0, 12.0, seed=21, noise_sd=0.002)
    path = tmp_path / "vol.bin"
    archive.save_bin(ph.field, str(path))
    lo, hi = ph.field.lo, ph.field.hi
    peaks = [lo + c * (hi - lo) for c in ph.normalized_labels()]
    rc = bimm_peaks.main([str(path), "--peaks", "{" + ",".join(f"{x:.6f}" for x in peaks) + "}", "--batch", "5000",
                          "--batches", "4", "--restart-period", "40", "--full-iters", "400", "--quiet",
                          "--pdf-name", str(tmp_path / "fit")])
    out = capsys.readouterr().out
    assert rc == 0, out
    lines = out.strip().split("\n")
    fitted = [float(x) for x in lines[3:6]]
    print(out)
    for f, want in zip(fitted, peaks):
        assert abs(f - want) < 0.1 * (hi - lo)
    full = np.loadtxt(tmp_path / "fit-full.txt", skiprows=1)
    parts = sum(np.loadtxt(tmp_path / f"fit-{n}.txt", skiprows=1)[:, 1]
                for n in ("00", "01", "02", "00-01", "00-02", "01-02"))
    assert full.shape == (513, 2) and np.allclose(parts, full[:, 1], rtol=1e-12)
