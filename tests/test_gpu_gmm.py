"""gmm-peaks on the GPU: the fused E+M pass and the hard assignment against
the oracle's numpy passes (FP64, different summation order: 1e-11 relative),
determinism, the underflow rule, and the driver end to end on phantoms."""
import numpy as np
import pytest

from oracle import gmm as OG
import segmentctscans_amd as S
from segmentctscans_amd.gmm import Samples, gmm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [1, 2, 3, 5, 6, 9, 16])
def test_pass_and_assign_match_oracle(ctx, K):
    rng = np.random.default_rng(K)
    x = rng.normal(0.3, 0.1, 200_003)
    mu = np.sort(rng.uniform(0.0, 0.6, K))
    var = rng.uniform(0.0005, 0.01, K)
    pi = rng.uniform(0.5, 1.5, K)
    pi /= pi.sum()
    s = Samples(ctx, x)
    w, s1, s2, nll = s.em_pass(pi, mu, var)
    ow, os1, os2, onll = OG.em_pass(x, pi, mu, var)
    assert np.allclose(w, ow, rtol=1e-11, atol=1e-9)
    assert np.allclose(s1, os1, rtol=1e-11, atol=1e-9)
    assert np.allclose(s2, os2, rtol=1e-11, atol=1e-9)
    assert nll == pytest.approx(onll, rel=1e-11)
    again = s.em_pass(pi, mu, var)   # deterministic: fixed grid, ordered sums
    assert all(np.array_equal(a, b) for a, b in zip(again, (w, s1, s2, nll)))
    cl = s.assign(pi, mu, var)
    ocl = OG.assign(x, pi, mu, var)
    # ties aside (none in continuous data), the argmax is the oracle's
    assert np.mean(cl == ocl) > 0.99999
    s.close()


def test_underflow_and_empty(ctx):
    s = Samples(ctx, np.array([1e6, 2e6]))
    w, s1, _, nll = s.em_pass([0.5, 0.5], [0.0, 1.0], [1e-4, 1e-4])
    assert np.allclose(w, [1.0, 1.0]) and np.allclose(s1, [1.5e6, 1.5e6])
    assert nll == pytest.approx(-2 * np.log(np.finfo(np.float64).tiny))
    s.close()
    e = Samples(ctx, np.zeros(0))
    w, _, _, nll = e.em_pass([1.0], [0.0], [1.0])
    assert w[0] == 0 and nll == 0
    assert e.assign([1.0], [0.0], [1.0]).size == 0
    e.close()
    with pytest.raises(S.SctError):
        Samples(ctx, np.ones(3)).em_pass(np.ones(17) / 17, np.zeros(17), np.ones(17))


def test_em_on_gpu_equals_em_on_oracle_passes(ctx):
    """The host EM loop over the device passes and over the oracle's passes
    reaches the same fit (same seed, same starting clusters)."""
    rng = np.random.default_rng(11)
    mus, sds = np.array([0.17, 0.23, 0.42]), np.array([0.02, 0.018, 0.03])
    comp = rng.choice(3, 50_000, p=[0.5, 0.2, 0.3])
    x = rng.normal(mus[comp], sds[comp])
    init = np.array([0.15, 0.25, 0.40])
    cl0 = np.argmin(np.abs(init[None, :] - x[:, None]), axis=1)
    g = gmm(Samples(ctx, x), init, cl0, 2, seed=4)
    h = gmm(OG.HostSamples(x), init, cl0, 2, seed=4)
    assert g[0] == pytest.approx(h[0], rel=1e-9)
    assert np.allclose(g[1], h[1], rtol=1e-6) and np.allclose(g[2], h[2], rtol=1e-7)
    assert np.mean(g[3] == h[3]) > 0.9999


def test_gmm_peaks_driver_end_to_end(tmp_path, capsys):
    """A cylinder of three intensity levels (raw 20, 50, 80 over thirds of the
    columns) with N(0, 2^2) noise as the denoised volume and the same plus
    more noise as the noised one: the fitted centres and deviations and the
    raw statistics come out near the truth."""
    from segmentctscans_amd import Field
    from segmentctscans_amd.archive import save_bin
    from segmentctscans_amd.gmm_peaks import main
    rng = np.random.default_rng(5)
    f = Field.cylinder(40, 60, 60, 29.5, 29.5, 27.0)
    m = f.mapping
    lvl = np.zeros(m.span, np.float32)
    for col in range(m.col_begin, m.col_end):
        b, rows = int(m.base[col]), m.row_end(col) - m.row_begin(col)
        lvl[b:b + rows * m.nstride] = (0.2, 0.5, 0.8)[min(2, 3 * col // 60)]
    valid = m.valid_mask()
    den = np.where(valid, lvl + rng.normal(0, 0.02, m.span), 0).astype(np.float32)
    noi = np.where(valid, lvl + rng.normal(0, 0.05, m.span), 0).astype(np.float32)
    for name, d in (("d.bin", den), ("n.bin", noi)):
        save_bin(Field(m, d, lo=0.0, hi=100.0), tmp_path / name)
    rc = main(["--denoised", str(tmp_path / "d.bin"), "--noised", str(tmp_path / "n.bin"), "--peaks", "25", "45",
               "85", "--quiet", "--attempts", "2", "--samples", "100000"])
    assert rc == 0
    out = capsys.readouterr().out
    centres = [float(line.split(":")[1]) for line in out.splitlines() if "Center:" in line]
    devs = [float(line.split(":")[1]) for line in out.splitlines() if "Deviations" in line]
    assert np.allclose(centres, [20, 50, 80], atol=0.3)
    assert np.allclose(devs, [2, 2, 2], rtol=0.1)
    raw = [line for line in out.splitlines() if "(raw data)" in line]
    assert len(raw) == 3
    sig = [float(r.split("sigma = ")[1].split(",")[0]) for r in raw]
    assert np.allclose(sig, [5, 5, 5], rtol=0.1)
