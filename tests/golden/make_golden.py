"""Writes the golden fixtures of tests/golden/ with the oracle (FP32 mirror and
FP64), from inputs stored in the same files.  The reference has no tests or
fixtures of its own, so these vectors come from the oracle, which the
tests/test_oracle_*.py suites pin by mathematics; committed, they freeze the
expected bits for the GPU path and guard the oracle against regressions.

    python tests/golden/make_golden.py      # rewrites potts_small.npz, potts_24.npz,
                                            # chambolle_24.npz, hist_edges.npz
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from segmentctscans_amd import Mapping  # noqa: E402
from segmentctscans_amd.phantom import make_phantom  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
TAU = float(np.float32(1.0 / np.sqrt(12.0)))
TAU_REF = float(np.float32(0.99) * np.float32(np.sqrt(np.float32(1.0 / 12.0))))
TAU_TV = float(np.float32(1.0 / 6.0) * np.float32(0.99))


def mapping_fields(m):
    return dict(n_images=m.n_images, height=m.height, width=m.width, nstride=m.nstride,
                start_row=m.start_row, base=m.base)


def potts_small():
    """16 x 12 cylinder, n = 10, K = 2..6, tau = 0.99/sqrt(12), 10 iterations."""
    geo = dict(xc=7.5, yc=5.5, r=5.0)
    ph = make_phantom(10, 16, 2.0, 4.0, seed=101, height=12, width=16, **geo)
    m = ph.field.mapping
    lay = O.Layout.from_mapping(m)
    out = dict(u=ph.field.data, xc=geo["xc"], yc=geo["yc"], r=geo["r"], **mapping_fields(m))
    for K in range(2, 7):
        c = np.linspace(0.2, 0.75, K).astype(np.float32)
        r32 = O.potts(lay, ph.field.data, c, 30.0, TAU_REF, TAU_REF, 10, precision="f32", want_state=True)
        r64 = O.potts(lay, ph.field.data, c, 30.0, TAU_REF, TAU_REF, 10, precision="f64", want_state=True)
        out.update({f"c{K}": c, f"V{K}": r32["V"], f"xi{K}": r32["xi"], f"res{K}": r32["res"],
                    f"dq{K}": r32["dq"], f"res64_{K}": r64["res"], f"V64_{K}": r64["V"]})
    np.savez_compressed(os.path.join(HERE, "potts_small.npz"), **out)


def potts_24():
    """24^3 cylinder, K = 3, tau = 1/sqrt(12), 50 iterations."""
    ph = make_phantom(24, 24, 2.0, 5.0, seed=102)
    m = ph.field.mapping
    lay = O.Layout.from_mapping(m)
    c = ph.normalized_labels()
    r32 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 50, precision="f32", want_state=True)
    r64 = O.potts(lay, ph.field.data, c, 30.0, TAU, TAU, 50, precision="f64", want_state=True)
    np.savez_compressed(os.path.join(HERE, "potts_24.npz"), u=ph.field.data, c=c, V=r32["V"], res=r32["res"],
                        dq=r32["dq"], res64=r64["res"], V64=r64["V"], xc=11.5, yc=11.5, r=0.45 * 24,
                        **mapping_fields(m))


def chambolle_24():
    """24^3 cylinder, lambda in {0.04, 0.12}, tau = 0.99/6, 50 iterations."""
    ph = make_phantom(24, 24, 2.0, 5.0, seed=103)
    m = ph.field.mapping
    lay = O.Layout.from_mapping(m)
    out = dict(g=ph.field.data, xc=11.5, yc=11.5, r=0.45 * 24, **mapping_fields(m))
    for tag, lam in (("004", 0.04), ("012", 0.12)):
        got, _, dq = O.chambolle(lay, ph.field.data, lam, TAU_TV, 50, precision="f32")
        out[f"out_{tag}"] = got
        out[f"dq_{tag}"] = dq
    np.savez_compressed(os.path.join(HERE, "chambolle_24.npz"), **out)


def hist_edges():
    """Values on bin edges, both range ends, out-of-range and NaN; 64 and 65536 bins."""
    m = Mapping.cylinder(32, 16, 16, 7.5, 7.5, 7.0)
    rng = np.random.default_rng(104)
    rs, re = np.float32(0.25), np.float32(1.75)
    vals = []
    for nb in (64, 65536):
        k = rng.integers(0, nb + 1, 200)
        vals.append((rs + k.astype(np.float64) * (float(re) - float(rs)) / nb).astype(np.float32))
    vals.append(np.array([rs, re, np.nextafter(rs, -1), np.nextafter(re, 2), np.nan, -5.0, 9.0], np.float32))
    pool = np.concatenate(vals + [rng.uniform(0.0, 2.0, 4000).astype(np.float32)])
    u = np.zeros(m.span, np.float32)
    valid = m.valid_mask()
    u[valid] = pool[rng.integers(0, pool.size, int(valid.sum()))]
    lay = O.Layout.from_mapping(m)
    out = dict(u=u, rs=rs, re=re, xc=7.5, yc=7.5, r=7.0, **mapping_fields(m))
    for nb in (64, 65536):
        out[f"counts_{nb}"] = O.calc_hist(lay, u, float(rs), float(re), nb)
    np.savez_compressed(os.path.join(HERE, "hist_edges.npz"), **out)


if __name__ == "__main__":
    potts_small()
    potts_24()
    chambolle_24()
    hist_edges()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
