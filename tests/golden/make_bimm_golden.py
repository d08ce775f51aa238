"""Writes tests/golden/bimm_pdfs.npz from the marginal densities the reference
ships with bimm-peaks (bimm-peaks/pdf-0[67]-*.txt: two fits, each with its
full marginal p(u) and the weighted marginal of every pure phase and every
interface pair, 513 points in raw intensity units).  They are the reference's
own output of marginal_p / marginal_p_i / marginal_p_ij (bimm-peaks/puv.h,
written by bimm-peaks.cpp:600-690) and pin the oracle's marginals.

    python tests/golden/make_bimm_golden.py [/root/reference]
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def main(ref: str = "/root/reference") -> None:
    src = Path(ref) / "bimm-peaks"
    out = {}
    for run in ("06", "07"):
        for part in ("full", "00", "01", "02", "00-01", "00-02", "01-02"):
            rows = (src / f"pdf-{run}-{part}.txt").read_text().split("\n")[1:]
            xy = np.array([[float(t) for t in r.split()] for r in rows if r.strip()], np.float64)
            out[f"{run}_{part.replace('-', '_')}"] = xy
    np.savez_compressed(HERE / "bimm_pdfs.npz", **out)
    print("wrote", HERE / "bimm_pdfs.npz", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(*sys.argv[1:])
