"""Loads the fixtures of tests/golden/ (written by tests/golden/make_golden.py)."""
import os

import numpy as np

from segmentctscans_amd import Field, Mapping

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TAU = float(np.float32(1.0 / np.sqrt(12.0)))
TAU_REF = float(np.float32(0.99) * np.float32(np.sqrt(np.float32(1.0 / 12.0))))
TAU_TV = float(np.float32(1.0 / 6.0) * np.float32(0.99))


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mapping(z) -> Mapping:
    """The cylinder mapping of a fixture, rebuilt and checked against the stored tables."""
    m = Mapping.cylinder(int(z["n_images"]), int(z["height"]), int(z["width"]), float(z["xc"]),
                         float(z["yc"]), float(z["r"]))
    assert int(z["nstride"]) == m.nstride
    assert np.array_equal(m.start_row, z["start_row"]) and np.array_equal(m.base, z["base"])
    return m


def field(z, key="u") -> Field:
    return Field(mapping(z), z[key])
