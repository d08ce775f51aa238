"""Seeded synthetic micro-CT phantoms (see csrc/phantom.h for the construction).

Input tooling for tests and the benchmark; no counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import Field, Mapping
from . import _native as N


class PhantomParams(C.Structure):
    _fields_ = [
        ("n_images", C.c_uint32), ("height", C.c_uint32), ("width", C.c_uint32),
        ("xc", C.c_double), ("yc", C.c_double), ("r", C.c_double),
        ("rmin", C.c_double), ("rmax", C.c_double), ("porosity", C.c_double),
        ("c_matrix", C.c_double), ("c_phase1", C.c_double), ("c_phase2", C.c_double),
        ("p_phase2", C.c_double), ("blur_sigma", C.c_double), ("blur_radius", C.c_int32),
        ("noise_sd", C.c_double), ("lo", C.c_double), ("hi", C.c_double),
        ("seed", C.c_uint64), ("max_attempts", C.c_int64),
    ]


class PhantomStats(C.Structure):
    _fields_ = [("n_spheres", C.c_uint32), ("porosity", C.c_double)]


@dataclass
class Phantom:
    field: Field            # normalised intensities u = (x - lo) / (hi - lo)
    labels: np.ndarray      # uint8 [span]: 0 matrix, 1 brine, 2 CO2 (pre-blur truth)
    params: PhantomParams
    n_spheres: int
    porosity: float

    def normalized_labels(self) -> np.ndarray:
        """The three phase intensities in the normalised units of `field`."""
        p = self.params
        return np.array([(c - p.lo) / (p.hi - p.lo) for c in (p.c_matrix, p.c_phase1, p.c_phase2)], np.float32)


def default_params(n_images: int, side: int, rmin: float, rmax: float, seed: int) -> PhantomParams:
    p = PhantomParams()
    N.phantom_lib().sct_phantom_defaults(C.byref(p), C.c_uint32(n_images), C.c_uint32(side),
                                         C.c_double(rmin), C.c_double(rmax), C.c_uint64(seed))
    return p


def make_phantom(n_images: int, side: int, rmin: float, rmax: float, seed: int = 0,
                 nthreads: int = 0, **overrides) -> Phantom:
    """Cylinder radius 0.45*side about ((side-1)/2, (side-1)/2); spheres U[rmin, rmax]."""
    p = default_params(n_images, side, rmin, rmax, seed)
    for k, v in overrides.items():
        setattr(p, k, v)
    m = Mapping.cylinder(p.n_images, p.height, p.width, p.xc, p.yc, p.r)
    u = np.zeros(m.span, np.float32)
    lab = np.zeros(m.span, np.uint8)
    st = PhantomStats()
    sr = np.ascontiguousarray(m.start_row, np.uint32)
    base = np.ascontiguousarray(m.base, np.uint32)
    L = N.phantom_lib()
    rc = L.sct_phantom_generate(C.byref(p), sr.ctypes.data_as(N.u32p), base.ctypes.data_as(N.u32p),
                                C.c_uint32(m.nstride), C.c_uint32(m.col_begin), C.c_uint32(m.col_end),
                                u.ctypes.data_as(N.f32p), lab.ctypes.data_as(C.POINTER(C.c_uint8)),
                                C.byref(st), C.c_int(nthreads))
    if rc:
        raise RuntimeError(f"sct_phantom_generate failed ({rc})")
    f = Field(m, u, lo=p.lo, hi=p.hi)
    return Phantom(f, lab, p, st.n_spheres, st.porosity)


def crop_box(ph: Phantom, n0: int, r0: int, c0: int, size: int) -> Field:
    """A size^3 box cut out of a phantom (rows/cols/n starting at r0/c0/n0),
    stored in a box layout (cylinder radius larger than the box diagonal).
    Every (row, col) of the box must lie inside the phantom's mask."""
    m = ph.field.mapping
    box = Mapping.cylinder(size, size, size, (size - 1) / 2, (size - 1) / 2, 4.0 * size)
    assert box.lines == size * size
    out = np.zeros(box.span, np.float32)
    src = ph.field.data
    for col in range(size):
        cc = c0 + col
        rb, re = m.row_begin(cc), m.row_end(cc)
        if not (rb <= r0 and r0 + size <= re):
            raise ValueError("crop leaves the mask")
        line0 = int(m.base[cc]) // m.nstride + (r0 - rb)
        blk = src[line0 * m.nstride:(line0 + size) * m.nstride].reshape(size, m.nstride)[:, n0:n0 + size]
        out[col * size * box.nstride:(col + 1) * size * box.nstride].reshape(size, box.nstride)[:, :size] = blk
    return Field(box, out, lo=ph.field.lo, hi=ph.field.hi)


# The batched workload: 64 sub-volumes x 9 lambdas, 300 iterations.
C4_LAMBDAS = tuple(20.0 + 5.0 * k for k in range(9))


@dataclass
class SubVolume:
    field: Field
    c: np.ndarray           # K normalised label intensities, ascending
    n: int
    height: int
    width: int
    box: bool               # box domain (True) or cylinder of radius 0.45 min(h, w)


def c4_volumes(seed: int = 3, count: int = 64, workers: int = 8) -> list[SubVolume]:
    """The sub-volumes of the batched workload: edges n, height, width ~
    U[48, 192] (seeded), a box or a cylinder domain with p = 1/2, spheres
    U[4, 12], K ~ U{3, 4, 5}.  Labels are the three phase intensities plus,
    for K > 3, extra phases evenly spaced in [0.030, 0.060] (raw units),
    all normalised and sorted.  The phantoms are generated on `workers`
    host threads (sphere placement is serial per phantom)."""
    import concurrent.futures as cf
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n, h, w = (int(x) for x in rng.integers(48, 193, 3))
        box = bool(rng.integers(0, 2))
        K = int(rng.integers(3, 6))
        specs.append((i, n, h, w, box, K))

    def one(spec):
        i, n, h, w, box, K = spec
        r = 4.0 * max(h, w) if box else 0.45 * min(h, w)
        ph = make_phantom(n, max(h, w), 4.0, 12.0, seed=1000 + i, nthreads=1, height=h, width=w,
                          xc=(w - 1) / 2, yc=(h - 1) / 2, r=r)
        lo, hi = ph.field.lo, ph.field.hi
        extra = np.linspace(0.030, 0.060, K - 1)[1:-1] if K > 3 else np.zeros(0)
        c = np.sort(np.concatenate([ph.normalized_labels(), (extra - lo) / (hi - lo)])).astype(np.float32)
        c = c[:K] if c.size >= K else np.linspace(0.1, 0.9, K).astype(np.float32)
        return SubVolume(ph.field, c, n, h, w, box)

    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(one, specs))
