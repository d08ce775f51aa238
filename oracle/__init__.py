"""ctypes binding of the CPU oracle (oracle/_build/libsct_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker / baseline, never as the
product path.  See sct_oracle.h for the reference files each function follows
and oracle_body.h for how each function is pinned.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libsct_oracle.so"

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class OrLayout(C.Structure):
    _fields_ = [
        ("n_images", C.c_uint32), ("height", C.c_uint32), ("width", C.c_uint32),
        ("nstride", C.c_uint32), ("col_begin", C.c_uint32), ("col_end", C.c_uint32),
        ("start_row", u32p), ("base", u32p),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB_PATH))
        L.or_cylinder.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                  C.c_uint32, u32p, u32p, u32p, u32p, C.POINTER(C.c_uint64)]
        L.or_grad_d.argtypes = [C.POINTER(OrLayout), f64p, f64p, f64p, f64p]
        L.or_div_d.argtypes = [C.POINTER(OrLayout), f64p, f64p, f64p, f64p]
        L.or_proj_simplex_d.argtypes = [C.c_int, f64p]
        L.or_proj_simplex_f.argtypes = [C.c_int, f32p]
        L.or_dykstra_d.argtypes = [C.c_int, f64p, C.c_int, C.c_double]
        L.or_dykstra_f.argtypes = [C.c_int, f32p, C.c_int, C.c_double]
        L.or_potts_d.argtypes = [C.POINTER(OrLayout), f32p, C.c_int, f32p, C.c_float, C.c_float, C.c_float,
                                 C.c_int, C.c_int, f32p, f64p, f64p, f64p, C.c_int, C.c_int]
        L.or_potts_f.argtypes = [C.POINTER(OrLayout), f32p, C.c_int, f32p, C.c_float, C.c_float, C.c_float,
                                 C.c_int, C.c_int, f32p, f32p, f32p, f32p, C.c_int, C.c_int]
        L.or_support_d.argtypes = [C.c_int, f64p]
        L.or_support_d.restype = C.c_double
        L.or_support_irls_d.argtypes = [C.c_int, f64p]
        L.or_support_irls_d.restype = C.c_double
        L.or_potts_gap_d.argtypes = [C.POINTER(OrLayout), f32p, C.c_int, f32p, C.c_float, f64p, f64p, f64p, f64p]
        i8p = C.POINTER(C.c_int8)
        L.or_adj_loss_d.argtypes = [C.POINTER(OrLayout), f64p, f32p, C.c_int, f64p, C.c_double, C.c_double,
                                    C.c_double, C.c_int, i8p, C.c_double]
        L.or_adj_loss_d.restype = C.c_double
        L.or_adj_grad_d.argtypes = [C.POINTER(OrLayout), f64p, f32p, C.c_int, f64p, C.c_double, C.c_double,
                                    C.c_double, C.c_int, i8p, C.c_double, C.c_int, f64p, f64p, f64p]
        L.or_adj_grad_d.restype = C.c_double
        u16p = C.POINTER(C.c_uint16)
        L.or_u16_hist.argtypes = [C.POINTER(OrLayout), u16p, C.POINTER(C.c_uint64)]
        L.or_find_intensity_range.argtypes = [C.POINTER(OrLayout), u16p, C.c_double, u16p, u16p]
        L.or_convert_to_fp.argtypes = [C.POINTER(OrLayout), C.c_void_p, C.c_int, C.c_float, C.c_float, f32p]
        L.or_bf16_round.argtypes = [C.c_float]
        L.or_bf16_round.restype = C.c_float
        L.or_f16_round.argtypes = [C.c_float]
        L.or_f16_round.restype = C.c_float
        L.or_chambolle_d.argtypes = [C.POINTER(OrLayout), f64p, C.c_float, C.c_float, C.c_int, f64p, f64p, C.c_int]
        L.or_chambolle_f.argtypes = [C.POINTER(OrLayout), f32p, C.c_float, C.c_float, C.c_int, f32p, f32p, C.c_int]
        L.or_calc_hist.argtypes = [C.POINTER(OrLayout), f32p, C.c_float, C.c_float, C.c_uint32, u32p]
        L.or_min_max.argtypes = [C.POINTER(OrLayout), f32p, f32p, f32p]
        L.or_grad_min_max.argtypes = [C.POINTER(OrLayout), f32p, f32p, f32p]
        L.or_grad_magn.argtypes = [C.POINTER(OrLayout), f32p, f32p]
        L.or_uv_hist.argtypes = [C.POINTER(OrLayout), f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float,
                                 C.c_uint32, C.c_uint32, u32p]
        for name in ("or_cylinder", "or_dykstra_d", "or_dykstra_f", "or_potts_d", "or_potts_f",
                     "or_chambolle_d", "or_chambolle_f"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Layout:
    """Cylinder layout built by the oracle's own transcription of the constructor."""

    def __init__(self, n_images, height, width, xc, yc, r, nstride=None):
        self.nstride = ((n_images + 31) // 32) * 32 if nstride is None else nstride
        self.n_images, self.height, self.width = n_images, height, width
        self.start_row = np.zeros(width, np.uint32)
        self.base = np.zeros(width + 1, np.uint32)
        cb, ce, ds = C.c_uint32(), C.c_uint32(), C.c_uint64()
        lib().or_cylinder(n_images, height, width, xc, yc, r, self.nstride, _ptr(self.start_row, u32p),
                          _ptr(self.base, u32p), C.byref(cb), C.byref(ce), C.byref(ds))
        self.col_begin, self.col_end, self.data_size = cb.value, ce.value, ds.value
        if self.col_begin == 0xFFFFFFFF:
            self.col_begin = self.col_end = 0
        self.span = int(self.base[width])

    @classmethod
    def from_mapping(cls, m) -> "Layout":
        self = cls.__new__(cls)
        for k in ("n_images", "height", "width", "nstride", "col_begin", "col_end", "data_size"):
            setattr(self, k, int(getattr(m, k)))
        self.start_row = np.ascontiguousarray(m.start_row, np.uint32)
        self.base = np.ascontiguousarray(m.base, np.uint32)
        self.span = int(self.base[self.width])
        return self

    def c(self) -> OrLayout:
        return OrLayout(self.n_images, self.height, self.width, self.nstride, self.col_begin, self.col_end,
                        _ptr(self.start_row, u32p), _ptr(self.base, u32p))


STORAGE = {"fp32": 0, "bf16": 1, "fp16": 2}


def potts(layout: Layout, u: np.ndarray, c, lam, tau1, tau2, iters, *, precision="f64", slab_w=0,
          want_state=False, nthreads=0, storage="fp32"):
    """Returns dict(res, V, xi, dq, rc)."""
    c = np.ascontiguousarray(c, np.float32)
    K = c.size
    dt = np.float64 if precision == "f64" else np.float32
    tp = f64p if precision == "f64" else f32p
    res = np.zeros(layout.span, np.float32)
    V = np.zeros((K - 1) * layout.span, dt) if want_state else None
    xi = np.zeros(3 * (K - 1) * layout.span, dt) if want_state else None
    dq = np.zeros(max(iters, 1), dt)
    fn = lib().or_potts_d if precision == "f64" else lib().or_potts_f
    lay = layout.c()
    u = np.ascontiguousarray(u, np.float32)
    rc = fn(C.byref(lay), _ptr(u, f32p), K, _ptr(c, f32p), lam, tau1, tau2, iters, slab_w, _ptr(res, f32p),
            _ptr(V, tp), _ptr(xi, tp), _ptr(dq, tp), STORAGE[storage], nthreads)
    out = dict(rc=rc, res=res, dq=dq[:iters])
    if want_state:
        out["V"] = V.reshape(K - 1, layout.span)
        out["xi"] = xi.reshape(3 * (K - 1), layout.span)
    return out


def support(p, method: str = "auto") -> float:
    """h_K(p) for p of shape (K, 3) with rows summing to 0 (K = 2..6)."""
    p = np.ascontiguousarray(p, np.float64)
    fn = lib().or_support_irls_d if method == "irls" else lib().or_support_d
    return fn(p.shape[0], _ptr(p, f64p))


def potts_gap(layout: Layout, u, c, lam, V, xi) -> tuple[float, float]:
    """FP64 primal and dual energies (E, D) of the Potts saddle problem for the
    state V ((K-1) x span) and Xi (3(K-1) x span)."""
    c = np.ascontiguousarray(c, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    V = np.ascontiguousarray(V, np.float64)
    xi = np.ascontiguousarray(xi, np.float64)
    E, D = C.c_double(), C.c_double()
    lay = layout.c()
    rc = lib().or_potts_gap_d(C.byref(lay), _ptr(u, f32p), c.size, _ptr(c, f32p), lam, _ptr(V, f64p),
                              _ptr(xi, f64p), C.byref(E), C.byref(D))
    assert rc == 0
    return E.value, D.value


def nearest_labels(u, c) -> np.ndarray:
    """Start labels of the iteration: index of the nearest c (lowest index on ties)."""
    u = np.asarray(u, np.float64)
    c = np.asarray(c, np.float64)
    return np.argmin(np.abs(c[None, :] - u[:, None]), axis=1).astype(np.int8)


def _adj_args(layout, u, target, c, init):
    u = np.ascontiguousarray(u, np.float64)
    target = np.ascontiguousarray(target, np.float32)
    c = np.ascontiguousarray(c, np.float64)
    init = None if init is None else np.ascontiguousarray(init, np.int8)
    return u, target, c, init


def adj_loss(layout: Layout, u, target, c, lam, tau1, tau2, iters, *, init=None, tol=1e-6) -> float:
    """FP64 loss 1/(N K) sum (V - onehot(target))^2 after iters iterations."""
    u, target, c, init = _adj_args(layout, u, target, c, init)
    lay = layout.c()
    return lib().or_adj_loss_d(C.byref(lay), _ptr(u, f64p), _ptr(target, f32p), c.size, _ptr(c, f64p), lam, tau1,
                               tau2, iters, None if init is None else init.ctypes.data_as(C.POINTER(C.c_int8)), tol)


def adj_grad(layout: Layout, u, target, c, lam, tau1, tau2, iters, *, init=None, tol=1e-6, ck=16):
    """FP64 adjoint: (loss, dL/dc[K], dL/dlambda, dL/du[span])."""
    u, target, c, init = _adj_args(layout, u, target, c, init)
    gc = np.zeros(c.size)
    gl = C.c_double()
    gu = np.zeros(layout.span)
    lay = layout.c()
    loss = lib().or_adj_grad_d(C.byref(lay), _ptr(u, f64p), _ptr(target, f32p), c.size, _ptr(c, f64p), lam, tau1,
                               tau2, iters, None if init is None else init.ctypes.data_as(C.POINTER(C.c_int8)), tol,
                               ck, _ptr(gc, f64p), C.byref(gl), _ptr(gu, f64p))
    return loss, gc, gl.value, gu


def u16_hist(layout: Layout, u) -> np.ndarray:
    u = np.ascontiguousarray(u, np.uint16)
    h = np.zeros(65536, np.uint64)
    lay = layout.c()
    lib().or_u16_hist(C.byref(lay), u.ctypes.data_as(C.POINTER(C.c_uint16)), h.ctypes.data_as(C.POINTER(C.c_uint64)))
    return h


def find_intensity_range(layout: Layout, u, tail: float) -> tuple[int, int]:
    u = np.ascontiguousarray(u, np.uint16)
    lo, hi = C.c_uint16(), C.c_uint16()
    lay = layout.c()
    lib().or_find_intensity_range(C.byref(lay), u.ctypes.data_as(C.POINTER(C.c_uint16)), tail, C.byref(lo),
                                  C.byref(hi))
    return lo.value, hi.value


def convert_to_fp(layout: Layout, x, lo: float, hi: float) -> np.ndarray:
    x = np.ascontiguousarray(x)
    assert x.dtype in (np.uint16, np.float32)
    out = np.zeros(layout.span, np.float32)
    lay = layout.c()
    lib().or_convert_to_fp(C.byref(lay), x.ctypes.data, int(x.dtype == np.uint16), lo, hi, _ptr(out, f32p))
    return out


def store_round(x, storage: str) -> np.ndarray:
    """Round float32 values through bfloat16 / half storage (oracle's rounding)."""
    fn = {"bf16": lib().or_bf16_round, "fp16": lib().or_f16_round}[storage]
    x = np.asarray(x, np.float32)
    return np.array([fn(float(v)) for v in x.reshape(-1)], np.float32).reshape(x.shape)


def chambolle(layout: Layout, g: np.ndarray, lam, tau, iters, *, precision="f32", nthreads=0):
    dt = np.float64 if precision == "f64" else np.float32
    tp = f64p if precision == "f64" else f32p
    gg = np.array(g, dt, copy=True)
    xi = np.zeros(3 * layout.span, dt)
    dq = np.zeros(max(iters, 1), dt)
    fn = lib().or_chambolle_d if precision == "f64" else lib().or_chambolle_f
    lay = layout.c()
    rc = fn(C.byref(lay), _ptr(gg, tp), lam, tau, iters, _ptr(xi, tp), _ptr(dq, tp), nthreads)
    assert rc == 0
    return gg, xi.reshape(3, layout.span), dq[:iters]


def calc_hist(layout: Layout, u, rs, re, nbins) -> np.ndarray:
    cnt = np.zeros(nbins, np.uint32)
    lay = layout.c()
    lib().or_calc_hist(C.byref(lay), _ptr(np.ascontiguousarray(u, np.float32), f32p), rs, re, nbins, _ptr(cnt, u32p))
    return cnt


def min_max(layout: Layout, u, positive=False):
    a, b = C.c_float(), C.c_float()
    lay = layout.c()
    fn = lib().or_grad_min_max if positive else lib().or_min_max
    fn(C.byref(lay), _ptr(np.ascontiguousarray(u, np.float32), f32p), C.byref(a), C.byref(b))
    return a.value, b.value


def grad_magn(layout: Layout, u) -> np.ndarray:
    v = np.zeros(layout.span, np.float32)
    lay = layout.c()
    lib().or_grad_magn(C.byref(lay), _ptr(np.ascontiguousarray(u, np.float32), f32p), _ptr(v, f32p))
    return v


def uv_hist(layout: Layout, u, v, u0, u1, v0, v1, nbins):
    stride = ((nbins + 31) // 32) * 32
    cnt = np.zeros((nbins - 1) * stride + nbins, np.uint32)
    lay = layout.c()
    lib().or_uv_hist(C.byref(lay), _ptr(np.ascontiguousarray(u, np.float32), f32p),
                     _ptr(np.ascontiguousarray(v, np.float32), f32p), u0, u1, v0, v1, nbins, stride, _ptr(cnt, u32p))
    return cnt, stride


def grad(layout: Layout, v):
    p = [np.zeros(layout.span) for _ in range(3)]
    lay = layout.c()
    lib().or_grad_d(C.byref(lay), _ptr(np.ascontiguousarray(v, np.float64), f64p), *[_ptr(x, f64p) for x in p])
    return p


def div(layout: Layout, p):
    d = np.zeros(layout.span)
    lay = layout.c()
    pp = [np.ascontiguousarray(x, np.float64) for x in p]
    lib().or_div_d(C.byref(lay), *[_ptr(x, f64p) for x in pp], _ptr(d, f64p))
    return d


def proj_simplex(q, precision="f64"):
    dt, tp = (np.float64, f64p) if precision == "f64" else (np.float32, f32p)
    a = np.array(q, dt, copy=True)
    (lib().or_proj_simplex_d if precision == "f64" else lib().or_proj_simplex_f)(a.size, _ptr(a, tp))
    return a


def dykstra(q, max_sweeps=20, tol=1e-6, precision="f64"):
    dt, tp = (np.float64, f64p) if precision == "f64" else (np.float32, f32p)
    a = np.array(q, dt, copy=True).reshape(-1, 3)
    K = a.shape[0]
    a = np.ascontiguousarray(a)
    s = (lib().or_dykstra_d if precision == "f64" else lib().or_dykstra_f)(K, _ptr(a, tp), max_sweeps, tol)
    return a, s
