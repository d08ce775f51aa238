"""MI355X-native segmentctscans hot path: Potts (CCP), Chambolle TV and the
histogram / extrema reductions on the packed cylinder layout.

Python mirror of the reference's C++ interface (names and argument meaning as
in the reference headers), calling the HIP library through the C ABI of
include/sct.h:

    make_context()                          include/cuda-context.h:11-14
    upload(ctx, field) / download(...)      include/cuda-context.h:24-31
    device_memory(ctx)                      include/cuda-context.h:33-37
    potts_min_partition(ctx, g, ci, ...)    potts-model/potts-min-partition.h:8-16
    chambolle(ctx, g, lam, tau, iters)      totalvarmin/chambolle.h:5-6
    calc_hist(ctx, f, ...)                  include/calc_hist.h:8-12
    min_max(ctx, f)                         include/min_max.h:7
    grad_magn(ctx, u)                       include/grad_magn.h:5-6
    uv_hist(ctx, u, v, ...)                 uv-hist/uv-hist.h:18-44
    Mapping / Field                         include/layout_cylinder.h, include/field.h
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

__all__ = [
    "SctError", "Mapping", "Field", "Context", "DeviceField", "PottsSolver",
    "make_context", "upload", "download", "device_memory", "potts_min_partition",
    "PottsTape", "chambolle", "calc_hist", "min_max", "grad_magn", "grad_min_max", "uv_hist",
    "UVHistReturn", "get_stride", "prev_stride", "max_phases",
]

max_phases = 6   # include/cuda-context.h:7


class SctError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[sct error {code}] {msg}")
        self.code = code


def get_stride(n: int) -> int:
    """Field<float>::get_stride, field.h:37-40."""
    return ((n + 31) // 32) * 32


def prev_stride(n: int) -> int:
    """Field<float>::prev_stride, field.h:42-45."""
    return (n // 32) * 32


# ---------------------------------------------------------------------------
# Layout and host field
# ---------------------------------------------------------------------------
@dataclass
class Mapping:
    """layout_cylinder::mapping (layout_cylinder.h:29-112), host tables."""
    n_images: int
    height: int
    width: int
    nstride: int
    col_begin: int
    col_end: int
    start_row: np.ndarray   # uint32 [width]
    base: np.ndarray        # uint32 [width + 1]
    data_size: int

    @classmethod
    def cylinder(cls, n_images: int, height: int, width: int, xc: float, yc: float, r: float,
                 nstride: int | None = None) -> "Mapping":
        """mapping_storage(extents{n, height, width}, xc, yc, r, nstride), layout_cylinder.h:152-195."""
        nstride = get_stride(n_images) if nstride is None else nstride
        sr = np.zeros(width, np.uint32)
        base = np.zeros(width + 1, np.uint32)
        cb, ce, ds = C.c_uint32(), C.c_uint32(), C.c_uint64()
        rc = N.lib().sct_cylinder_layout(n_images, height, width, xc, yc, r, nstride,
                                         sr.ctypes.data_as(N.u32p), base.ctypes.data_as(N.u32p),
                                         C.byref(cb), C.byref(ce), C.byref(ds))
        if rc:
            raise SctError(rc, "invalid cylinder layout")
        col_begin, col_end = cb.value, ce.value
        if col_begin == 0xFFFFFFFF:
            col_begin = col_end = 0
        return cls(n_images, height, width, nstride, col_begin, col_end, sr, base, ds.value)

    @property
    def span(self) -> int:
        return int(self.base[self.width])

    @property
    def lines(self) -> int:
        return self.span // self.nstride

    def row_begin(self, col: int) -> int:
        return int(self.start_row[col])

    def row_end(self, col: int) -> int:
        return int(self.start_row[col]) + int(self.base[col + 1] - self.base[col]) // self.nstride

    def line_rc(self) -> tuple[np.ndarray, np.ndarray]:
        """(row, col) of every line, in storage order."""
        rows, cols = [], []
        for col in range(self.col_begin, self.col_end):
            r0, r1 = self.row_begin(col), self.row_end(col)
            rows.append(np.arange(r0, r1))
            cols.append(np.full(r1 - r0, col))
        if not rows:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        return np.concatenate(rows), np.concatenate(cols)

    def valid_mask(self) -> np.ndarray:
        """Boolean [span]: True on in-mask voxels (n < n_images), False on padding."""
        m = np.zeros((self.lines, self.nstride), bool)
        m[:, : self.n_images] = True
        return m.reshape(-1)

    def c_layout(self) -> N.SctLayout:
        self._keep = (np.ascontiguousarray(self.start_row, np.uint32), np.ascontiguousarray(self.base, np.uint32))
        return N.SctLayout(self.n_images, self.height, self.width, self.nstride, self.col_begin,
                           self.col_end, self._keep[0].ctypes.data_as(N.u32p), self._keep[1].ctypes.data_as(N.u32p))

    def to_dense(self, data: np.ndarray, fill=0.0) -> np.ndarray:
        """Scatter a span array into a dense [width(col), height(row), n] box."""
        out = np.full((self.width, self.height, self.n_images), fill, dtype=data.dtype)
        rows, cols = self.line_rc()
        lines = data.reshape(self.lines, self.nstride)[:, : self.n_images]
        out[cols, rows, :] = lines
        return out


class Field:
    """Host volume Field<T> (field.h:23-153): span storage + mapping + lo/hi."""

    def __init__(self, mapping: Mapping, data: np.ndarray | None = None, dtype=np.float32,
                 lo: float = 0.0, hi: float = 0.0):
        self.mapping = mapping
        self.data = np.zeros(mapping.span, dtype) if data is None else np.ascontiguousarray(data)
        if self.data.shape != (mapping.span,):
            raise ValueError("data must have shape (span,)")
        self.lo, self.hi = lo, hi

    @classmethod
    def cylinder(cls, num_images: int, width: int, height: int, xc: float, yc: float, r: float) -> "Field":
        """Field(num_images, width, height, xc, yc, r), field.h:47-58."""
        return cls(Mapping.cylinder(num_images, height, width, xc, yc, r))


# ---------------------------------------------------------------------------
# Device context and fields
# ---------------------------------------------------------------------------
class Context:
    """HIP context: one device, one stream (replaces CudaContext)."""

    def __init__(self, device: int = 0):
        self._L = N.lib()
        h = C.c_void_p()
        rc = self._L.sct_ctx_create(device, C.byref(h))
        if rc:
            raise SctError(rc, f"sct_ctx_create(device={device}) failed (is a GPU visible?)")
        self.handle = h

    def check(self, rc: int) -> None:
        if rc:
            msg = self._L.sct_ctx_last_error(self.handle)
            raise SctError(rc, msg.decode() if msg else "")

    @property
    def stream(self) -> int:
        return self._L.sct_ctx_stream(self.handle)

    def synchronize(self) -> None:
        self.check(self._L.sct_ctx_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            self._L.sct_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_context(device: int = 0) -> Context:
    """make_context(), cuda-context.cpp:18-20."""
    return Context(device)


def device_memory(ctx: Context) -> tuple[int, int]:
    """device_memory(ctx) -> (available, total), cuda-context.cpp:22-41."""
    a, t = C.c_size_t(), C.c_size_t()
    ctx.check(N.lib().sct_device_memory(ctx.handle, C.byref(a), C.byref(t)))
    return a.value, t.value


class DeviceField:
    """DeviceField<T> (device-field-impl.h:14-54): voxels + device mapping tables."""

    def __init__(self, ctx: Context, handle: C.c_void_p, mapping: Mapping, dtype=np.float32):
        self.ctx, self.handle, self.mapping, self.dtype = ctx, handle, mapping, np.dtype(dtype)

    @property
    def data_ptr(self) -> int:
        return N.lib().sct_field_data(self.handle)

    @property
    def span(self) -> int:
        return N.lib().sct_field_span(self.handle)

    @property
    def data_size(self) -> int:
        return N.lib().sct_field_data_size(self.handle)

    def to_host(self) -> np.ndarray:
        out = np.empty(self.span, self.dtype)
        self.ctx.check(N.lib().sct_download(self.ctx.handle, self.handle, out.ctypes.data, out.size))
        return out

    def close(self) -> None:
        if self.handle:
            N.lib().sct_field_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dt(dtype) -> int:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return N.SCT_F32
    if dtype == np.uint16:
        return N.SCT_U16
    raise ValueError(f"unsupported dtype {dtype}")


def upload(ctx: Context, field: Field) -> DeviceField:
    """upload(ctx, mdspan), cuda-context.h:24-27."""
    data = np.ascontiguousarray(field.data)
    lay = field.mapping.c_layout()
    h = C.c_void_p()
    ctx.check(N.lib().sct_upload(ctx.handle, C.byref(lay), _dt(data.dtype), data.ctypes.data, C.byref(h)))
    return DeviceField(ctx, h, field.mapping, data.dtype)


def empty_like(ctx: Context, mapping: Mapping, dtype=np.float32) -> DeviceField:
    lay = mapping.c_layout()
    h = C.c_void_p()
    ctx.check(N.lib().sct_field_create(ctx.handle, C.byref(lay), _dt(dtype), C.byref(h)))
    return DeviceField(ctx, h, mapping, dtype)


def download(ctx: Context, dfield: DeviceField, out: Field) -> None:
    """download(ctx, field, mdspan), cuda-context.cpp:59-77 (raises on span mismatch)."""
    if out.data.dtype != dfield.dtype:
        raise ValueError("dtype mismatch")
    ctx.check(N.lib().sct_download(ctx.handle, dfield.handle, out.data.ctypes.data, out.data.size))


# ---------------------------------------------------------------------------
# Potts
# ---------------------------------------------------------------------------
def _opts(storage="fp32", kernel="auto", slab_w=None, mem_available=None, dq=None,
          gap=None, gap_every=0, gap_tol=0.0, iters_done=None) -> N.SctPottsOpts:
    """gap: host float64 array receiving the relative gap every gap_every iterations;
    gap_tol > 0 stops at the first evaluation below it, iters_done (c_int32)
    receiving the iterations performed."""
    o = N.SctPottsOpts()
    N.lib().sct_potts_opts_init(C.byref(o))
    o.storage = N.STORAGE[storage]
    o.kernel = N.KERNEL[kernel]
    o.slab_w = 0 if slab_w is None else int(slab_w)
    o.mem_available = -1 if mem_available is None else int(mem_available)
    if dq is not None:
        o.dq_trace = dq.ctypes.data_as(N.f32p)
    if gap is not None:
        o.gap_trace = gap.ctypes.data_as(N.f64p)
        o.gap_capacity = int(gap.size)
    o.gap_every = int(gap_every)
    o.gap_tol = float(gap_tol)
    if iters_done is not None:
        o.iters_done = C.pointer(iters_done)
    return o


def potts_min_partition(ctx: Context, g: DeviceField, ci, lam: float, tau1: float, tau2: float,
                        iters: int, *, storage="fp32", kernel="auto", slab_w=None,
                        mem_available=None, dq_trace=False, gap_every=0, gap_tol=0.0):
    """potts_min_partition<K>(ctx, g, cpu_map, ci, lambda, tau1, tau2, iters).

    Returns a DeviceField of c[argmax] labels, or None when the slab cannot
    hold the halo (the reference's null return, potts-min-partition.cpp:571-577).
    With dq_trace=True returns (field, dq[iters]); with gap_every > 0 the
    relative FP64 gap after every gap_every-th iteration is appended to the
    returned tuple.  gap_tol > 0 (with gap_every > 0) stops at the first
    evaluation whose relative gap is below gap_tol and appends the number of
    iterations performed.
    """
    c = np.ascontiguousarray(ci, np.float32)
    K = c.size
    if not 2 <= K <= max_phases:
        raise ValueError("K must be in [2, 6]")
    dq = np.zeros(max(iters, 1), np.float32) if dq_trace else None
    gap = np.full(max(iters // gap_every, 1), np.nan) if gap_every > 0 else None
    done = C.c_int32(-1)
    o = _opts(storage, kernel, slab_w, mem_available, dq, gap, gap_every, gap_tol, done if gap_tol > 0 else None)
    h = C.c_void_p()
    rc = N.lib().sct_potts_min_partition(ctx.handle, g.handle, K, c.ctypes.data_as(N.f32p), lam, tau1,
                                         tau2, iters, C.byref(o), C.byref(h))
    n_gap = iters // gap_every if gap is not None else 0
    if gap is not None and gap_tol > 0 and done.value >= 0:
        n_gap = done.value // gap_every
    extra = ([dq[:iters]] if dq_trace else []) + ([gap[:n_gap]] if gap is not None else []) + \
        ([done.value] if gap_tol > 0 else [])
    if rc == N.SCT_ERR_NOMEM:
        return (None, *[None] * len(extra)) if extra else None
    ctx.check(rc)
    res = DeviceField(ctx, h, g.mapping)
    return (res, *extra) if extra else res


def convert_to_fp(ctx: Context, f: DeviceField, lo: float, hi: float) -> DeviceField:
    """convert_to_fp<ElementType>(ctx, in, lo, hi), convert_to_fp.cpp:48-71."""
    h = C.c_void_p()
    ctx.check(N.lib().sct_convert_to_fp(ctx.handle, f.handle, lo, hi, C.byref(h)))
    return DeviceField(ctx, h, f.mapping)


def find_intensity_range(ctx: Context, f: DeviceField, tail_thr: float, *, hist: bool = False):
    """find_intensity_range(ctx, field, tail_thr), find_intensity_range.cpp:63-110.
    Returns (lo, hi), or (lo, hi, counts[65536]) with hist=True."""
    lo, hi = C.c_uint16(), C.c_uint16()
    h = np.zeros(65536, np.uint64) if hist else None
    ctx.check(N.lib().sct_find_intensity_range(ctx.handle, f.handle, tail_thr, C.byref(lo), C.byref(hi),
                                               h.ctypes.data_as(C.POINTER(C.c_uint64)) if hist else None))
    return (lo.value, hi.value, h) if hist else (lo.value, hi.value)


def potts_batch(ctx: Context, problems, tau1: float, tau2: float, iters: int, *, storage="fp32",
                graph=True, autotune=False, persistent=False) -> tuple[list, float]:
    """Many Potts solves in shared launches (no counterpart in the reference).

    problems: sequence of (DeviceField g, labels c, lambda).  Returns the
    label fields (one per problem, as potts_min_partition would return them)
    and the event time of the iterations in ms; graph=True runs all
    iterations as one captured hipGraph launch; persistent=True runs each K
    group's iterations in one persistent launch (tiles from one ordered
    ticket queue, each half-step waiting for the one before); autotune=True first picks each K
    group's Xi workgroup tile and, persistent, its workgroups per CU by timing
    (potts_batch_tiles and potts_batch_occupancy report them)."""
    n = len(problems)
    arr = (N.SctPottsProblem * max(n, 1))()
    for i, (g, c, lam) in enumerate(problems):
        c = np.ascontiguousarray(c, np.float32)
        if not 2 <= c.size <= max_phases:
            raise ValueError("K must be in [2, 6]")
        arr[i].g = g.handle
        arr[i].K = c.size
        for k in range(c.size):
            arr[i].c[k] = float(c[k])
        arr[i].lam = lam
    outs = (C.c_void_p * max(n, 1))()
    ms = C.c_float()
    ctx.check(N.lib().sct_potts_batch(ctx.handle, n, arr, tau1, tau2, iters, N.STORAGE[storage],
                                      (N.SCT_BATCH_GRAPH if graph else 0) | (N.SCT_BATCH_AUTOTUNE if autotune else 0)
                                      | (N.SCT_BATCH_PERSISTENT if persistent else 0),
                                      outs, C.byref(ms)))
    return [DeviceField(ctx, C.c_void_p(outs[i]), problems[i][0].mapping) for i in range(n)], ms.value


def potts_batch_set_tiles(ctx: Context, tiles: dict) -> None:
    """Fix the Xi workgroup tile ({K: waves in 1, 2, 4}) of later potts_batch
    calls; an empty dict returns every K to the default or the autotuner."""
    w = (C.c_int * 5)(*[int(tiles.get(K, 0)) for K in range(2, 7)])
    ctx.check(N.lib().sct_potts_batch_set_tiles(ctx.handle, w))


def potts_batch_set_occupancy(ctx: Context, per_cu: dict) -> None:
    """Fix the workgroups per CU ({K: 1..64}) of later persistent potts_batch
    launches; an empty dict returns every K to the default or the autotuner.
    Counts above what a CU holds give grids that are not resident at once."""
    w = (C.c_int * 5)(*[int(per_cu.get(K, 0)) for K in range(2, 7)])
    ctx.check(N.lib().sct_potts_batch_set_occupancy(ctx.handle, w))


def potts_batch_tiles(ctx: Context) -> dict:
    """{K: waves per Xi workgroup} used by the last potts_batch call."""
    w = (C.c_int * 5)()
    ctx.check(N.lib().sct_potts_batch_tiles(ctx.handle, w))
    return {K: w[K - 2] for K in range(2, 7) if w[K - 2]}


def potts_batch_occupancy(ctx: Context) -> dict:
    """{K: workgroups per CU} of the persistent launches of the last
    potts_batch call (0 for every K of a call without persistent=True)."""
    w = (C.c_int * 5)()
    ctx.check(N.lib().sct_potts_batch_occupancy(ctx.handle, w))
    t = potts_batch_tiles(ctx)
    return {K: w[K - 2] for K in t}


def potts_adjoint(ctx: Context, g: DeviceField, target: DeviceField, ci, lam: float, tau1: float, tau2: float,
                  iters: int, *, checkpoint_every: int = 0, want_grad_u: bool = True) -> dict:
    """Gradient of the mean squared error between V (all K components) and the
    one-hot target after iters iterations, with respect to c, lambda and u
    (no counterpart in the reference).  Returns dict(loss, grad_c, grad_lambda,
    grad_u (host span array or None), forward_ms, backward_ms)."""
    c = np.ascontiguousarray(ci, np.float32)
    if not 2 <= c.size <= max_phases:
        raise ValueError("K must be in [2, 6]")
    loss, gl = C.c_double(), C.c_double()
    gc = np.zeros(c.size, np.float64)
    h = C.c_void_p()
    t = np.zeros(2, np.float32)
    ctx.check(N.lib().sct_potts_adjoint(ctx.handle, g.handle, target.handle, c.size, c.ctypes.data_as(N.f32p), lam,
                                        tau1, tau2, iters, checkpoint_every, C.byref(loss),
                                        gc.ctypes.data_as(C.POINTER(C.c_double)), C.byref(gl),
                                        C.byref(h) if want_grad_u else None, t.ctypes.data_as(N.f32p)))
    gu = None
    if want_grad_u:
        f = DeviceField(ctx, h, g.mapping)
        gu = f.to_host()
        f.close()
    return dict(loss=loss.value, grad_c=gc, grad_lambda=gl.value, grad_u=gu, forward_ms=float(t[0]),
                backward_ms=float(t[1]))


class PottsTape:
    """Forward / backward over the unrolled Potts iteration (no counterpart in
    the reference).  forward() runs the iterations from the nearest-label start
    and keeps the checkpoints; backward(dV) maps dL/dV_i (K device fields) to
    dL/dc, dL/dlambda and dL/du, any number of times."""

    def __init__(self, ctx: Context, g: DeviceField, ci, lam: float, tau1: float, tau2: float, iters: int, *,
                 checkpoint_every: int = 0, want_v: bool = True):
        self.ctx, self.g = ctx, g
        c = np.ascontiguousarray(ci, np.float32)
        if not 2 <= c.size <= max_phases:
            raise ValueError("K must be in [2, 6]")
        self.K = c.size
        vh = (C.c_void_p * self.K)()
        h = C.c_void_p()
        t = C.c_float()
        ctx.check(N.lib().sct_potts_tape_forward(ctx.handle, g.handle, self.K, c.ctypes.data_as(N.f32p), lam, tau1,
                                                 tau2, iters, checkpoint_every, vh if want_v else None, C.byref(h),
                                                 C.byref(t)))
        self.handle = h
        self.forward_ms = t.value
        self.V = [DeviceField(ctx, C.c_void_p(vh[i]), g.mapping) for i in range(self.K)] if want_v else None

    def backward(self, dV, *, want_grad_u: bool = True) -> dict:
        if len(dV) != self.K:
            raise ValueError("backward needs K fields")
        arr = (C.c_void_p * self.K)(*[d.handle.value if isinstance(d.handle, C.c_void_p) else d.handle for d in dV])
        gc = np.zeros(self.K, np.float64)
        gl = C.c_double()
        h = C.c_void_p()
        t = C.c_float()
        self.ctx.check(N.lib().sct_potts_tape_backward(self.handle, arr, gc.ctypes.data_as(C.POINTER(C.c_double)),
                                                       C.byref(gl), C.byref(h) if want_grad_u else None, C.byref(t)))
        gu = None
        if want_grad_u:
            f = DeviceField(self.ctx, h, self.g.mapping)
            gu = f.to_host()
            f.close()
        return dict(grad_c=gc, grad_lambda=gl.value, grad_u=gu, backward_ms=t.value)

    def close(self) -> None:
        if self.V:
            for f in self.V:
                f.close()
            self.V = None
        if self.handle:
            N.lib().sct_potts_tape_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PottsSolver:
    """Stepping interface over the same kernels (single slab)."""

    def __init__(self, ctx: Context, g: DeviceField, ci, lam: float, tau1: float, tau2: float, *,
                 storage="fp32", kernel="auto", dq_trace=False, gap_every=0, gap_capacity=0):
        self.ctx, self.g = ctx, g
        self.c = np.ascontiguousarray(ci, np.float32)
        self.K = self.c.size
        self._dq = np.zeros(1, np.float32) if dq_trace else None
        self._gap = np.full(max(gap_capacity, 1), np.nan) if gap_every else None
        self._o = _opts(storage, kernel, None, None, self._dq, self._gap, gap_every)
        h = C.c_void_p()
        ctx.check(N.lib().sct_potts_solver_create(ctx.handle, g.handle, self.K, self.c.ctypes.data_as(N.f32p),
                                                  lam, tau1, tau2, C.byref(self._o), C.byref(h)))
        self.handle = h
        self.iters = 0

    def iterate(self, n: int) -> None:
        self.ctx.check(N.lib().sct_potts_solver_iterate(self.handle, n))
        self.iters += n

    def iterate_timed(self, n: int) -> np.ndarray:
        """iterate(n), returning the duration in ms of every kernel launch,
        shape (n, kernels per iteration): (Xi, V) for the two-pass kernels,
        (fused,) for the fused kernel.  HIP events on the library stream."""
        nk = N.lib().sct_potts_solver_kernels(self.handle)
        out = np.zeros((n, nk), np.float32)
        self.ctx.check(N.lib().sct_potts_solver_iterate_timed(self.handle, n, out.ctypes.data_as(N.f32p)))
        self.iters += n
        return out

    def labels(self) -> DeviceField:
        res = empty_like(self.ctx, self.g.mapping)
        self.ctx.check(N.lib().sct_potts_solver_labels(self.handle, res.handle))
        return res

    def state(self) -> tuple[np.ndarray, np.ndarray]:
        span = self.g.mapping.span
        V = np.zeros((self.K - 1, span), np.float32)
        xi = np.zeros((3 * (self.K - 1), span), np.float32)
        self.ctx.check(N.lib().sct_potts_solver_state(self.handle, V.ctypes.data_as(N.f32p),
                                                      xi.ctypes.data_as(N.f32p)))
        return V, xi

    def gap(self) -> tuple[float, float]:
        """FP64 primal and dual energies (E, D) of the current state."""
        E, D = C.c_double(), C.c_double()
        self.ctx.check(N.lib().sct_potts_solver_gap(self.handle, C.byref(E), C.byref(D)))
        return E.value, D.value

    def gap_trace(self) -> np.ndarray:
        """Relative gaps (E - D) / E recorded every gap_every iterations so far."""
        if self._gap is None:
            return np.zeros(0)
        k = min(self.iters // self._o.gap_every, self._gap.size)
        return self._gap[:k].copy()

    def dq(self) -> np.ndarray:
        out = np.zeros(self.iters, np.float32)
        self.ctx.check(N.lib().sct_potts_solver_dq(self.handle, out.ctypes.data_as(N.f32p), self.iters))
        return out

    def close(self) -> None:
        if self.handle:
            N.lib().sct_potts_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Chambolle and reductions
# ---------------------------------------------------------------------------
def chambolle(ctx: Context, g: DeviceField, lam: float, tau: float, iters: int, *, dq_trace=False):
    """chambolle(ctx, g, lambda, tau, iters): overwrites g (chambolle.cpp:225-276)."""
    dq = np.zeros(max(iters, 1), np.float32) if dq_trace else None
    ctx.check(N.lib().sct_chambolle(ctx.handle, g.handle, lam, tau, iters,
                                    dq.ctypes.data_as(N.f32p) if dq is not None else None))
    return dq[:iters] if dq_trace else None


def min_max(ctx: Context, f: DeviceField) -> tuple[float, float]:
    """min_max(ctx, in), min_max.cpp:58-81."""
    a, b = C.c_float(), C.c_float()
    ctx.check(N.lib().sct_min_max(ctx.handle, f.handle, C.byref(a), C.byref(b)))
    return a.value, b.value


def grad_min_max(ctx: Context, v: DeviceField) -> tuple[float, float]:
    """Min/max over strictly positive v, uv-hist-kernel.cpp:31-71."""
    a, b = C.c_float(), C.c_float()
    ctx.check(N.lib().sct_grad_min_max(ctx.handle, v.handle, C.byref(a), C.byref(b)))
    return a.value, b.value


def calc_hist(ctx: Context, f: DeviceField, *args, seed: int = 0, counts: bool = False):
    """calc_hist, calc_hist.h:8-12.

    calc_hist(ctx, f, nbins, delta)                         -> (density, min, max)
    calc_hist(ctx, f, range_start, range_end, nbins, delta) -> density
    counts=True returns integer counts in place of the density.
    """
    L = N.lib()
    if len(args) in (2, 4):
        nbins = int(args[-2])
        cnt = np.empty(nbins, np.uint32)
        # the density is computed only when it is what the caller asked for
        den = None if counts else np.empty(nbins, np.float64)
        den_p = None if den is None else den.ctypes.data_as(N.f64p)
    if len(args) == 2:
        delta = float(args[1])
        a, b = C.c_float(), C.c_float()
        ctx.check(L.sct_calc_hist_auto(ctx.handle, f.handle, nbins, delta, seed, cnt.ctypes.data_as(N.u32p),
                                       den_p, C.byref(a), C.byref(b)))
        return (cnt if counts else den), a.value, b.value
    if len(args) == 4:
        rs, re_, delta = float(args[0]), float(args[1]), float(args[3])
        ctx.check(L.sct_calc_hist(ctx.handle, f.handle, rs, re_, nbins, delta, seed,
                                  cnt.ctypes.data_as(N.u32p), den_p))
        return cnt if counts else den
    raise TypeError("calc_hist(ctx, f, nbins, delta) or calc_hist(ctx, f, rs, re, nbins, delta)")


def grad_magn(ctx: Context, u: DeviceField) -> DeviceField:
    """grad_magn(ctx, u), grad_magn.cpp:60-79."""
    h = C.c_void_p()
    ctx.check(N.lib().sct_grad_magn(ctx.handle, u.handle, C.byref(h)))
    return DeviceField(ctx, h, u.mapping)


@dataclass
class UVHistReturn:
    """uv-hist.h:11-16."""
    hist: np.ndarray      # uint32 [nbins, stride] view of the flat histogram
    stride: int
    u_start: float
    u_end: float
    v_start: float
    v_end: float


def _uv_shape(flat: np.ndarray, nbins: int, stride: int) -> np.ndarray:
    full = np.zeros(nbins * stride, np.uint32)
    full[: flat.size] = flat
    return full.reshape(nbins, stride)


def uv_hist(ctx: Context, u: DeviceField, v: DeviceField, *args):
    """uv_hist(ctx, u, v, nbins) -> UVHistReturn;
    uv_hist(ctx, u, v, u0, u1, v0, v1, nbins) -> (hist[nbins, stride], stride)."""
    L = N.lib()
    stride = C.c_uint32()
    if len(args) == 1:
        nbins = int(args[0])
        flat = np.zeros((nbins - 1) * get_stride(nbins) + nbins, np.uint32)
        rng = np.zeros(4, np.float32)
        ctx.check(L.sct_uv_hist_auto(ctx.handle, u.handle, v.handle, nbins, flat.ctypes.data_as(N.u32p),
                                     C.byref(stride), rng.ctypes.data_as(N.f32p)))
        return UVHistReturn(_uv_shape(flat, nbins, stride.value), stride.value, *map(float, rng))
    if len(args) == 5:
        u0, u1, v0, v1 = map(float, args[:4])
        nbins = int(args[4])
        flat = np.zeros((nbins - 1) * get_stride(nbins) + nbins, np.uint32)
        ctx.check(L.sct_uv_hist(ctx.handle, u.handle, v.handle, u0, u1, v0, v1, nbins,
                                flat.ctypes.data_as(N.u32p), C.byref(stride)))
        return _uv_shape(flat, nbins, stride.value), stride.value
    raise TypeError("uv_hist(ctx, u, v, nbins) or uv_hist(ctx, u, v, u0, u1, v0, v1, nbins)")
