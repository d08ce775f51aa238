"""ctypes declarations of the C ABI in include/sct.h.

The product path is the HIP library segmentctscans_amd/_lib/libsct.so; loading
fails loudly when it has not been built (there is no fallback).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "_lib"

SCT_OK, SCT_ERR_INVALID, SCT_ERR_HIP, SCT_ERR_SIZE, SCT_ERR_NOMEM = 0, 1, 2, 3, 4
SCT_F32, SCT_U16 = 0, 1
STORAGE = {"fp32": 0, "bf16": 1, "fp16": 2}
KERNEL = {"auto": 0, "twopass": 1, "fused": 2}

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class SctLayout(C.Structure):
    _fields_ = [
        ("n_images", C.c_uint32), ("height", C.c_uint32), ("width", C.c_uint32),
        ("nstride", C.c_uint32), ("col_begin", C.c_uint32), ("col_end", C.c_uint32),
        ("start_row", u32p), ("base", u32p),
    ]


class SctBinInfo(C.Structure):
    _fields_ = [("n_images", C.c_uint32), ("height", C.c_uint32), ("width", C.c_uint32),
                ("nstride", C.c_uint32), ("col_begin", C.c_uint32), ("col_end", C.c_uint32),
                ("span", C.c_uint64), ("data_size", C.c_uint64), ("lo", C.c_float), ("hi", C.c_float),
                ("header_bits", C.c_uint64)]


class SctPottsProblem(C.Structure):
    _fields_ = [("g", C.c_void_p), ("K", C.c_int32), ("c", C.c_float * 6), ("lam", C.c_float)]


SCT_BATCH_GRAPH = 1
SCT_BATCH_AUTOTUNE = 2
SCT_BATCH_PERSISTENT = 4


class SctPottsOpts(C.Structure):
    _fields_ = [
        ("size", C.c_uint32), ("storage", C.c_int32), ("kernel", C.c_int32),
        ("slab_w", C.c_int32), ("mem_available", C.c_int64),
        ("dq_trace", f32p), ("gap_trace", f64p), ("gap_every", C.c_int32),
        ("gap_capacity", C.c_int32), ("gap_tol", C.c_double), ("iters_done", C.POINTER(C.c_int32)),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "sct_cylinder_layout": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                      C.c_double, C.c_uint32, u32p, u32p, u32p, u32p,
                                      C.POINTER(C.c_uint64)]),
    "sct_get_stride": (C.c_uint32, [C.c_uint32]),
    "sct_prev_stride": (C.c_uint32, [C.c_uint32]),
    "sct_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "sct_ctx_destroy": (None, [C.c_void_p]),
    "sct_ctx_last_error": (C.c_char_p, [C.c_void_p]),
    "sct_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "sct_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "sct_device_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sct_field_create": (C.c_int, [C.c_void_p, C.POINTER(SctLayout), C.c_int, C.POINTER(C.c_void_p)]),
    "sct_upload": (C.c_int, [C.c_void_p, C.POINTER(SctLayout), C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sct_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "sct_field_destroy": (None, [C.c_void_p]),
    "sct_field_data": (C.c_void_p, [C.c_void_p]),
    "sct_field_span": (C.c_size_t, [C.c_void_p]),
    "sct_field_data_size": (C.c_uint64, [C.c_void_p]),
    "sct_potts_opts_init": (None, [C.POINTER(SctPottsOpts)]),
    "sct_potts_min_partition": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, f32p, C.c_float, C.c_float,
                                          C.c_float, C.c_int, C.POINTER(SctPottsOpts), C.POINTER(C.c_void_p)]),
    "sct_potts_solver_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, f32p, C.c_float, C.c_float,
                                          C.c_float, C.POINTER(SctPottsOpts), C.POINTER(C.c_void_p)]),
    "sct_potts_solver_iterate": (C.c_int, [C.c_void_p, C.c_int]),
    "sct_potts_solver_kernels": (C.c_int, [C.c_void_p]),
    "sct_bin_parse": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, u32p, u32p]),
    "sct_bin_read": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p]),
    "sct_bin_bound": (C.c_size_t, [C.c_void_p, C.c_int]),
    "sct_bin_write": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t,
                                C.POINTER(C.c_size_t)]),
    "sct_convert_to_fp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.POINTER(C.c_void_p)]),
    "sct_find_intensity_range": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_uint16),
                                           C.POINTER(C.c_uint16), C.POINTER(C.c_uint64)]),
    "sct_potts_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, f32p, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_void_p), f32p]),
    "sct_potts_tape_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, f32p, C.c_float, C.c_float, C.c_float,
                                         C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), f32p]),
    "sct_potts_tape_backward": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_void_p), f32p]),
    "sct_potts_tape_destroy": (None, [C.c_void_p]),
    "sct_samples_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_void_p)]),
    "sct_samples_destroy": (None, [C.c_void_p]),
    "sct_gmm1d_pass": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_double)] * 7),
    "sct_gmm1d_assign": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_double)] * 3 + [C.POINTER(C.c_int32)]),
    "sct_potts_batch_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "sct_potts_batch_set_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "sct_potts_batch_occupancy": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "sct_potts_batch_set_occupancy": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "sct_potts_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_float)]),
    "sct_potts_solver_gap": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "sct_uv_samples_create": (C.c_int, [C.c_void_p, f64p, f64p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "sct_uv_samples_destroy": (None, [C.c_void_p]),
    "sct_bimm_nll": (C.c_int, [C.c_void_p, C.c_int, f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    "sct_selftest_store_round": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p]),
    "sct_selftest_cr_math": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "sct_potts_solver_iterate_timed": (C.c_int, [C.c_void_p, C.c_int, f32p]),
    "sct_potts_solver_labels": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sct_potts_solver_state": (C.c_int, [C.c_void_p, f32p, f32p]),
    "sct_potts_solver_dq": (C.c_int, [C.c_void_p, f32p, C.c_int]),
    "sct_potts_solver_destroy": (None, [C.c_void_p]),
    "sct_chambolle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_int, f32p]),
    "sct_min_max": (C.c_int, [C.c_void_p, C.c_void_p, f32p, f32p]),
    "sct_calc_hist": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                C.c_uint64, u32p, f64p]),
    "sct_calc_hist_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                        C.c_uint64, u32p, f64p]),
    "sct_calc_hist_auto": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_uint64, u32p,
                                     f64p, f32p, f32p]),
    "sct_grad_magn": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sct_grad_min_max": (C.c_int, [C.c_void_p, C.c_void_p, f32p, f32p]),
    "sct_uv_hist": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float,
                              C.c_float, C.c_uint32, u32p, u32p]),
    "sct_uv_hist_auto": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, u32p, u32p, f32p]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libsct.so (raises if it is missing: the HIP path has no fallback)."""
    global _lib
    if _lib is None:
        path = LIB_DIR / "libsct.so"
        if not path.exists():
            raise RuntimeError(
                f"{path} is missing: build the HIP library first "
                "(python -m segmentctscans_amd._build or __graft_entry__.build())")
        L = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_phantom = None


def phantom_lib() -> C.CDLL:
    global _phantom
    if _phantom is None:
        path = LIB_DIR / "libsct_phantom.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run python -m segmentctscans_amd._build")
        _phantom = C.CDLL(str(path))
    return _phantom
