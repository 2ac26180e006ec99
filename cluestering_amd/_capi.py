"""ctypes binding of libclue_amd.so (declared in include/clue_amd.h).

The library is the product's only compute path: if it is missing this module
raises at import, and nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; CLUE_AMD_LIBRARY selects another build of the same library.
LIB_PATH = os.environ.get("CLUE_AMD_LIBRARY") or os.path.join(_PKG, "libclue_amd.so")

MAX_DIM = 10

KERNEL_FLAT, KERNEL_EXP, KERNEL_GAUSS = 0, 1, 2
(METRIC_EUCLIDEAN, METRIC_WEIGHTED_EUCLIDEAN, METRIC_PERIODIC_EUCLIDEAN, METRIC_MANHATTAN,
 METRIC_CHEBYSHEV, METRIC_WEIGHTED_CHEBYSHEV, METRIC_MAHALANOBIS) = range(7)

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_OUT_OF_RANGE = 2
ERR_RUNTIME = 3
ERR_HIP = 4
ERR_CAPACITY = 5
ERR_NO_DEVICE = 6


class ClueError(RuntimeError):
    """A HIP or capacity failure inside libclue_amd."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[clue status {status}] {message}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("kernel", C.c_int32),
        ("metric", C.c_int32),
        ("points_per_tile", C.c_int32),
        ("wrapped", C.c_int32 * MAX_DIM),
        ("dc", C.c_float),
        ("rhoc", C.c_float),
        ("dm", C.c_float),
        ("seed_dc", C.c_float),
        ("kparam", C.c_float * 2),
        ("mparam", C.c_float * MAX_DIM),
    ]


class Points(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("coords", C.c_void_p * MAX_DIM),
        ("weight", C.c_void_p),
        ("sigma", C.c_void_p * MAX_DIM),
        ("density_uncertainty", C.c_void_p),
        ("tags", C.c_void_p),
    ]


class Results(C.Structure):
    _fields_ = [
        ("cluster_index", C.c_void_p),
        ("is_seed", C.c_void_p),
        ("rho", C.c_void_p),
        ("nearest_higher", C.c_void_p),
        ("delta", C.c_void_p),
    ]


class ParamsF64(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("kernel", C.c_int32),
        ("metric", C.c_int32),
        ("points_per_tile", C.c_int32),
        ("wrapped", C.c_int32 * MAX_DIM),
        ("dc", C.c_double),
        ("rhoc", C.c_double),
        ("dm", C.c_double),
        ("seed_dc", C.c_double),
        ("kparam", C.c_double * 2),
        ("mparam", C.c_double * MAX_DIM),
    ]


class PointsF64(C.Structure):
    _fields_ = Points._fields_


class ResultsF64(C.Structure):
    _fields_ = Results._fields_


class Tuning(C.Structure):
    _fields_ = [
        ("group_queries", C.c_int32),
        ("xcd_chunk_log2", C.c_int32),
        ("block_waves", C.c_int32),
        ("nh_plain_scan", C.c_int32),
        ("persistent", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("n_lanes", C.c_int32),
        ("hbm_budget_bytes", C.c_int64),
        ("max_batch_points", C.c_int64),
        ("max_batch_events", C.c_int64),
        ("staging_bytes", C.c_int64),
    ]


class ServerTiming(C.Structure):
    _fields_ = [
        ("submit_ns", C.c_int64),
        ("launch_ns", C.c_int64),
        ("done_ns", C.c_int64),
        ("batch_events", C.c_int32),
        ("reserved", C.c_int32),
        ("batch_points", C.c_int64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("events_submitted", C.c_int64),
        ("events_done", C.c_int64),
        ("points_done", C.c_int64),
        ("batches", C.c_int64),
        ("lane_points", C.c_int64),
    ]


_SIGNATURES = {
    "clue_last_error": (C.c_char_p, []),
    "clue_version": (C.c_char_p, []),
    "clue_device_count": (C.c_int32, [C.POINTER(C.c_int32)]),
    "clue_device_name": (C.c_int32, [C.c_int32, C.c_char_p, C.c_int32]),
    "clue_context_create": (C.c_int32, [C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]),
    "clue_context_destroy": (C.c_int32, [C.c_void_p]),
    "clue_context_stream": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "clue_synchronize": (C.c_int32, [C.c_void_p]),
    "clue_validate_params": (C.c_int32, [C.POINTER(Params)]),
    "clue_make_clusters": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Points),
                                       C.POINTER(Results)]),
    "clue_make_clusters_async": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Points),
                                             C.POINTER(Results)]),
    "clue_make_clusters_noalloc": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Points),
                                               C.POINTER(Results)]),
    "clue_make_clusters_batched": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Points),
                                               C.POINTER(C.c_uint32), C.c_int32,
                                               C.POINTER(Results), C.c_void_p,
                                               C.POINTER(C.c_int64)]),
    "clue_make_clusters_host": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_void_p,
                                            C.c_void_p, C.POINTER(C.c_uint32), C.c_int32,
                                            C.c_void_p, C.c_void_p, C.c_void_p]),
    "clue_validate_params_f64": (C.c_int32, [C.POINTER(ParamsF64)]),
    "clue_make_clusters_f64": (C.c_int32, [C.c_void_p, C.POINTER(ParamsF64), C.POINTER(PointsF64),
                                           C.POINTER(ResultsF64)]),
    "clue_make_clusters_batched_f64": (C.c_int32, [C.c_void_p, C.POINTER(ParamsF64),
                                                   C.POINTER(PointsF64), C.POINTER(C.c_uint32),
                                                   C.c_int32, C.POINTER(ResultsF64), C.c_void_p,
                                                   C.POINTER(C.c_int64)]),
    "clue_make_clusters_host_f64": (C.c_int32, [C.c_void_p, C.POINTER(ParamsF64), C.c_int32,
                                                C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                                C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "clue_get_seeds": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "clue_get_n_clusters": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    "clue_make_association": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                          C.c_int32, C.c_void_p, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)]),
    "clue_set_tuning": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "clue_get_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.c_int32,
                                    C.POINTER(C.c_uint32), C.c_int32, C.POINTER(Tuning)]),
    "clue_set_profiling": (C.c_int32, [C.c_void_p, C.c_int32]),
    "clue_stage_times": (C.c_int32, [C.c_void_p, C.POINTER(C.c_char_p), C.POINTER(C.c_float),
                                     C.c_int32, C.POINTER(C.c_int32)]),
    "clue_plan_create": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(Points),
                                     C.POINTER(C.c_uint32), C.c_int32, C.POINTER(Results),
                                     C.c_void_p, C.POINTER(C.c_void_p)]),
    "clue_plan_launch": (C.c_int32, [C.c_void_p]),
    "clue_plan_destroy": (C.c_int32, [C.c_void_p]),
    "clue_malloc": (C.c_int32, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "clue_free": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "clue_memcpy_h2d": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "clue_memcpy_d2h": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "clue_memset_d8": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    "clue_debug_math": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "clue_debug_copy": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "clue_debug_poison": (C.c_int32, [C.c_void_p]),
    "clue_debug_math_f64": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
    "clue_server_create": (C.c_int32, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "clue_server_destroy": (C.c_int32, [C.c_void_p]),
    "clue_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_void_p,
                                       C.POINTER(C.c_uint64)]),
    "clue_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int64,
                                     C.POINTER(ServerTiming)]),
    "clue_server_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    "clue_now_ns": (C.c_int64, []),
}

# Every symbol include/clue_amd.h declares.
EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load libclue_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (the HIP library is the only compute path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().clue_last_error().decode(errors="replace")


def check(status: int) -> None:
    """Map a clue_status to the exception the reference raises for it."""
    if status == OK:
        return
    msg = last_error()
    if status == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if status == ERR_RUNTIME:
        raise RuntimeError(msg)
    raise ClueError(status, msg)


def make_params(ndim: int, dc: float, rhoc: float, dm: Optional[float] = None,
                seed_dc: Optional[float] = None, kernel: int = KERNEL_FLAT,
                kparams: Sequence[float] = (0.5,), metric: int = METRIC_EUCLIDEAN,
                mparams: Sequence[float] = (), wrapped: Sequence[int] = (),
                points_per_tile: int = 0, dtype=np.float32):
    """clue_params, or clue_params_f64 for dtype float64 (Clusterer<Ndim, double>)."""
    p = ParamsF64() if np.dtype(dtype) == np.float64 else Params()
    p.ndim = int(ndim)
    p.kernel = int(kernel)
    p.metric = int(metric)
    p.points_per_tile = int(points_per_tile)
    for d, w in enumerate(list(wrapped)[:MAX_DIM]):
        p.wrapped[d] = int(w)
    p.dc = float(dc)
    p.rhoc = float(rhoc)
    p.dm = float(dc if dm is None else dm)
    p.seed_dc = float(dc if seed_dc is None else seed_dc)
    for k, v in enumerate(list(kparams)[:2]):
        p.kparam[k] = float(v)
    for k, v in enumerate(list(mparams)[:MAX_DIM]):
        p.mparam[k] = float(v)
    return p


def as_u32_array(sizes) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(sizes, dtype=np.uint32))
