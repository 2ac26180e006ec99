"""Drop-in for the reference's ``CLUE_GPU_HIP`` nanobind module
(CLUEstering/BindingModules/hip/binding_gpu_hip.cpp:15-28, Run.hpp:25-142),
running on libclue_amd.so.

mainRun takes exactly the reference's arguments and writes the cluster ids into
``results`` in place.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi
from .CLUE_Convolutional_Kernels import MetricDescriptor

_contexts: dict = {}


def _context(device_id: int):
    from .device import Context
    ctx = _contexts.get(device_id)
    if ctx is None:
        ctx = Context(device_id)
        _contexts[device_id] = ctx
    return ctx


def listDevices(backend: str) -> None:
    """Prints the HIP devices (Run.hpp:53-65)."""
    lib = _capi.load()
    n = C.c_int32(0)
    lib.clue_device_count(C.byref(n))
    if n.value == 0:
        print(f"No devices found for the {backend} backend.")
        return
    print(f"{backend} devices found: ")
    buf = C.create_string_buffer(256)
    for i in range(n.value):
        lib.clue_device_name(i, buf, 256)
        print(f"\tdevice {i}: {buf.value.decode()}")


def mainRun(dc: float, rhoc: float, dm: float, seed_dc: float, wrapped: Sequence[int],
            data: np.ndarray, results: np.ndarray, kernel, Ndim: int,
            batch_sample_sizes: Optional[np.ndarray], n_points: int, device_id: int,
            metric_desc: MetricDescriptor) -> None:
    """Clusters ``data`` ((Ndim+1) x n_points: coordinate rows then the weight row)
    and writes the labels into ``results`` (int32[n_points]) in place.

    Follows Run.hpp:69-142: Ndim outside 1..10 prints a message and returns; the
    metric descriptor is resolved per MetricDescriptor.hpp:30-78 (empty weights
    mean unit weights)."""
    if not 1 <= int(Ndim) <= _capi.MAX_DIM:
        print("This library only works up to 10 dimensions")
        return
    Ndim = int(Ndim)
    n_points = int(n_points)
    d = np.ascontiguousarray(data, dtype=np.float32)
    if d.size < (Ndim + 1) * n_points:
        raise ValueError("data must hold (Ndim + 1) x n_points values")
    d = d.reshape(-1)[: (Ndim + 1) * n_points]
    if not (isinstance(results, np.ndarray) and results.dtype == np.int32
            and results.flags.c_contiguous and results.flags.writeable
            and results.size >= n_points):
        raise TypeError("results must be a writable C-contiguous int32 array of n_points")
    mid, mp = metric_desc.resolve(Ndim)
    p = _capi.make_params(Ndim, dc, rhoc, dm, seed_dc, kernel.kind, kernel.params, mid, mp,
                          list(wrapped)[:Ndim])
    ctx = _context(int(device_id))
    lib = _capi.load()
    if batch_sample_sizes is not None:
        sz = _capi.as_u32_array(batch_sample_sizes)
        sp, ne = sz.ctypes.data_as(C.POINTER(C.c_uint32)), len(sz)
    else:
        sz, sp, ne = None, None, 0
    _capi.check(lib.clue_make_clusters_host(ctx.handle, C.byref(p), n_points,
                                            d.ctypes.data_as(C.c_void_p),
                                            results.ctypes.data_as(C.c_void_p), sp, ne, None, None,
                                            None))
