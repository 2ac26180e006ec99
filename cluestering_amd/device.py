"""Device-resident API over the C ABI: a queue-like Context, device arrays,
PointsDevice and a Clusterer with the reference's C++ surface
(core/Clusterer.hpp:37-312, data_structures/PointsDevice.hpp:32).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi
from .CLUE_Convolutional_Kernels import EuclideanMetric, FlatKernel, MetricDescriptor


class Context:
    """A device and a HIP stream (the reference's Queue, utils/get_queue.hpp:19-48)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        lib = _capi.load()
        h = C.c_void_p()
        _capi.check(lib.clue_context_create(int(device), C.c_void_p(stream), C.byref(h)))
        self._h = h
        self.device = int(device)
        self._lib = lib

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("context is closed")
        return self._h

    def close(self) -> None:
        if self._h is not None:
            self._lib.clue_context_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        _capi.check(self._lib.clue_context_stream(self.handle, C.byref(s)))
        return s.value or 0

    def synchronize(self) -> None:
        _capi.check(self._lib.clue_synchronize(self.handle))

    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a if dtype is None else a.astype(dtype, copy=False))
        d = DeviceArray(self, a.shape, a.dtype)
        d.copy_from_host(a)
        return d

    def set_tuning(self, group_queries: Optional[bool] = None,
                   xcd_chunk_log2: Optional[int] = None,
                   block_waves: Optional[int] = None,
                   nh_plain_scan: Optional[bool] = None,
                   persistent: Optional[int] = None) -> None:
        """Fix the launch tuning of later runs (speed only; results never change).
        persistent: waves per CU that pull work items from device queues (0: one
        wave per item). With no arguments the built-in table chooses again."""
        if (group_queries is None and xcd_chunk_log2 is None and block_waves is None
                and nh_plain_scan is None and persistent is None):
            _capi.check(self._lib.clue_set_tuning(self.handle, None))
            return
        t = _capi.Tuning()
        t.group_queries = 1 if group_queries or group_queries is None else 0
        t.xcd_chunk_log2 = 4 if xcd_chunk_log2 is None else int(xcd_chunk_log2)
        t.block_waves = 4 if block_waves is None else int(block_waves)
        t.nh_plain_scan = 1 if nh_plain_scan else 0
        t.persistent = 0 if persistent is None else int(persistent)
        _capi.check(self._lib.clue_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self, params: _capi.Params, sizes: Sequence[int]) -> dict:
        """The tuning a run of a batch with these event sizes would use."""
        sz = _capi.as_u32_array(sizes)
        t = _capi.Tuning()
        _capi.check(self._lib.clue_get_tuning(self.handle, C.byref(params), int(sz.sum()),
                                              sz.ctypes.data_as(C.POINTER(C.c_uint32)),
                                              len(sz), C.byref(t)))
        return {"group_queries": bool(t.group_queries), "xcd_chunk_log2": int(t.xcd_chunk_log2),
                "block_waves": int(t.block_waves), "nh_plain_scan": bool(t.nh_plain_scan),
                "persistent": int(t.persistent)}

    def set_profiling(self, enabled: bool) -> None:
        _capi.check(self._lib.clue_set_profiling(self.handle, 1 if enabled else 0))

    def stage_times(self) -> dict:
        names = (C.c_char_p * 16)()
        us = (C.c_float * 16)()
        cnt = C.c_int32()
        _capi.check(self._lib.clue_stage_times(self.handle, names, us, 16, C.byref(cnt)))
        return {names[k].decode(): float(us[k]) for k in range(cnt.value)}


class DeviceArray:
    """A block of device memory with a numpy shape and dtype."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _capi.check(ctx._lib.clue_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            self.ctx._lib.clue_free(self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx._h is not None:
                self.free()
        except Exception:
            pass

    def address(self, offset_elems: int = 0) -> int:
        return self.ptr + offset_elems * self.dtype.itemsize

    def copy_from_host(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _capi.check(self.ctx._lib.clue_memcpy_h2d(self.ctx.handle, C.c_void_p(self.ptr),
                                                  a.ctypes.data_as(C.c_void_p), a.nbytes))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _capi.check(self.ctx._lib.clue_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                                  C.c_void_p(self.ptr), self.nbytes))
        return out

    def fill_bytes(self, value: int) -> None:
        _capi.check(self.ctx._lib.clue_memset_d8(self.ctx.handle, C.c_void_p(self.ptr),
                                                 int(value), self.nbytes))


class PointsDevice:
    """Device SoA points and results (data_structures/PointsDevice.hpp:32,
    device partition detail/DeviceViewPartition.hpp:14-20): ndim coordinate columns,
    weights, and rho / nearest_higher / is_seed / cluster_index outputs, plus the
    optional per-dimension sigma, density uncertainty and tags."""

    def __init__(self, ctx: Context, n_points: int, ndim: int, dtype=np.float32):
        if n_points < 1:
            raise ValueError("the number of points must be positive")
        if not 1 <= ndim <= _capi.MAX_DIM:
            raise IndexError("This library only works up to 10 dimensions")
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise TypeError("points are float32 or float64")
        self.ctx = ctx
        self.n = int(n_points)
        self.ndim = int(ndim)
        self.coords = ctx.empty((ndim, n_points), self.dtype)
        self.weight = ctx.empty(n_points, self.dtype)
        self.cluster_index = ctx.empty(n_points, np.int32)
        self.is_seed = ctx.empty(n_points, np.int32)
        self.rho = ctx.empty(n_points, self.dtype)
        self.nearest_higher = ctx.empty(n_points, np.int32)
        self.delta = ctx.empty(n_points, self.dtype)
        self.sigma: Optional[DeviceArray] = None
        self.density_uncertainty: Optional[DeviceArray] = None
        self.tags: Optional[DeviceArray] = None
        self.clustered = False

    @classmethod
    def from_host(cls, ctx: Context, coords: np.ndarray, weight: np.ndarray, sigma=None, u=None,
                  tags=None, dtype=np.float32) -> "PointsDevice":
        c = np.ascontiguousarray(coords, dtype=dtype)
        if c.ndim == 1:
            c = c.reshape(1, -1)
        p = cls(ctx, c.shape[1], c.shape[0], dtype)
        p.coords.copy_from_host(c)
        p.weight.copy_from_host(np.ascontiguousarray(weight, dtype=dtype))
        if sigma is not None:
            p.set_sigmas(sigma)
        if u is not None:
            p.set_density_uncertainty(u)
        if tags is not None:
            p.set_tags(tags)
        return p

    def set_sigmas(self, sigma: np.ndarray) -> None:
        s = np.ascontiguousarray(sigma, dtype=self.dtype).reshape(self.ndim, self.n)
        self.sigma = self.ctx.to_device(s)

    def set_density_uncertainty(self, u: np.ndarray) -> None:
        self.density_uncertainty = self.ctx.to_device(np.asarray(u, self.dtype).reshape(self.n))

    def set_tags(self, tags: np.ndarray) -> None:
        self.tags = self.ctx.to_device(np.asarray(tags, np.uint32).reshape(self.n))

    @property
    def f64(self) -> bool:
        return self.dtype == np.float64

    def c_points(self):
        pts = _capi.PointsF64() if self.f64 else _capi.Points()
        pts.n = self.n
        for d in range(self.ndim):
            pts.coords[d] = self.coords.address(d * self.n)
            if self.sigma is not None:
                pts.sigma[d] = self.sigma.address(d * self.n)
        pts.weight = self.weight.ptr
        pts.density_uncertainty = self.density_uncertainty.ptr if self.density_uncertainty else None
        pts.tags = self.tags.ptr if self.tags else None
        return pts

    def c_results(self):
        r = _capi.ResultsF64() if self.f64 else _capi.Results()
        r.cluster_index = self.cluster_index.ptr
        r.is_seed = self.is_seed.ptr
        r.rho = self.rho.ptr
        r.nearest_higher = self.nearest_higher.ptr
        r.delta = self.delta.ptr
        return r

    def results(self) -> dict:
        return {
            "cluster_index": self.cluster_index.to_host(),
            "is_seed": self.is_seed.to_host(),
            "rho": self.rho.to_host(),
            "nearest_higher": self.nearest_higher.to_host(),
            "delta": self.delta.to_host(),
        }

    def n_clusters(self) -> int:
        ci = self.cluster_index.to_host()
        return int(ci.max()) + 1 if ci.size else 0


def _kernel_params(kernel) -> tuple:
    return kernel.kind, kernel.params


class Clusterer:
    """clue::Clusterer<Ndim> (core/Clusterer.hpp:37-312) over the C ABI.

    Parameters are validated as in core/detail/Clusterer.hpp:42-48; outlier and
    seeding distances default to the density radius.
    """

    def __init__(self, ctx: Context, density_radius: float, min_density: float,
                 outlier_distance: Optional[float] = None,
                 seeding_distance: Optional[float] = None):
        self.ctx = ctx
        self.wrapped: Sequence[int] = ()
        self.points_per_tile = 0
        self.set_parameters(density_radius, min_density, outlier_distance, seeding_distance)

    def set_parameters(self, density_radius, min_density, outlier_distance=None,
                       seeding_distance=None) -> None:
        dc = float(density_radius)
        dm = dc if outlier_distance is None else float(outlier_distance)
        sdc = dc if seeding_distance is None else float(seeding_distance)
        rhoc = float(min_density)
        if not (dc > 0 and rhoc >= 0 and dm > 0 and sdc > 0):
            raise ValueError("Invalid clustering parameters. The parameters must be positive.")
        self.dc, self.rhoc, self.dm, self.seed_dc = dc, rhoc, dm, sdc

    def set_wrapped_coordinates(self, *wrapped) -> None:
        if len(wrapped) == 1 and not isinstance(wrapped[0], (int, np.integer)):
            wrapped = tuple(wrapped[0])
        self.wrapped = tuple(int(w) for w in wrapped)

    def params(self, ndim: int, metric: Optional[MetricDescriptor] = None,
               kernel=None, dtype=np.float32):
        metric = metric if metric is not None else MetricDescriptor("euclidean_plain")
        kernel = kernel if kernel is not None else FlatKernel(0.5)
        if metric.tag == "euclidean_plain":
            mid, mp = _capi.METRIC_EUCLIDEAN, []
        else:
            mid, mp = metric.resolve(ndim)
        kid, kp = _kernel_params(kernel)
        p = _capi.make_params(ndim, self.dc, self.rhoc, self.dm, self.seed_dc, kid, kp, mid, mp,
                              self.wrapped, self.points_per_tile, dtype)
        lib = _capi.load()
        if np.dtype(dtype) == np.float64:
            _capi.check(lib.clue_validate_params_f64(C.byref(p)))
        else:
            _capi.check(lib.clue_validate_params(C.byref(p)))
        return p

    def make_clusters(self, points: PointsDevice, metric: Optional[MetricDescriptor] = None,
                      kernel=None) -> None:
        """make_clusters(queue, d_points, metric, kernel) (core/detail/Clusterer.hpp:143-151)."""
        p = self.params(points.ndim, metric, kernel, points.dtype)
        pts = points.c_points()
        res = points.c_results()
        lib = _capi.load()
        fn = lib.clue_make_clusters_f64 if points.f64 else lib.clue_make_clusters
        _capi.check(fn(self.ctx.handle, C.byref(p), C.byref(pts), C.byref(res)))
        points.clustered = True

    def make_clusters_batched(self, points: PointsDevice, sizes: Sequence[int],
                              metric: Optional[MetricDescriptor] = None, kernel=None,
                              event_of_cluster: Optional[DeviceArray] = None) -> int:
        """make_clusters(queue, d_points, batch_item_sizes, ...) (core/detail/Clusterer.hpp:171-183).
        Returns the number of clusters."""
        p = self.params(points.ndim, metric, kernel, points.dtype)
        sz = _capi.as_u32_array(sizes)
        pts = points.c_points()
        res = points.c_results()
        nc = C.c_int64()
        lib = _capi.load()
        fn = lib.clue_make_clusters_batched_f64 if points.f64 else lib.clue_make_clusters_batched
        _capi.check(fn(
            self.ctx.handle, C.byref(p), C.byref(pts), sz.ctypes.data_as(C.POINTER(C.c_uint32)),
            len(sz), C.byref(res), C.c_void_p(event_of_cluster.ptr if event_of_cluster else None),
            C.byref(nc)))
        points.clustered = True
        return int(nc.value)

    def get_seeds(self) -> np.ndarray:
        """getSeeds() (core/detail/Clusterer.hpp:198-203)."""
        lib = _capi.load()
        cnt = C.c_int64()
        _capi.check(lib.clue_get_n_clusters(self.ctx.handle, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), np.int32)
        _capi.check(lib.clue_get_seeds(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                       out.size, C.byref(cnt)))
        return out[: cnt.value]


class AssociationMap:
    """Key -> members CSR built on the device (AssociationMap<Device>,
    data_structures/internal/MakeAssociator.hpp:15-45): ``offsets`` (bins + 1)
    and ``indexes`` (the members, ascending within a key) are device arrays."""

    def __init__(self, ctx: Context, keys: DeviceArray, n: Optional[int] = None):
        n = int(np.prod(keys.shape)) if n is None else int(n)
        lib = _capi.load()
        cap = n + 2
        self.offsets = ctx.empty(cap, np.int32)
        self.indexes = ctx.empty(max(n, 1), np.int32)
        bins, members = C.c_int32(), C.c_int32()
        st = lib.clue_make_association(ctx.handle, C.c_void_p(keys.ptr), n,
                                       C.c_void_p(self.offsets.ptr), cap,
                                       C.c_void_p(self.indexes.ptr), C.byref(bins),
                                       C.byref(members))
        if st == _capi.ERR_CAPACITY and bins.value + 1 > cap:   # keys beyond n: size and retry
            cap = bins.value + 1
            self.offsets = ctx.empty(cap, np.int32)
            st = lib.clue_make_association(ctx.handle, C.c_void_p(keys.ptr), n,
                                           C.c_void_p(self.offsets.ptr), cap,
                                           C.c_void_p(self.indexes.ptr), C.byref(bins),
                                           C.byref(members))
        _capi.check(st)
        self.n_bins = int(bins.value)
        self.n_members = int(members.value)

    def size(self) -> int:
        return self.n_bins

    def to_host(self):
        """(offsets, indexes) as host arrays."""
        off = self.offsets.to_host()[: self.n_bins + 1]
        idx = self.indexes.to_host()[: self.n_members]
        return off, idx


def get_clusters(points: "PointsDevice") -> AssociationMap:
    """get_clusters(queue, d_points) (utils/detail/get_clusters.hpp:44-48): the
    cluster -> points map of clustered device points, built on the device."""
    if not points.clustered:
        raise RuntimeError("the points have not been clustered")
    return AssociationMap(points.ctx, points.cluster_index, points.n)


class Plan:
    """A batch captured once in a hipGraph and replayed with one launch."""

    def __init__(self, clusterer: Clusterer, points: PointsDevice, sizes: Sequence[int],
                 metric: Optional[MetricDescriptor] = None, kernel=None,
                 event_of_cluster: Optional[DeviceArray] = None):
        self.ctx = clusterer.ctx
        self._p = clusterer.params(points.ndim, metric, kernel)
        self._sizes = _capi.as_u32_array(sizes)
        self._pts = points.c_points()
        self._res = points.c_results()
        h = C.c_void_p()
        lib = _capi.load()
        _capi.check(lib.clue_plan_create(
            self.ctx.handle, C.byref(self._p), C.byref(self._pts),
            self._sizes.ctypes.data_as(C.POINTER(C.c_uint32)), len(self._sizes),
            C.byref(self._res), C.c_void_p(event_of_cluster.ptr if event_of_cluster else None),
            C.byref(h)))
        self._h = h
        self._lib = lib

    def launch(self) -> None:
        _capi.check(self._lib.clue_plan_launch(self._h))

    def close(self) -> None:
        if self._h is not None:
            self._lib.clue_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    lib = _capi.load()
    c = C.c_int32()
    st = lib.clue_device_count(C.byref(c))
    return c.value if st == 0 else 0


def debug_math_f64(ctx: Context, a: np.ndarray, b: np.ndarray):
    """Device sqrt(a), exp(a) and a / b in double precision, as the stages compute them."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    da, db = ctx.to_device(a), ctx.to_device(b)
    o1, o2, o3 = (ctx.empty(a.shape, np.float64) for _ in range(3))
    _capi.check(_capi.load().clue_debug_math_f64(ctx.handle, a.size, C.c_void_p(da.ptr),
                                                 C.c_void_p(db.ptr), C.c_void_p(o1.ptr),
                                                 C.c_void_p(o2.ptr), C.c_void_p(o3.ptr)))
    return o1.to_host(), o2.to_host(), o3.to_host()


def debug_math(ctx: Context, a: np.ndarray, b: np.ndarray):
    """Device sqrtf(a), expf(a) and a / b, for pinning device math bits."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    da, db = ctx.to_device(a), ctx.to_device(b)
    o1, o2, o3 = (ctx.empty(a.shape, np.float32) for _ in range(3))
    _capi.check(_capi.load().clue_debug_math(ctx.handle, a.size, C.c_void_p(da.ptr),
                                             C.c_void_p(db.ptr), C.c_void_p(o1.ptr),
                                             C.c_void_p(o2.ptr), C.c_void_p(o3.ptr)))
    return o1.to_host(), o2.to_host(), o3.to_host()
