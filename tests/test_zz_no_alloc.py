"""The allocation-free entry point (clue_make_clusters_noalloc) against
clue_make_clusters, through the C ABI: every output bit for bit, on two calls in
a row with the second input smaller (the shared scratch keeps the first call's
contents), and an event beyond the scratch refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd import _capi, synthetic
from cluestering_amd.device import Clusterer, PointsDevice

pytestmark = pytest.mark.gpu


def _run(ctx, fn, clus, pts, metric, kernel):
    lib = _capi.load()
    p = clus.params(pts.ndim, metric, kernel)
    cp, cr = pts.c_points(), pts.c_results()
    for arr in (pts.cluster_index, pts.is_seed, pts.rho, pts.nearest_higher, pts.delta):
        arr.fill_bytes(0xA5)  # neither call may pass on the other's outputs
    _capi.check(getattr(lib, fn)(ctx.handle, C.byref(p), C.byref(cp), C.byref(cr)))
    out = pts.results()
    seeds = Clusterer.get_seeds(clus)
    return out, seeds


def _same(a, b, what):
    for k in ("cluster_index", "is_seed", "nearest_higher"):
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"
    for k in ("rho", "delta"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), f"{what}: {k} bits differ"


def test_noalloc_equals_make_clusters_twice_then_smaller(gpu_ctx):
    big = synthetic.blobs2d_1m(n=1 << 19)                 # 2-D, flat kernel, 524288 points
    small = synthetic.detector3d_1m(n=40000)              # 3-D periodic, exponential kernel
    print(f"[noalloc] cases: {big.name} n={big.n}, {small.name} n={small.n}", flush=True)
    for w in (big, small):
        pts = PointsDevice.from_host(gpu_ctx, w.coords, w.weight)
        clus = Clusterer(gpu_ctx, w.dc, w.rhoc, w.dm)
        clus.set_wrapped_coordinates(w.wrapped)
        ref, ref_seeds = _run(gpu_ctx, "clue_make_clusters", clus, pts, w.metric, w.kernel)
        got, got_seeds = _run(gpu_ctx, "clue_make_clusters_noalloc", clus, pts, w.metric, w.kernel)
        _same(got, ref, w.name)
        assert np.array_equal(got_seeds, ref_seeds)
        assert ref["cluster_index"].max() > 0


def test_noalloc_refuses_events_beyond_the_scratch(gpu_ctx):
    n = 3 << 20  # more than the 160 MiB scratch holds in 2-D
    pts = PointsDevice(gpu_ctx, n, 2)
    clus = Clusterer(gpu_ctx, 1.0, 1.0)
    p = clus.params(2, None, K.FlatKernel(0.5))
    cp, cr = pts.c_points(), pts.c_results()
    st = _capi.load().clue_make_clusters_noalloc(gpu_ctx.handle, C.byref(p), C.byref(cp), C.byref(cr))
    assert st == _capi.ERR_CAPACITY
    assert "scratch" in _capi.load().clue_last_error().decode()
