"""Device association maps (getClusters / get_clusters / getSampleAssociations on
device points) against a host CSR built with numpy: offsets from the key counts,
members in ascending index within a key (the host fill order of
data_structures/detail/AssociationMap.hpp:359-383)."""
import numpy as np
import pytest

from cluestering_amd import synthetic
from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd.device import AssociationMap, Clusterer, PointsDevice, get_clusters

pytestmark = pytest.mark.gpu


def _host_csr(keys):
    keys = np.asarray(keys, np.int64)
    bins = max(int(keys.max()) + 1, 0) if keys.size else 0
    counts = np.bincount(keys[keys >= 0], minlength=bins) if bins else np.zeros(0, np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    members = np.nonzero(keys >= 0)[0]
    idx = members[np.argsort(keys[members], kind="stable")].astype(np.int32)
    return off, idx


def _check(ctx, keys):
    d = ctx.to_device(np.asarray(keys, np.int32))
    m = AssociationMap(ctx, d)
    off, idx = m.to_host()
    ref_off, ref_idx = _host_csr(keys)
    assert m.size() == ref_off.size - 1
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(idx, ref_idx)


def test_association_keys_edge_cases(gpu_ctx):
    rng = np.random.default_rng(3)
    _check(gpu_ctx, [-1, -1, -1])                       # nothing associated: 0 bins
    _check(gpu_ctx, [0])
    _check(gpu_ctx, [5, -1, 5, 0, -1, 2])               # empty bins in between
    _check(gpu_ctx, rng.integers(-1, 3, 100000))        # few large bins
    _check(gpu_ctx, rng.integers(-1, 70000, 300000))    # many bins, 3 radix passes
    keys = rng.integers(-1, 1000, 9000)
    keys[::7] = 999_999                                  # sparse high key
    _check(gpu_ctx, keys)


def test_get_clusters_of_device_points(gpu_ctx):
    w = synthetic.blobs2d_16k()
    pts = PointsDevice.from_host(gpu_ctx, w.coords, w.weight)
    algo = Clusterer(gpu_ctx, w.dc, w.rhoc, w.dm)
    algo.make_clusters(pts, kernel=K.FlatKernel(0.5))
    labels = pts.cluster_index.to_host()
    off, idx = get_clusters(pts).to_host()
    ref_off, ref_idx = _host_csr(labels)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(idx, ref_idx)
    assert off.size - 1 == labels.max() + 1


def test_sample_associations_of_a_batch(gpu_ctx):
    w = synthetic.batch2d_4096(n_events=64)
    pts = PointsDevice.from_host(gpu_ctx, w.coords, w.weight)
    algo = Clusterer(gpu_ctx, w.dc, w.rhoc, w.dm)
    eoc = gpu_ctx.empty(w.coords.shape[1], np.int32)
    nc = algo.make_clusters_batched(pts, w.sizes, kernel=w.kernel, event_of_cluster=eoc)
    m = AssociationMap(gpu_ctx, eoc, nc)
    off, idx = m.to_host()
    labels = pts.cluster_index.to_host()
    bounds = np.concatenate([[0], np.cumsum(np.asarray(w.sizes, np.int64))])
    # Event e owns the clusters whose labels appear among its points, in order.
    for e in range(len(w.sizes)):
        lab = labels[bounds[e]:bounds[e + 1]]
        own = np.unique(lab[lab >= 0])
        np.testing.assert_array_equal(idx[off[e]:off[e + 1]], own)
    assert m.n_members == nc
