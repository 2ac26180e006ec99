"""Helpers shared by the tests: running one case on the GPU (through the C ABI)
and on the oracle with the same inputs and parameters, and comparing them."""
from __future__ import annotations

import numpy as np

from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd import _capi

KERNEL_NAMES = {_capi.KERNEL_FLAT: "flat", _capi.KERNEL_EXP: "exp", _capi.KERNEL_GAUSS: "gaus"}
METRIC_NAMES = {
    _capi.METRIC_EUCLIDEAN: "euclidean",
    _capi.METRIC_WEIGHTED_EUCLIDEAN: "weighted_euclidean",
    _capi.METRIC_PERIODIC_EUCLIDEAN: "periodic_euclidean",
    _capi.METRIC_MANHATTAN: "manhattan",
    _capi.METRIC_CHEBYSHEV: "chebyshev",
    _capi.METRIC_WEIGHTED_CHEBYSHEV: "weighted_chebyshev",
    _capi.METRIC_MAHALANOBIS: "mahalanobis",
}


class Case:
    """One clustering problem: points, parameters, kernel and metric."""

    def __init__(self, coords, weight, dc, rhoc, dm=None, seed_dc=None, kernel=None,
                 metric=None, wrapped=(), sigma=None, u=None, tags=None, sizes=None,
                 points_per_tile=0, name="case", dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.coords = np.ascontiguousarray(coords, self.dtype)
        if self.coords.ndim == 1:
            self.coords = self.coords.reshape(1, -1)
        self.weight = np.ascontiguousarray(weight, self.dtype)
        self.dc, self.rhoc = float(dc), float(rhoc)
        self.dm = float(dc if dm is None else dm)
        self.seed_dc = float(dc if seed_dc is None else seed_dc)
        self.kernel = kernel if kernel is not None else K.FlatKernel(0.5)
        self.metric = metric  # None: plain Euclidean
        self.wrapped = tuple(wrapped)
        self.sigma = None if sigma is None else np.ascontiguousarray(sigma, self.dtype)
        self.u = None if u is None else np.ascontiguousarray(u, self.dtype)
        self.tags = tags
        self.sizes = sizes
        self.points_per_tile = points_per_tile
        self.name = name

    @property
    def ndim(self):
        return self.coords.shape[0]

    @property
    def n(self):
        return self.coords.shape[1]

    def metric_ids(self):
        if self.metric is None:
            return _capi.METRIC_EUCLIDEAN, []
        return self.metric.resolve(self.ndim)

    def oracle_config(self, oracle):
        mid, mp = self.metric_ids()
        return oracle.Config(ndim=self.ndim, dc=self.dc, rhoc=self.rhoc, dm=self.dm,
                             seed_dc=self.seed_dc, kernel=KERNEL_NAMES[self.kernel.kind],
                             kparams=self.kernel.params, metric=METRIC_NAMES[mid],
                             mparams=[float(self.dtype.type(x)) for x in mp], wrapped=self.wrapped,
                             points_per_tile=self.points_per_tile or 128)

    def run_oracle(self, oracle, omp=True):
        cfg = self.oracle_config(oracle)
        if self.sizes is None:
            return oracle.run(cfg, self.coords, self.weight, sigma=self.sigma, u=self.u,
                              tags=self.tags, omp=omp, dtype=self.dtype)
        if self.dtype == np.float32:
            return oracle.run_batched(cfg, self.coords, self.weight, self.sizes, sigma=self.sigma,
                                      u=self.u, tags=self.tags, omp=omp)
        return self._run_oracle_per_event(oracle, cfg, omp)

    def _run_oracle_per_event(self, oracle, cfg, omp):
        """Batched semantics from single-event oracle runs: each event alone, ids
        global and event-major, nearest-higher as packed indices, default tags the
        packed index (BatchedClusteringKernels.hpp:183-187, 291-336)."""
        outs, first, off = [], 0, 0
        for n in (int(x) for x in self.sizes):
            sl = slice(off, off + n)
            tags = self.tags[sl] if self.tags is not None else np.arange(off, off + n, dtype=np.uint32)
            r = oracle.run(cfg, self.coords[:, sl], self.weight[sl],
                           sigma=None if self.sigma is None else self.sigma[:, sl],
                           u=None if self.u is None else self.u[sl], tags=tags, omp=omp,
                           dtype=self.dtype)
            nh = np.where(r.nh >= 0, r.nh + off, -1).astype(np.int32)
            cl = np.where(r.cluster >= 0, r.cluster + first, -1).astype(np.int32)
            outs.append((r.rho, nh, r.delta, r.is_seed, cl))
            first += r.n_seeds
            off += n
        cat = [np.concatenate([o[k] for o in outs]) for k in range(5)]
        return oracle.Result(cat[0], cat[1], cat[2], cat[3], cat[4], first, None, None)

    def clusterer(self, ctx):
        from cluestering_amd.device import Clusterer
        c = Clusterer(ctx, self.dc, self.rhoc, self.dm, self.seed_dc)
        c.set_wrapped_coordinates(self.wrapped)
        c.points_per_tile = self.points_per_tile
        return c

    def points(self, ctx):
        from cluestering_amd.device import PointsDevice
        return PointsDevice.from_host(ctx, self.coords, self.weight, sigma=self.sigma, u=self.u,
                                      tags=self.tags, dtype=self.dtype)

    def run_gpu(self, ctx):
        c = self.clusterer(ctx)
        p = self.points(ctx)
        if self.sizes is None:
            c.make_clusters(p, self.metric, self.kernel)
            nclus = None
        else:
            nclus = c.make_clusters_batched(p, self.sizes, self.metric, self.kernel)
        out = p.results()
        out["n_clusters"] = nclus
        out["seeds"] = c.get_seeds()
        return out


def assert_same(gpu: dict, ref, name: str = "") -> None:
    """Bit-exact comparison of every output (rho and delta by their bits)."""
    where = f" [{name}]" if name else ""
    for key, a, b in (("rho", gpu["rho"], ref.rho), ("delta", gpu["delta"], ref.delta)):
        assert a.dtype == b.dtype, f"{key}: gpu {a.dtype} oracle {b.dtype}{where}"
        bits = np.uint32 if a.dtype == np.float32 else np.uint64
        bad = np.nonzero(a.view(bits) != b.view(bits))[0]
        assert bad.size == 0, (f"{key} differs at {bad.size} points{where}, first {bad[:5]}: "
                               f"gpu {a[bad[:5]]} oracle {b[bad[:5]]}")
    for key, a, b in (("nearest_higher", gpu["nearest_higher"], ref.nh),
                      ("is_seed", gpu["is_seed"], ref.is_seed),
                      ("cluster_index", gpu["cluster_index"], ref.cluster)):
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, (f"{key} differs at {bad.size} points{where}, first {bad[:5]}: "
                               f"gpu {a[bad[:5]]} oracle {b[bad[:5]]}")
