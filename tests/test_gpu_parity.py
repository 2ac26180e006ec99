"""GPU parity: every output of the HIP path (rho, nearest_higher, delta, is_seed,
cluster_index, seeds) is compared bit for bit with the oracle on the same inputs."""
import numpy as np
import pytest

from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd import synthetic
from cluestering_amd.device import Plan, debug_math

from _support import Case, assert_same
from conftest import read_points

pytestmark = pytest.mark.gpu


def _workload_case(w, name):
    return Case(w.coords, w.weight, w.dc, w.rhoc, w.dm, kernel=w.kernel, metric=w.metric,
                wrapped=w.wrapped, sizes=w.sizes, name=name)


def _check(case, ctx, oracle):
    print(f"[case] {case.name} n={case.n} ndim={case.ndim}", flush=True)
    gpu = case.run_gpu(ctx)
    ref = case.run_oracle(oracle)
    assert_same(gpu, ref, case.name)
    seeds = np.nonzero(ref.is_seed)[0]
    np.testing.assert_array_equal(gpu["seeds"], seeds)
    return gpu, ref


def test_device_math_bits(gpu_ctx, oracle_mod):
    rng = np.random.default_rng(7)
    a = np.concatenate([
        rng.uniform(0, 10, 20000), rng.uniform(0, 1e-3, 5000), 10.0 ** rng.uniform(-40, 38, 5000),
        np.array([0.0, 1.0, 2.0, 2.25, 4.0, np.inf], np.float64)]).astype(np.float32)
    b = rng.uniform(0.1, 10, a.size).astype(np.float32)
    s, _, q = debug_math(gpu_ctx, a, b)
    np.testing.assert_array_equal(s.view(np.uint32), np.sqrt(a).view(np.uint32))
    np.testing.assert_array_equal(q.view(np.uint32), (a / b).view(np.uint32))
    x = np.concatenate([rng.uniform(-110, 90, 20000), rng.uniform(-2, 0, 5000),
                        np.array([0.0, -0.0, -87.3, 88.7, 89.0, -104.0, -103.9], np.float64)])
    x = x.astype(np.float32)
    _, e, _ = debug_math(gpu_ctx, x, np.ones_like(x))
    np.testing.assert_array_equal(e.view(np.uint32), oracle_mod.expf_array(x).view(np.uint32))


def _bits_range(lo_bits, hi_bits):
    return np.arange(lo_bits, hi_bits, dtype=np.uint32).view(np.float32)


def test_device_sqrt_correctly_rounded(gpu_ctx):
    """The stages' square root (hardware estimate + rounding test) equals the
    correctly rounded IEEE sqrt. Every float in [1, 4) covers both exponent
    parities of every mantissa; every float in [2^-97, 2^-94) and [2^126, +inf]
    covers the ends of the fast range; then specials and a random sample."""
    ranges = [(0x3F800000, 0x40800000),   # [1, 4)
              (0x0F000000, 0x10800000),   # [2^-97, 2^-94): the switch at 2^-96
              (0x7E800000, 0x7F800001)]   # [2^126, +inf]
    for lo, hi in ranges:
        a = _bits_range(lo, hi)
        s, _, _ = debug_math(gpu_ctx, a, np.ones_like(a))
        bad = np.nonzero(s.view(np.uint32) != np.sqrt(a).view(np.uint32))[0]
        assert bad.size == 0, f"range {lo:#x}: first mismatch at {a[bad[0]]!r}"
    rng = np.random.default_rng(8)
    a = np.concatenate([rng.integers(0, 0x7F800000, 1 << 22, dtype=np.uint32).view(np.float32),
                        np.array([0.0, -0.0, 1e-45, 1e-40, 1.17549435e-38, np.inf], np.float32)])
    s, _, _ = debug_math(gpu_ctx, a, np.ones_like(a))
    np.testing.assert_array_equal(s.view(np.uint32), np.sqrt(a).view(np.uint32))
    neg = np.array([-1.0, -1e-30, -np.inf, np.nan], np.float32)
    s, _, _ = debug_math(gpu_ctx, neg, np.ones_like(neg))
    assert np.isnan(s).all()


def test_blobs2d_16k(gpu_ctx, oracle_mod):
    _check(_workload_case(synthetic.blobs2d_16k(), "blobs2d_16k"), gpu_ctx, oracle_mod)


def test_blobs2d_uniform_weights_64k(gpu_ctx, oracle_mod):
    _check(_workload_case(synthetic.blobs2d_1m(n=1 << 16), "blobs2d_64k"), gpu_ctx, oracle_mod)


def test_detector3d_periodic_exp_64k(gpu_ctx, oracle_mod):
    _check(_workload_case(synthetic.detector3d_1m(n=1 << 16), "detector3d_64k"), gpu_ctx,
           oracle_mod)


def test_batched_gaussian(gpu_ctx, oracle_mod):
    w = synthetic.batch2d_4096(n_events=24, lo=256, hi=8192)
    gpu, ref = _check(_workload_case(w, "batch24"), gpu_ctx, oracle_mod)
    assert gpu["n_clusters"] == ref.n_seeds


@pytest.mark.parametrize("tuning", [(0, -1, 4, 0), (1, -1, 4, 0), (1, 0, 4, 0), (1, 4, 4, 0),
                                    (0, 6, 4, 0), (1, 12, 4, 0), (1, 2, 8, 0), (1, 4, 16, 0),
                                    (0, -1, 16, 0), (1, 4, 1, 1), (1, 6, 1, 32), (1, -1, 2, 8),
                                    (0, 12, 4, 16), (0, -1, 1, 1)])
def test_tunings_do_not_change_results(gpu_ctx, oracle_mod, tuning):
    """Every launch tuning (query grouping on/off, XCD chunking, waves per
    workgroup, persistent waves pulling from the work-item queues) gives the same
    bits: a 2-D gaussian batch with dense tiles and a 3-D periodic exponential
    event. (0, -1, 1, 1): one persistent wave per CU, far fewer than the batch's
    work items, beside the nearest-higher plain scan, which has no persistent
    form and keeps one wave per item."""
    from cluestering_amd import _capi
    group, xcd, waves, persistent = tuning
    gpu_ctx.set_tuning(group_queries=bool(group), xcd_chunk_log2=xcd, block_waves=waves,
                       persistent=persistent)
    try:
        for w, name in [(synthetic.batch2d_4096(n_events=12, lo=4096, hi=30000), "batch12"),
                        (synthetic.detector3d_1m(n=1 << 15), "detector3d_32k")]:
            # No values from an earlier run may survive in the workspace.
            _capi.check(_capi.load().clue_debug_poison(gpu_ctx.handle))
            _check(_workload_case(w, f"{name} tuning={tuning}"), gpu_ctx, oracle_mod)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("kernel", [K.FlatKernel(0.5), K.ExponentialKernel(0.5, 1.0),
                                    K.GaussianKernel(1.0, 1.0), K.FlatKernel(2.0),
                                    K.GaussianKernel(0.7, 1.3), K.ExponentialKernel(1.7, 0.8),
                                    K.GaussianKernel(0.5, 2.0)])
@pytest.mark.parametrize("metric", [None, K.EuclideanMetric(), K.EuclideanMetric([1.0, 2.0]),
                                    K.ManhattanMetric(), K.ChebyshevMetric(),
                                    K.ChebyshevMetric([1.0, 2.0])])
def test_kernels_and_metrics(gpu_ctx, oracle_mod, kernel, metric):
    c, w = read_points("data_4096.csv")
    w = np.random.default_rng(3).uniform(0.5, 1.5, w.size).astype(np.float32)
    _check(Case(c, w, 1.5, 10.0, 1.5, kernel=kernel, metric=metric, name=f"{kernel}-{metric}"),
           gpu_ctx, oracle_mod)


@pytest.mark.parametrize("ndim", list(range(1, 11)))
def test_dimensions(gpu_ctx, oracle_mod, ndim):
    rng = np.random.default_rng(100 + ndim)
    n = 3000
    centres = rng.uniform(0, 20, size=(6, ndim))
    lab = rng.integers(0, 6, n)
    x = (centres[lab] + rng.normal(0, 0.6, size=(n, ndim))).T.astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    _check(Case(x, w, 1.2 + 0.3 * ndim, 3.0, 2.0 + 0.3 * ndim, name=f"dim{ndim}"), gpu_ctx,
           oracle_mod)


def test_options_tags_uncertainty_sigma(gpu_ctx, oracle_mod):
    c, w = read_points("data_4096.csv")
    rng = np.random.default_rng(11)
    n = w.size
    tags = rng.permutation(n).astype(np.uint32)
    u = rng.uniform(0.5, 1.5, n).astype(np.float32)
    sigma = rng.uniform(0.0, 1.0, size=(2, n)).astype(np.float32)
    _check(Case(c, w, 1.5, 10.0, 1.5, tags=tags, name="tags"), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, u=u, name="uncertainty"), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, sigma=sigma, name="sigma"), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, sigma=sigma + 0.5, metric=K.MahalanobisMetric(),
                name="mahalanobis"), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, seed_dc=3.0, name="seed_dc"), gpu_ctx, oracle_mod)


def test_periodic_opposite_angles(gpu_ctx, oracle_mod):
    c, w = read_points("opposite_angles.csv")
    metric = K.PeriodicEuclideanMetric([0.0, float(2 * np.float32(np.pi))])
    gpu, _ = _check(Case(c, w, 0.2, 5.0, 0.2, metric=metric, wrapped=(0, 1), name="periodic"),
                    gpu_ctx, oracle_mod)
    assert gpu["cluster_index"].max() + 1 == 1  # tests/test_clustering.cpp:237-270


def test_tag_tie_break_known_answer(gpu_ctx, oracle_mod):
    # tests/test_clustering.cpp:517-567
    x = np.zeros((2, 2), np.float32)
    w = np.ones(2, np.float32)
    gpu, _ = _check(Case(x, w, 1.0, 0.0, 1.0, name="tie"), gpu_ctx, oracle_mod)
    assert list(gpu["nearest_higher"]) == [1, -1] and list(gpu["is_seed"]) == [0, 1]
    gpu, _ = _check(Case(x, w, 1.0, 0.0, 1.0, tags=np.array([100, 1], np.uint32), name="tie-tags"),
                    gpu_ctx, oracle_mod)
    assert list(gpu["nearest_higher"]) == [-1, 0] and list(gpu["is_seed"]) == [1, 0]


def test_mahalanobis_known_answer(gpu_ctx, oracle_mod):
    # tests/test_mahalanobis.cpp:10-43
    x = np.array([[2.0, 4.0]], np.float32)
    w = np.ones(2, np.float32)
    gpu, _ = _check(Case(x, w, 1.0, 0.3, name="maha-none"), gpu_ctx, oracle_mod)
    assert gpu["cluster_index"].max() + 1 == 2
    gpu, _ = _check(Case(x, w, 1.0, 0.3, metric=K.MahalanobisMetric(),
                         sigma=np.array([[2.0, 2.0]], np.float32), name="maha-sigma"),
                    gpu_ctx, oracle_mod)
    assert gpu["cluster_index"].max() + 1 == 1


def test_edge_cases(gpu_ctx, oracle_mod):
    _check(Case(np.array([[1.0], [2.0]], np.float32), np.ones(1, np.float32), 1.0, 0.5,
                name="single"), gpu_ctx, oracle_mod)
    x = np.full((2, 500), 3.0, np.float32)  # zero range in every dimension
    _check(Case(x, np.ones(500, np.float32), 1.0, 2.0, name="degenerate"), gpu_ctx, oracle_mod)
    c, w = read_points("data_1024.csv")
    for ppt in (1, 7, 100000):
        _check(Case(c, w, 1.5, 10.0, 1.5, points_per_tile=ppt, name=f"ppt{ppt}"), gpu_ctx,
               oracle_mod)
    _check(Case(c, w, 1.5, 1e9, 1.5, name="no-seeds"), gpu_ctx, oracle_mod)


def test_batched_with_empty_and_tiny_events(gpu_ctx, oracle_mod):
    from cluestering_amd.device import Clusterer, PointsDevice
    c, w = read_points("data_1024.csv")
    sizes = np.array([0, 300, 1, 0, 2, 721], np.uint32)
    _, ref = _check(Case(c, w, 1.3, 10.0, 1.3, sizes=sizes, name="ragged"), gpu_ctx, oracle_mod)
    # getSampleAssociations: the event of every cluster, in the caller's numbering.
    pts = PointsDevice.from_host(gpu_ctx, c, w)
    eoc = gpu_ctx.empty(w.size, np.int32)
    k = Clusterer(gpu_ctx, 1.3, 10.0, 1.3).make_clusters_batched(pts, sizes, event_of_cluster=eoc)
    assert k == ref.n_seeds
    np.testing.assert_array_equal(eoc.to_host()[:k], ref.event_of_cluster)


def test_batched_event_size_boundaries(gpu_ctx, oracle_mod):
    """Events on both sides of the seed stage's LDS limit (at most 32,768 points:
    links and labels resolved in LDS; above: jumping passes in memory and the label
    chase), with the event of every cluster, and lone events on both sides of the
    8,192 points up to which a lone event takes the fused per-event stages."""
    from cluestering_amd.device import Clusterer, PointsDevice
    rng = np.random.default_rng(17)
    sizes = np.array([32768, 32769, 45000, 100, 32767], np.uint32)
    parts = [synthetic._event2d(rng, int(n)) for n in sizes]
    c = np.ascontiguousarray(np.concatenate([p[0] for p in parts], axis=1))
    w = np.ascontiguousarray(np.concatenate([p[1] for p in parts]))
    kern = K.GaussianKernel(1.0, 1.0)
    _, ref = _check(Case(c, w, 1.5, 5.0, 2.0, kernel=kern, sizes=sizes, name="lds-boundary"),
                    gpu_ctx, oracle_mod)
    pts = PointsDevice.from_host(gpu_ctx, c, w)
    eoc = gpu_ctx.empty(w.size, np.int32)
    k = Clusterer(gpu_ctx, 1.5, 5.0, 2.0).make_clusters_batched(pts, sizes, None, kern,
                                                              event_of_cluster=eoc)
    assert k == ref.n_seeds
    np.testing.assert_array_equal(eoc.to_host()[:k], ref.event_of_cluster)
    for n in (8192, 8193):
        cn, wn = synthetic._event2d(rng, n)
        _check(Case(np.ascontiguousarray(cn), np.ascontiguousarray(wn), 1.5, 5.0, 2.0, kernel=kern,
                    name=f"lone-{n}"), gpu_ctx, oracle_mod)


def test_batched_reference_fixture(gpu_ctx, oracle_mod):
    # tests/test_batched_clustering.cpp:8-65: 10 copies of data_1024 give 10x the clusters.
    c, w = read_points("data_1024.csv")
    cb, wb = np.tile(c, (1, 10)), np.tile(w, 10)
    gpu, ref = _check(Case(cb, wb, 1.3, 10.0, 1.3, sizes=[1024] * 10, name="batched-fixture"),
                      gpu_ctx, oracle_mod)
    single = Case(c, w, 1.3, 10.0, 1.3).run_gpu(gpu_ctx)
    assert gpu["n_clusters"] == 10 * (single["cluster_index"].max() + 1)


def test_determinism_and_graph_replay(gpu_ctx, oracle_mod):
    w = synthetic.batch2d_4096(n_events=12, lo=256, hi=4096)
    case = _workload_case(w, "replay")
    ref = case.run_oracle(oracle_mod)
    clus = case.clusterer(gpu_ctx)
    pts = case.points(gpu_ctx)
    plan = Plan(clus, pts, w.sizes, case.metric, case.kernel)
    for rep in range(3):
        pts.cluster_index.fill_bytes(0xFF)
        pts.rho.fill_bytes(0)
        plan.launch()
        gpu_ctx.synchronize()
        out = pts.results()
        assert_same(out, ref, f"replay{rep}")
        # eager work between replays
        clus.make_clusters_batched(pts, w.sizes, case.metric, case.kernel)
        assert_same(pts.results(), ref, f"eager{rep}")
    plan.close()


def _blobs(ndim, n, seed, spread=0.6, n_centres=8):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 20, size=(n_centres, ndim))
    lab = rng.integers(0, n_centres, n)
    x = (centres[lab] + rng.normal(0, spread, size=(n, ndim))).T
    return x, rng.uniform(0.5, 1.5, n)


@pytest.mark.parametrize("metric", [None, K.EuclideanMetric([1.0, 2.5, 0.5]), K.ManhattanMetric(),
                                    K.ChebyshevMetric(), K.ChebyshevMetric([2.0, 1.0, 0.5]),
                                    K.PeriodicEuclideanMetric([0.0, 12.0, 0.0]),
                                    K.PeriodicEuclideanMetric([7.0, 9.0, 25.0]),
                                    K.EuclideanMetric([1.0, 0.0, 2.0])])
def test_nh_block_layout_matches_plain_scan(gpu_ctx, oracle_mod, metric):
    """The nearest-higher stage over Morton-ordered candidate blocks with box and
    key culling gives the plain scan's bits, for every metric with a box bound
    (a zero weight included), including flat-kernel densities full of exact ties
    and a batch of events."""
    from cluestering_amd import _capi
    x, w = _blobs(3, 12000, 41)
    wrapped = ()
    if metric is not None and metric.tag == "periodic_euclidean":
        wrapped = tuple(int(p > 0) for p in metric.params)
        for d, p in enumerate(metric.params):
            if p > 0:
                x[d] = np.mod(x[d], p)
    x = x.astype(np.float32)
    cases = [Case(x, w, 1.1, 6.0, 1.6, kernel=K.GaussianKernel(1.0, 1.0), metric=metric,
                  wrapped=wrapped, name=f"{metric}-gauss"),
             Case(x, np.ones_like(w), 0.9, 4.0, 1.4, kernel=K.FlatKernel(1.0), metric=metric,
                  wrapped=wrapped, name=f"{metric}-flat-ties"),
             Case(x, w, 1.1, 6.0, 2.2, seed_dc=2.8, kernel=K.ExponentialKernel(0.5, 1.0),
                  metric=metric, wrapped=wrapped, sizes=[3000, 64, 1, 8935], name=f"{metric}-batch"),
             # A search box wider than the grid: with wrapping, tiles recur in the walk.
             Case(x[:, :3000], w[:3000], 0.8, 2.0, 4.5, kernel=K.FlatKernel(0.5), metric=metric,
                  wrapped=wrapped, name=f"{metric}-wide-box")]
    try:
        for plain in (False, True):
            gpu_ctx.set_tuning(group_queries=True, xcd_chunk_log2=4, block_waves=1,
                               nh_plain_scan=plain)
            for case in cases:
                _capi.check(_capi.load().clue_debug_poison(gpu_ctx.handle))
                case.name = f"{case.name.split(' plain=')[0]} plain={plain}"
                _check(case, gpu_ctx, oracle_mod)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("ndim", [1, 2, 5, 10])
def test_nh_block_layout_dimensions(gpu_ctx, oracle_mod, ndim):
    from cluestering_amd import _capi
    x, w = _blobs(ndim, 6000, 60 + ndim, spread=0.5)
    gpu_ctx.set_tuning(group_queries=True, xcd_chunk_log2=4, block_waves=1)
    try:
        for dtype in (np.float32, np.float64):
            _capi.check(_capi.load().clue_debug_poison(gpu_ctx.handle))
            _check(Case(x, w, 0.9 + 0.3 * ndim, 3.0, 1.2 + 0.3 * ndim, name=f"dim{ndim}-{dtype.__name__}",
                        dtype=dtype), gpu_ctx, oracle_mod)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("plain", [False, True])
def test_duplicates_zero_and_negative_weights(gpu_ctx, oracle_mod, plain, dtype):
    """Exact duplicate points (zero distances, equal densities broken by tag) and
    weights that are zero or negative (densities <= 0, whose keys carry no tag),
    through both nearest-higher scans, bit for bit against the oracle."""
    from cluestering_amd import _capi
    rng = np.random.default_rng(77)
    base = rng.uniform(0, 12, size=(2, 1500)).astype(np.float32)
    x = np.concatenate([base, base[:, :700], base[:, :300]], axis=1)  # duplicates
    n = x.shape[1]
    w = rng.choice(np.array([1.0, 0.0, -0.5, 2.0], np.float32), size=n, p=[0.55, 0.15, 0.15, 0.15])
    gpu_ctx.set_tuning(group_queries=True, xcd_chunk_log2=4, block_waves=1, nh_plain_scan=plain)
    try:
        for kernel in (K.FlatKernel(0.5), K.GaussianKernel(1.0, 1.0)):
            _capi.check(_capi.load().clue_debug_poison(gpu_ctx.handle))
            _check(Case(x, w, 0.9, 0.5, 1.3, kernel=kernel, dtype=dtype,
                        name=f"dups-{kernel} plain={plain} {np.dtype(dtype).name}"),
                   gpu_ctx, oracle_mod)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("plain", [False, True])
def test_large_tiles(gpu_ctx, oracle_mod, plain):
    """Tiles of thousands of points (points_per_tile = 3000): the block scan reads a
    tile's block boxes in several rounds of 64."""
    from cluestering_amd import _capi
    x, w = _blobs(2, 20000, 88, spread=1.5, n_centres=4)
    gpu_ctx.set_tuning(group_queries=True, xcd_chunk_log2=4, block_waves=1, nh_plain_scan=plain)
    try:
        _capi.check(_capi.load().clue_debug_poison(gpu_ctx.handle))
        _check(Case(x, w, 0.7, 4.0, 1.1, kernel=K.GaussianKernel(1.0, 1.0), points_per_tile=3000,
                    name=f"large-tiles plain={plain}"), gpu_ctx, oracle_mod)
    finally:
        gpu_ctx.set_tuning()
