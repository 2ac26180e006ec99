"""Double precision (Clusterer<Ndim, double>, the precision of the reference's
own tests/test_clustering.cpp): every output of the f64 HIP path, compared bit
for bit with the oracle's double-precision run on genuinely double inputs."""
import numpy as np
import pytest

from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd import synthetic
from cluestering_amd.device import debug_math_f64

from _support import Case, assert_same
from conftest import read_points

pytestmark = pytest.mark.gpu

F64 = np.float64


def _check(case, ctx, oracle):
    print(f"[case] {case.name} n={case.n} ndim={case.ndim} f64", flush=True)
    gpu = case.run_gpu(ctx)
    ref = case.run_oracle(oracle)
    assert gpu["rho"].dtype == F64
    assert_same(gpu, ref, case.name)
    np.testing.assert_array_equal(gpu["seeds"], np.nonzero(ref.is_seed)[0])
    return gpu, ref


def _jitter(rng, c):
    """Coordinates that no float32 represents: the f64 path must keep them."""
    return np.asarray(c, F64) + rng.uniform(-1e-9, 1e-9, np.shape(c))


def test_device_math_f64_bits(gpu_ctx, oracle_mod):
    rng = np.random.default_rng(17)
    a = np.concatenate([rng.uniform(0, 10, 20000), 10.0 ** rng.uniform(-300, 300, 20000),
                        np.array([0.0, 1.0, 2.0, 4.0, 1e-310, np.inf])])
    b = rng.uniform(0.1, 10, a.size)
    s, _, q = debug_math_f64(gpu_ctx, a, b)
    np.testing.assert_array_equal(s.view(np.uint64), np.sqrt(a).view(np.uint64))
    np.testing.assert_array_equal(q.view(np.uint64), (a / b).view(np.uint64))
    x = np.concatenate([rng.uniform(-746, 710, 40000), rng.uniform(-2, 0, 10000),
                        rng.uniform(-745.2, -708, 5000),
                        np.array([0.0, -0.0, 709.78, 709.8, -745.13, -745.2, np.inf, -np.inf])])
    _, e, _ = debug_math_f64(gpu_ctx, x, np.ones_like(x))
    np.testing.assert_array_equal(e.view(np.uint64), oracle_mod.exp_f64_array(x).view(np.uint64))


def test_f64_blobs_flat_and_gaussian(gpu_ctx, oracle_mod):
    rng = np.random.default_rng(31)
    w = synthetic.blobs2d_16k()
    c = _jitter(rng, w.coords)
    wt = rng.uniform(0.5, 1.5, c.shape[1])
    _check(Case(c, wt, 1.5, 5.0, 2.0, kernel=K.FlatKernel(0.5), name="flat", dtype=F64),
           gpu_ctx, oracle_mod)
    _check(Case(c, wt, 1.5, 5.0, 2.0, kernel=K.GaussianKernel(1.0, 1.0), name="gauss", dtype=F64),
           gpu_ctx, oracle_mod)
    _check(Case(c, wt, 1.5, 5.0, 2.0, kernel=K.GaussianKernel(0.7, 1.3), name="gauss-np2",
                dtype=F64), gpu_ctx, oracle_mod)


def test_f64_detector3d_periodic_exp(gpu_ctx, oracle_mod):
    w = synthetic.detector3d_1m(n=1 << 15)
    rng = np.random.default_rng(32)
    c = _jitter(rng, w.coords)
    _check(Case(c, w.weight, w.dc, w.rhoc, w.dm, kernel=w.kernel, metric=w.metric,
                wrapped=w.wrapped, name="detector3d", dtype=F64), gpu_ctx, oracle_mod)


@pytest.mark.parametrize("metric", [K.ManhattanMetric(), K.ChebyshevMetric([1.0, 2.0]),
                                    K.EuclideanMetric([1.0, 2.5])])
def test_f64_metrics(gpu_ctx, oracle_mod, metric):
    c, w = read_points("data_4096.csv")
    rng = np.random.default_rng(33)
    _check(Case(_jitter(rng, c), rng.uniform(0.5, 1.5, w.size), 1.5, 10.0, 1.5,
                kernel=K.ExponentialKernel(0.5, 1.0), metric=metric, name=str(metric), dtype=F64),
           gpu_ctx, oracle_mod)


def test_f64_options_tags_uncertainty_sigma(gpu_ctx, oracle_mod):
    c, w = read_points("data_4096.csv")
    rng = np.random.default_rng(34)
    c = _jitter(rng, c)
    n = w.size
    tags = rng.permutation(n).astype(np.uint32)
    u = rng.uniform(0.5, 1.5, n)
    sigma = rng.uniform(0.0, 1.0, size=(2, n))
    _check(Case(c, w, 1.5, 10.0, 1.5, tags=tags, name="tags", dtype=F64), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, u=u, name="uncertainty", dtype=F64), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, sigma=sigma, name="sigma", dtype=F64), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, sigma=sigma + 0.5, metric=K.MahalanobisMetric(),
                name="mahalanobis", dtype=F64), gpu_ctx, oracle_mod)
    _check(Case(c, w, 1.5, 10.0, 1.5, seed_dc=3.0, name="seed_dc", dtype=F64), gpu_ctx,
           oracle_mod)


@pytest.mark.parametrize("ndim", [1, 3, 5, 10])
def test_f64_dimensions(gpu_ctx, oracle_mod, ndim):
    rng = np.random.default_rng(200 + ndim)
    n = 3000
    centres = rng.uniform(0, 20, size=(6, ndim))
    lab = rng.integers(0, 6, n)
    x = (centres[lab] + rng.normal(0, 0.6, size=(n, ndim))).T
    w = rng.uniform(0.5, 1.5, n)
    _check(Case(x, w, 1.2 + 0.3 * ndim, 3.0, 2.0 + 0.3 * ndim, name=f"dim{ndim}", dtype=F64),
           gpu_ctx, oracle_mod)


def test_f64_batched(gpu_ctx, oracle_mod):
    w = synthetic.batch2d_4096(n_events=20, lo=256, hi=8192)
    rng = np.random.default_rng(35)
    case = Case(_jitter(rng, w.coords), w.weight, w.dc, w.rhoc, w.dm, kernel=w.kernel,
                sizes=w.sizes, name="batch20", dtype=F64)
    gpu, ref = _check(case, gpu_ctx, oracle_mod)
    assert gpu["n_clusters"] == ref.n_seeds


def test_f64_tag_tie_break_known_answer(gpu_ctx, oracle_mod):
    # tests/test_clustering.cpp:517-567, in double as the reference runs it
    x = np.zeros((2, 2))
    w = np.ones(2)
    gpu, _ = _check(Case(x, w, 1.0, 0.0, 1.0, name="tie", dtype=F64), gpu_ctx, oracle_mod)
    np.testing.assert_array_equal(gpu["nearest_higher"], [1, -1])
    gpu, _ = _check(Case(x, w, 1.0, 0.0, 1.0, tags=np.array([100, 1], np.uint32), name="tie-tags",
                         dtype=F64), gpu_ctx, oracle_mod)
    np.testing.assert_array_equal(gpu["nearest_higher"], [-1, 0])
