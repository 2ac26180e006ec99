"""The reference's own known answers and data fixtures through the HIP path.

- Metric values of tests/test_distance_metrics.cpp:8-63 (and hand-computed
  periodic and Mahalanobis values), read as delta of a 2-point run; the GPU bits
  must equal the value and the oracle's.
- The silhouette thresholds of tests/test_clustering.cpp:14-235 on the data files
  the reference ships, in single and double precision, with every output bit for
  bit against the oracle, and the cluster counts of data/truth_files.
"""
import numpy as np
import pandas as pd
import pytest

from cluestering_amd import CLUE_Convolutional_Kernels as K

from _support import Case, assert_same
from conftest import DATA, read_points, silhouette

pytestmark = pytest.mark.gpu

# The expected value is a square root of an exact value (correctly rounded in the
# precision of the run) or exact.
METRIC_CASES = [
    ("euclidean", (1, 2), (4, 6), None, (), None, 25.0),
    ("weighted_euclidean", (1, 2), (4, 6), K.EuclideanMetric([1.0, 2.0]), (), None, 41.0),
    ("manhattan", (-1, 3), (0, 6), K.ManhattanMetric(), (), None, -4.0),
    ("chebyshev", (-1, 4), (0, 6), K.ChebyshevMetric(), (), None, -2.0),
    ("weighted_chebyshev", (-1, 4), (0, 6), K.ChebyshevMetric([1.0, 2.0]), (), None, -4.0),
    ("periodic", (1, 1), (9, 6), K.PeriodicEuclideanMetric([10.0, 8.0]), (1, 1), None, 13.0),
    ("mahalanobis", (2,), (4,), K.MahalanobisMetric(), (), [[2.0, 2.0]], 0.5),
]


def _expected(v, dtype):
    """-x: the exact value x; x: sqrt(x) rounded in dtype."""
    return dtype(-v) if v < 0 else np.sqrt(dtype(v))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", METRIC_CASES, ids=lambda c: c[0])
def test_metric_known_values(gpu_ctx, oracle_mod, case, dtype):
    name, p1, p2, metric, wrapped, sigma, want = case
    print(f"[case] metric {name} {np.dtype(dtype).name}", flush=True)
    c = np.array([p1, p2], np.float64).T
    # Weights 1 and 2, dc below the distance, rhoc above both densities, dm
    # covering the pair: delta of point 0 is the metric's distance.
    case_ = Case(c, [1.0, 2.0], dc=0.1, rhoc=10.0, dm=50.0, metric=metric, wrapped=wrapped,
                 sigma=None if sigma is None else np.array(sigma), name=name, dtype=dtype)
    gpu = case_.run_gpu(gpu_ctx)
    assert list(gpu["nearest_higher"]) == [1, -1]
    assert gpu["delta"][0] == _expected(want, dtype)
    assert_same(gpu, case_.run_oracle(oracle_mod), name)


FIXTURES = [
    ("data_1024", 2, 1.5, 10, 1.5, 0.9, "data_1024_truth"),
    ("data_2048", 2, 1.5, 10, 1.5, 0.9, "data_2048_truth"),
    ("data_4096", 2, 1.5, 10, 1.5, 0.9, "data_4096_truth"),
    ("data_8192", 2, 1.5, 10, 1.5, 0.9, "data_8192_truth"),
    ("data_16384", 2, 1.5, 10, 1.5, 0.9, "data_16384_truth"),
    ("data_32768", 2, 1.5, 10, 1.5, 0.9, "data_32768_truth"),
    ("data_65536", 2, 1.5, 10, 1.5, 0.9, "data_65536_truth"),
    ("sissa_1000", 2, 20, 10, 20, 0.5, "sissa_1000_truth"),
    ("sissa_4000", 2, 20, 10, 20, 0.45, "sissa_4000_truth"),
    ("toyDetector_1000", 2, 4, 2.5, 4, 0.8, "toy_det_1000_truth"),
    ("blob", 3, 1, 5, 2, 0.8, None),
    # No threshold in the reference for this file (its truth test is disabled): parity only.
    ("toyDetector_5000", 2, 4, 2.5, 4, None, None),
]

_scores: dict = {}


def _silhouette_cached(name, coords, labels):
    key = (name, labels.tobytes())
    if key not in _scores:
        _scores[key] = silhouette(coords, labels)
    return _scores[key]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fx", FIXTURES, ids=[f[0] for f in FIXTURES])
def test_reference_fixture(gpu_ctx, oracle_mod, fx, dtype):
    name, ndim, dc, rhoc, dm, threshold, truth = fx
    print(f"[case] fixture {name} {np.dtype(dtype).name}", flush=True)
    c, w = read_points(f"{name}.csv", dtype=dtype)
    assert c.shape[0] == ndim
    case = Case(c, w, dc, rhoc, dm, name=name, dtype=dtype)
    gpu = case.run_gpu(gpu_ctx)
    assert_same(gpu, case.run_oracle(oracle_mod), name)
    labels = gpu["cluster_index"]
    if threshold is not None:
        assert _silhouette_cached(name, c, labels) >= threshold
    if truth is not None:
        t = pd.read_csv(f"{DATA}/{truth}.csv").cluster_ids.to_numpy()
        assert labels.max() + 1 == t.max() + 1
