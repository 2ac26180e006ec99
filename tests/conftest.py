"""Shared test setup: the `gpu` marker, repo paths and CPU-side helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DATA = os.path.join(ROOT, "tests", "golden", "reference_data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libclue_amd.so)")


def read_points(name, dtype=np.float32):
    """Reads one of the reference's CSV fixtures: coordinate columns then weight."""
    import pandas as pd
    df = pd.read_csv(os.path.join(DATA, name), dtype=dtype)
    a = df.to_numpy().T
    return np.ascontiguousarray(a[:-1]), np.ascontiguousarray(a[-1])


def silhouette(coords, labels, rows=512):
    """Mean silhouette over clustered points in clusters of >= 2 points, as
    utils/detail/scores.hpp:25-99 defines it (Euclidean, outliers excluded; b is
    the smallest mean distance to any other cluster, single-point ones included).

    Distances come in blocks of `rows` query points against all clustered points
    (float64, |q|^2 + |x|^2 - 2 q.x); the per-cluster sums are one matrix product
    with the membership matrix, so the reference's largest fixtures (65536
    points) take seconds."""
    x = np.asarray(coords, np.float64).T
    lab = np.asarray(labels)
    keep = np.nonzero(lab >= 0)[0]
    if keep.size == 0:
        return 0.0
    ids, inv = np.unique(lab[keep], return_inverse=True)
    k = ids.size
    counts = np.bincount(inv, minlength=k).astype(np.float64)
    xs = np.ascontiguousarray(x[keep])
    norm = (xs * xs).sum(1)
    member = np.zeros((keep.size, k))
    member[np.arange(keep.size), inv] = 1.0
    scores = []
    for b0 in range(0, keep.size, rows):
        q = xs[b0:b0 + rows]
        d2 = q @ xs.T
        d2 *= -2.0
        d2 += norm[b0:b0 + rows, None]
        d2 += norm[None, :]
        np.maximum(d2, 0.0, out=d2)
        np.sqrt(d2, out=d2)
        sums = d2 @ member  # (rows, clusters): sum of distances per cluster
        own = inv[b0:b0 + rows]
        valid = counts[own] >= 2
        r = np.arange(q.shape[0])
        a = sums[r, own] / np.maximum(counts[own] - 1, 1)
        mean = sums / counts[None, :]
        mean[r, own] = np.inf
        b = mean.min(1) if k > 1 else np.full(q.shape[0], np.inf)
        s = (b - a) / np.maximum(a, b)
        scores.append(s[valid])
    s = np.concatenate(scores)
    return float(s.mean()) if s.size else 0.0


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from cluestering_amd import device
    if device.device_count() < 1:
        pytest.fail("no HIP device visible to libclue_amd")
    ctx = device.Context(0)
    yield ctx
    ctx.close()
