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


def silhouette(coords, labels):
    """Mean silhouette over clustered points in clusters of >= 2 points, as
    utils/detail/scores.hpp:25-99 defines it (Euclidean, outliers excluded)."""
    x = np.asarray(coords, np.float64).T
    lab = np.asarray(labels)
    ids = [c for c in np.unique(lab) if c >= 0]
    members = {c: np.nonzero(lab == c)[0] for c in ids}
    scores = []
    for i in range(len(lab)):
        c = lab[i]
        if c < 0 or len(members[c]) < 2:
            continue
        d_own = np.sqrt(((x[members[c]] - x[i]) ** 2).sum(1))
        a = d_own.sum() / (len(members[c]) - 1)
        b = np.inf
        for o in ids:
            if o == c:
                continue
            b = min(b, np.sqrt(((x[members[o]] - x[i]) ** 2).sum(1)).mean())
        scores.append((b - a) / max(a, b))
    return float(np.mean(scores)) if scores else 0.0


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
