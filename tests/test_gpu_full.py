"""The benchmark configurations at their full sizes.

- 2-D, N = 2^20, 200 blobs, flat kernel; and 3-D detector-like, N = 2^20, periodic
  phi, exponential kernel: every output bit-exact against the oracle.
- The 4096-event batch (48.8M points): size-independent properties over every
  point (the seed rule, labels constant along nearest-higher chains, seeds
  numbered in ascending index, event-major ids, every point a member of its own
  event), and a sample of events bit-exact against the oracle.
"""
import numpy as np
import pytest

from cluestering_amd import synthetic

from _support import Case, assert_same

pytestmark = pytest.mark.gpu


def _case(w, name):
    return Case(w.coords, w.weight, w.dc, w.rhoc, w.dm, kernel=w.kernel, metric=w.metric,
                wrapped=w.wrapped, sizes=w.sizes, name=name)


@pytest.mark.parametrize("make", [synthetic.blobs2d_1m, synthetic.detector3d_1m])
def test_single_event_full_size(gpu_ctx, oracle_mod, make):
    w = make()
    case = _case(w, w.name)
    print(f"[case] {w.name} n={case.n}", flush=True)
    gpu = case.run_gpu(gpu_ctx)
    ref = case.run_oracle(oracle_mod)
    assert_same(gpu, ref, w.name)
    assert gpu["seeds"].size == ref.n_seeds


def test_batch_4096_properties_and_sample(gpu_ctx, oracle_mod):
    w = synthetic.batch2d_4096()
    case = _case(w, "batch2d_4096")
    print(f"[case] batch2d_4096 n={case.n}", flush=True)
    gpu = case.run_gpu(gpu_ctx)
    rho, nh, seed, cl = gpu["rho"], gpu["nearest_higher"], gpu["is_seed"], gpu["cluster_index"]
    sizes = np.asarray(w.sizes, np.int64)
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    event = np.repeat(np.arange(sizes.size), sizes)
    # Seed rule (ClusteringKernels.hpp:300-316) with u = 1.
    np.testing.assert_array_equal(seed.astype(bool), (nh == -1) & (rho >= np.float32(w.rhoc)))
    # Nearest-higher stays within the event and is higher in (rho, index) order.
    has = nh >= 0
    j = nh[has]
    assert (event[j] == event[has]).all()
    i = np.nonzero(has)[0]
    assert ((rho[j] > rho[i]) | ((rho[j] == rho[i]) & (rho[j] > 0) & (j > i))).all()
    # Labels: seeds numbered 0..S-1 in ascending index; followers take their
    # nearest-higher's label; points without one that are not seeds are -1.
    seeds = np.nonzero(seed)[0]
    np.testing.assert_array_equal(cl[seeds], np.arange(seeds.size))
    np.testing.assert_array_equal(gpu["seeds"], seeds)
    fol = has & ~seed.astype(bool)
    np.testing.assert_array_equal(cl[fol], cl[nh[fol]])
    assert (cl[(nh == -1) & ~seed.astype(bool)] == -1).all()
    assert gpu["n_clusters"] == seeds.size
    # Every cluster lives in one event, and clusters are numbered event by event.
    assert (np.diff(event[seeds]) >= 0).all()
    lab = cl >= 0
    assert (event[seeds[cl[lab]]] == event[lab]).all()
    # A sample of events (small, mid-size and the largest) against the oracle.
    order = np.argsort(sizes)
    picks = sorted(set(order[:3].tolist() + order[sizes.size // 2:sizes.size // 2 + 3].tolist()
                       + order[-2:].tolist()))
    cfg = case.oracle_config(oracle_mod)
    for e in picks:
        sl = slice(int(bounds[e]), int(bounds[e + 1]))
        ref = oracle_mod.run(cfg, w.coords[:, sl], w.weight[sl], omp=True)
        base = int(cl[sl][cl[sl] >= 0].min()) if (cl[sl] >= 0).any() else 0
        np.testing.assert_array_equal(rho[sl].view(np.uint32), ref.rho.view(np.uint32))
        np.testing.assert_array_equal(nh[sl], np.where(ref.nh >= 0, ref.nh + bounds[e], -1))
        np.testing.assert_array_equal(cl[sl], np.where(ref.cluster >= 0, ref.cluster + base, -1))
        np.testing.assert_array_equal(gpu["delta"][sl].view(np.uint32),
                                      ref.delta.view(np.uint32))
