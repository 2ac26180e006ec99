"""Served events: each result is the same alone and in the mix (bit-exact labels),
continuous batching folds concurrent arrivals into shared launches, and the
server handles mixed dimensions, tiny events and back-pressure."""
import ctypes as C

import numpy as np
import pytest

from cluestering_amd import _capi
from cluestering_amd.serving import Server, make_pool, run_stream, served_params

pytestmark = pytest.mark.gpu


def _alone(ctx, ev):
    """The same event through the single-event host path."""
    lib = _capi.load()
    out = np.empty(ev.n, np.int32)
    _capi.check(lib.clue_make_clusters_host(ctx.handle, C.byref(ev.params), ev.n,
                                            ev.data.ctypes.data_as(C.c_void_p),
                                            out.ctypes.data_as(C.c_void_p), None, 0, None, None,
                                            None))
    return out


def test_served_equals_alone_in_the_mix(gpu_ctx, oracle_mod):
    # Sizes up to the served stream's largest (262144 points), 2-D and 3-D mixed.
    pool = make_pool(40, seed=11, lo=64, hi=262144)
    assert max(ev.n for ev in pool) > 100000
    with Server(n_lanes=2, max_batch_points=600000) as srv:
        ids = [srv.submit(ev.params, ev.data) for ev in pool]      # all at once: a mix
        results = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["events_done"] == len(pool)
    assert st["batches"] < len(pool)            # events were folded into shared launches
    assert max(t["batch_events"] for _, t in results) > 1
    for ev, (labels, t) in zip(pool, results):
        np.testing.assert_array_equal(labels, _alone(gpu_ctx, ev))
    # A few small ones against the oracle too.
    for ev, (labels, _) in list(zip(pool, results))[:6]:
        if ev.n > 6000:
            continue
        p = ev.params
        kern = {0: "flat", 1: "exp", 2: "gaus"}[p.kernel]
        metric = "periodic_euclidean" if p.metric == _capi.METRIC_PERIODIC_EUCLIDEAN else "euclidean"
        cfg = oracle_mod.Config(ndim=p.ndim, dc=p.dc, rhoc=p.rhoc, dm=p.dm, kernel=kern,
                                kparams=tuple(p.kparam), metric=metric,
                                mparams=tuple(p.mparam)[:p.ndim], wrapped=tuple(p.wrapped)[:p.ndim])
        ref = oracle_mod.run(cfg, ev.data[:-1], ev.data[-1])
        np.testing.assert_array_equal(labels, ref.cluster)


def test_served_stream_and_backpressure(gpu_ctx):
    pool = make_pool(16, seed=12, lo=256, hi=30000)
    # A small staging ring forces submit to wait for earlier events to complete.
    with Server(n_lanes=2, staging_bytes=2 << 20) as srv:
        r = run_stream(srv, pool, rate=20000.0, n_events=200)
        assert srv.stats()["events_done"] == 200
    assert r["events_per_s"] > 0 and r["latency_ms"]["p99"] >= r["latency_ms"]["p50"] > 0


def test_server_rejects_oversized_and_bad_events():
    with Server(n_lanes=1, max_batch_points=1000) as srv:
        p = served_params(2)
        with pytest.raises(_capi.ClueError):
            srv.submit(p, np.zeros((3, 2000), np.float32))
        bad = served_params(2)
        bad.dc = -1.0
        with pytest.raises(ValueError):
            srv.submit(bad, np.zeros((3, 10), np.float32))
        rid = srv.submit(p, np.random.default_rng(0).uniform(0, 5, (3, 500)).astype(np.float32))
        labels, t = srv.wait(rid)
        assert labels.shape == (500,) and t["done_ns"] >= t["launch_ns"] >= t["submit_ns"]


def test_batches_respect_the_hbm_budget():
    """A lane's batch grows only while the device bytes it would hold (workspace,
    staging, columns, labels) stay within its share of the budget: with 48 MiB over
    2 lanes, no batch may exceed 24 MiB at the workspace's smallest per-point size
    (over 100 bytes a point in 2-D), and every event still completes."""
    pool = [ev for ev in make_pool(60, seed=13, lo=2000, hi=60000) if ev.ndim == 2][:20]
    budget = 48 << 20
    with Server(n_lanes=2, hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(ev.params, ev.data) for ev in pool]
        results = [srv.wait(i) for i in ids]
        # An event that cannot fit one lane's share is refused at submission.
        huge = make_pool(1, seed=14, lo=250000, hi=262144)[0]
        with pytest.raises(_capi.ClueError):
            srv.submit(huge.params, huge.data)
    assert all(t["batch_points"] * 100 <= budget // 2 for _, t in results)
    assert max(t["batch_events"] for _, t in results) >= 1
