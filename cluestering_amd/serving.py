"""Continuous-batching CLUE server (clue_server_* in include/clue_amd.h) and a
Poisson-arrival stream driver that measures per-event latency and events/s.

The served stream of the benchmark: arrivals at a fixed rate, ndim 2 or 3 with
equal probability, sizes log-uniform in [256, 262144]; 2-D events as the
batched workload (gaussian kernel), 3-D events as the detector workload
(exponential kernel, periodic phi), each class with its own parameters.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _capi
from .CLUE_Convolutional_Kernels import (ExponentialKernel, GaussianKernel,
                                         PeriodicEuclideanMetric)


ServerConfig = _capi.ServerConfig
Timing = _capi.ServerTiming
Stats = _capi.ServerStats


def now_ns() -> int:
    return int(_capi.load().clue_now_ns())


class Server:
    """One device, ``n_lanes`` concurrent launches, admission under an HBM budget."""

    def __init__(self, device: int = 0, n_lanes: int = 2, hbm_budget_bytes: int = 64 << 30,
                 max_batch_points: int = 0, max_batch_events: int = 0,
                 staging_bytes: int = 1 << 30):
        lib = _capi.load()
        cfg = ServerConfig(device, n_lanes, hbm_budget_bytes, max_batch_points, max_batch_events,
                           staging_bytes)
        h = C.c_void_p()
        _capi.check(lib.clue_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._lib = lib
        self._sizes = {}

    def submit(self, params: _capi.Params, data: np.ndarray) -> int:
        """data: (ndim+1) x n float32 SoA (coordinates, then weights)."""
        d = np.ascontiguousarray(data, dtype=np.float32)
        n = d.shape[1]
        rid = C.c_uint64()
        _capi.check(self._lib.clue_server_submit(self._h, C.byref(params), n,
                                                 d.ctypes.data_as(C.c_void_p), C.byref(rid)))
        self._sizes[rid.value] = n
        return rid.value

    def wait(self, rid: int, timeout_s: Optional[float] = None):
        """Labels (numbered per event from 0) and the timing of one event."""
        n = self._sizes.pop(rid)
        out = np.empty(n, np.int32)
        t = Timing()
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        _capi.check(self._lib.clue_server_wait(self._h, C.c_uint64(rid),
                                               out.ctypes.data_as(C.c_void_p), us, C.byref(t)))
        return out, {"submit_ns": t.submit_ns, "launch_ns": t.launch_ns, "done_ns": t.done_ns,
                     "batch_events": t.batch_events, "batch_points": t.batch_points}

    def stats(self) -> dict:
        s = Stats()
        _capi.check(self._lib.clue_server_get_stats(self._h, C.byref(s)))
        return {f: int(getattr(s, f)) for f, _ in Stats._fields_}

    def close(self) -> None:
        if self._h is not None:
            self._lib.clue_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServedEvent:
    ndim: int
    data: np.ndarray      # (ndim+1) x n float32
    params: _capi.Params

    @property
    def n(self) -> int:
        return self.data.shape[1]


def served_params(ndim: int) -> _capi.Params:
    """Parameters of the two classes of the served stream."""
    if ndim == 2:
        k = GaussianKernel(1.0, 1.0)
        return _capi.make_params(2, 1.5, 5.0, 2.0, None, k.kind, k.params)
    k = ExponentialKernel(0.5, 1.0)
    mid, mp = PeriodicEuclideanMetric([0.0, 0.0, float(np.float32(20 * np.pi))]).resolve(3)
    return _capi.make_params(3, 1.0, 4.0, 1.5, None, k.kind, k.params, mid, mp, (0, 0, 1))


def make_pool(n_events: int, seed: int = 0xC1E0 + 5, lo: int = 256, hi: int = 262144
              ) -> List[ServedEvent]:
    from .synthetic import served_event
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(n_events):
        ndim, c, w = served_event(rng, lo, hi)
        data = np.ascontiguousarray(np.vstack([c, w[None, :]]), dtype=np.float32)
        pool.append(ServedEvent(ndim, data, served_params(ndim)))
    return pool


def run_stream(server: Server, pool: Sequence[ServedEvent], rate: float, n_events: int,
               seed: int = 1, n_clients: int = 4) -> dict:
    """Submits n_events from the pool at Poisson arrival times of the given rate
    (events/s) while a collector thread waits for them in arrival order. Arrival k
    is submitted by client thread k mod n_clients, so one slow submit does not hold
    back the next arrivals (ctypes releases the GIL during the submit's copy).
    Latency = completion time - scheduled arrival (same clock as the server);
    late_submits counts arrivals submitted more than 1 ms after their time."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_events)
    arrivals = np.cumsum(gaps) * 1e9
    order = rng.integers(0, len(pool), size=n_events)
    ids: List[int] = [0] * n_events
    ready = [threading.Event() for _ in range(n_events)]
    lat = np.zeros(n_events)
    done = np.zeros(n_events)
    batch_ev = np.zeros(n_events)
    late = np.zeros(n_clients, np.int64)
    errors: List[BaseException] = []
    t0 = now_ns() + 2_000_000

    def collector():
        for k in range(n_events):
            ready[k].wait()
            if errors:
                return
            _, t = server.wait(ids[k])
            lat[k] = t["done_ns"] - (t0 + arrivals[k])
            done[k] = t["done_ns"]
            batch_ev[k] = t["batch_events"]

    def client(c):
        try:
            for k in range(c, n_events, n_clients):
                target = t0 + arrivals[k]
                while True:
                    dt = target - now_ns()
                    if dt <= 0:
                        break
                    time.sleep(min(dt, 2e6) / 1e9 if dt > 200_000 else 0)
                if now_ns() - target > 1_000_000:
                    late[c] += 1
                ev = pool[order[k]]
                ids[k] = server.submit(ev.params, ev.data)
                ready[k].set()
        except BaseException as e:  # surface in the caller, release the collector
            errors.append(e)
            for r in ready:
                r.set()

    th = threading.Thread(target=collector)
    th.start()
    clients = [threading.Thread(target=client, args=(c,)) for c in range(n_clients)]
    for t in clients:
        t.start()
    for t in clients:
        t.join()
    th.join()
    if errors:
        raise errors[0]
    span = (done.max() - t0) / 1e9
    pts = int(sum(pool[i].n for i in order))
    return {"rate": rate, "events": n_events, "points": pts,
            "events_per_s": n_events / span, "points_per_s": pts / span,
            "latency_ms": {"p50": float(np.percentile(lat, 50)) / 1e6,
                           "p99": float(np.percentile(lat, 99)) / 1e6,
                           "mean": float(lat.mean()) / 1e6},
            "mean_batch_events": float(batch_ev.mean()),
            "late_submits": int(late.sum()), "clients": n_clients}
