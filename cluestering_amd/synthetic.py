"""Synthetic CLUE workloads with fixed seeds (numpy PCG64).

Each generator returns float32 coordinates shaped (ndim, n), float32 weights and
the clustering parameters of the workload. The same arrays feed the GPU path and
the CPU checks.

Workloads:
  blobs2d_16k      2-D, N=16384: 20 Gaussian blobs (sigma 1) with centres uniform
                   in [0,100]^2, 80% of the points in blobs, 20% uniform noise,
                   unit weights; dc=1.5, rhoc=5, dm=2, flat kernel 0.5.
  blobs2d_1m       2-D, N=2^20: 200 blobs, centres in [0,1000]^2, sigma uniform in
                   [0.5,2], 85% blob points, weights uniform in [0.5,1.5];
                   dc=1.5, rhoc=8, dm=2.5, flat kernel 0.5.
  detector3d_1m    3-D, N=2^20: x, y uniform in [-150,150], periodic phi in
                   [-pi,pi); 2000 showers with spread (1.5, 1.5, 0.05), 20% noise,
                   lognormal(0,1) weights; exponential kernel (decay 0.5, amplitude
                   1). The per-dimension radii (2, 2, 0.1) are expressed with the
                   scalar radius of the algorithm by rescaling the coordinates to
                   (x/2, y/2, phi/0.1): dc'=1, dm'=1.5, phi period 20*pi, rhoc=4.
  batch2d_4096     4096 events, 2-D, sizes log-uniform in [256, 65536], each with
                   5-50 blobs (sigma 1) in [0,100]^2 and 20% noise, weights uniform
                   in [0.5,1.5]; gaussian kernel (sigma 1, amplitude 1);
                   dc=1.5, rhoc=5, dm=2.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .CLUE_Convolutional_Kernels import (ExponentialKernel, FlatKernel, GaussianKernel,
                                         MetricDescriptor, PeriodicEuclideanMetric)


@dataclass
class Workload:
    name: str
    coords: np.ndarray          # (ndim, n) float32
    weight: np.ndarray          # (n,) float32
    dc: float
    rhoc: float
    dm: float
    kernel: object
    metric: Optional[MetricDescriptor] = None
    wrapped: Sequence[int] = ()
    sizes: Optional[np.ndarray] = None  # batched workloads: points per event
    ndims: Optional[List[int]] = None
    meta: dict = field(default_factory=dict)

    @property
    def ndim(self) -> int:
        return self.coords.shape[0]

    @property
    def n(self) -> int:
        return self.coords.shape[1]


def _blobs2d(rng, n, n_blobs, box, sigma, blob_frac, weights):
    per = int(np.floor(blob_frac * n / n_blobs))
    centres = rng.uniform(box[0], box[1], size=(n_blobs, 2))
    sig = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n_blobs,))
    parts = [rng.normal(centres[b], sig[b], size=(per, 2)) for b in range(n_blobs)]
    noise = rng.uniform(box[0], box[1], size=(n - per * n_blobs, 2))
    xy = np.concatenate(parts + [noise]).T.astype(np.float32)
    if weights == "unit":
        w = np.ones(n, np.float32)
    else:
        w = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    return np.ascontiguousarray(xy), w


def blobs2d_16k(seed: int = 0xC1E0 + 1, n: int = 16384) -> Workload:
    rng = np.random.default_rng(seed)
    xy, w = _blobs2d(rng, n, 20, (0.0, 100.0), 1.0, 0.8, "unit")
    return Workload("blobs2d_16k", xy, w, 1.5, 5.0, 2.0, FlatKernel(0.5))


def blobs2d_1m(seed: int = 0xC1E0 + 2, n: int = 1 << 20) -> Workload:
    """At n below 2^20 the blob count and the box shrink with n (same density)."""
    rng = np.random.default_rng(seed)
    frac = n / float(1 << 20)
    n_blobs = max(1, int(round(200 * frac)))
    box = 1000.0 * np.sqrt(n_blobs / 200.0)
    sig = rng.uniform(0.5, 2.0, size=n_blobs)
    xy, w = _blobs2d(rng, n, n_blobs, (0.0, box), sig, 0.85, "uniform")
    return Workload("blobs2d_1m", xy, w, 1.5, 8.0, 2.5, FlatKernel(0.5))


def _detector3d(rng, n, n_showers=2000, noise_frac=0.2, half=150.0):
    per = int(np.floor((1.0 - noise_frac) * n / n_showers))
    cx = rng.uniform(-half, half, size=n_showers)
    cy = rng.uniform(-half, half, size=n_showers)
    cp = rng.uniform(-np.pi, np.pi, size=n_showers)
    x = rng.normal(np.repeat(cx, per), 1.5)
    y = rng.normal(np.repeat(cy, per), 1.5)
    p = rng.normal(np.repeat(cp, per), 0.05)
    nn = n - per * n_showers
    x = np.concatenate([x, rng.uniform(-half, half, size=nn)])
    y = np.concatenate([y, rng.uniform(-half, half, size=nn)])
    p = np.concatenate([p, rng.uniform(-np.pi, np.pi, size=nn)])
    p = (p + np.pi) % (2 * np.pi) - np.pi  # wrap into [-pi, pi)
    w = rng.lognormal(0.0, 1.0, size=n).astype(np.float32)
    # Rescale to one scalar radius: (x/2, y/2, phi/0.1).
    c = np.stack([x / 2.0, y / 2.0, p / 0.1]).astype(np.float32)
    return np.ascontiguousarray(c), w


def detector3d_1m(seed: int = 0xC1E0 + 3, n: int = 1 << 20) -> Workload:
    """At n below 2^20 the shower count and the x-y box shrink with n (same density)."""
    rng = np.random.default_rng(seed)
    frac = n / float(1 << 20)
    n_showers = max(1, int(round(2000 * frac)))
    c, w = _detector3d(rng, n, n_showers, half=150.0 * np.sqrt(n_showers / 2000.0))
    period = np.float32(20.0 * np.pi)
    return Workload("detector3d_1m", c, w, 1.0, 4.0, 1.5, ExponentialKernel(0.5, 1.0),
                    metric=PeriodicEuclideanMetric([0.0, 0.0, float(period)]),
                    wrapped=(0, 0, 1))


def event_sizes(rng, n_events, lo=256, hi=65536) -> np.ndarray:
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_events))).astype(np.uint32)


def _event2d(rng, n):
    n_blobs = int(rng.integers(5, 51))
    xy, w = _blobs2d(rng, n, n_blobs, (0.0, 100.0), 1.0, 0.8, "uniform")
    return xy, w


def batch2d_4096(seed: int = 0xC1E0 + 4, n_events: int = 4096, lo: int = 256,
                 hi: int = 65536) -> Workload:
    rng = np.random.default_rng(seed)
    sizes = event_sizes(rng, n_events, lo, hi)
    cs, ws = [], []
    for n in sizes:
        xy, w = _event2d(rng, int(n))
        cs.append(xy)
        ws.append(w)
    c = np.ascontiguousarray(np.concatenate(cs, axis=1))
    w = np.ascontiguousarray(np.concatenate(ws))
    return Workload("batch2d_4096", c, w, 1.5, 5.0, 2.0, GaussianKernel(1.0, 1.0), sizes=sizes)


def blobs_nd(rng, n, ndim, n_blobs=None, box=100.0, sigma=1.0, noise=0.2):
    """Gaussian blobs in ndim dimensions (centres uniform in [0, box]^ndim, 1-sigma
    spread) plus a uniform noise share, weights uniform in [0.5, 1.5]."""
    if n_blobs is None:
        n_blobs = int(rng.integers(5, 51))
    per = int(np.floor((1.0 - noise) * n / n_blobs))
    centres = rng.uniform(0.0, box, size=(n_blobs, ndim))
    parts = [rng.normal(centres[b], sigma, size=(per, ndim)) for b in range(n_blobs)]
    rest = rng.uniform(0.0, box, size=(n - per * n_blobs, ndim))
    c = np.concatenate(parts + [rest]).T.astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    return np.ascontiguousarray(c), w


def served_event(rng, lo=256, hi=262144):
    """One event of the served stream: ndim 2 or 3 with equal probability, size
    log-uniform in [lo, hi]; 2-D events as in batch2d_4096, 3-D as detector3d_1m."""
    n = int(np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    if rng.random() < 0.5:
        xy, w = _event2d(rng, n)
        return 2, xy, w
    showers = max(1, int(round(2000 * n / (1 << 20))))
    c, w = _detector3d(rng, n, showers, half=150.0 * np.sqrt(showers / 2000.0))
    return 3, c, w
