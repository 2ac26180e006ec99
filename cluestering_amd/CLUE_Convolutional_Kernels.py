"""Convolutional kernels and distance-metric descriptors.

Mirrors the reference's ``CLUE_Convolutional_Kernels`` module
(CLUEstering/BindingModules/binding_kernels.cpp:10-56) and the kernel classes of
core/ConvolutionalKernel.hpp:19-111, including the constructor checks of
core/detail/ConvolutionalKernel.hpp:14-18, 28-33, 45-50 (ValueError, the Python
face of std::invalid_argument).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

from . import _capi


class FlatKernel:
    """K(i, j) = 1 if i == j else flat (ConvolutionalKernel.hpp:21-25)."""

    kind = _capi.KERNEL_FLAT

    def __init__(self, flat: float):
        if not flat > 0:
            raise ValueError("Flat kernel parameter must be positive.")
        self.flat = float(flat)

    @property
    def params(self):
        return (self.flat, 0.0)

    def __repr__(self):
        return f"FlatKernel({self.flat})"


class ExponentialKernel:
    """K(d) = amplitude * exp(-avg * d), 1 on the diagonal (ConvolutionalKernel.hpp:53-57)."""

    kind = _capi.KERNEL_EXP

    def __init__(self, exp_avg: float, exp_amplitude: float):
        if not (exp_avg > 0 and exp_amplitude > 0):
            raise ValueError("Exponential kernel parameters must be positive.")
        self.exp_avg = float(exp_avg)
        self.exp_amplitude = float(exp_amplitude)

    @property
    def params(self):
        return (self.exp_avg, self.exp_amplitude)

    def __repr__(self):
        return f"ExponentialKernel({self.exp_avg}, {self.exp_amplitude})"


class GaussianKernel:
    """K(d) = amplitude * exp(-d*d / (2*std*std)), 1 on the diagonal (ConvolutionalKernel.hpp:36-41)."""

    kind = _capi.KERNEL_GAUSS

    def __init__(self, gaus_std: float, gaus_amplitude: float):
        if not (gaus_std > 0 and gaus_amplitude > 0):
            raise ValueError("Gaussian kernel parameters must be positive.")
        self.gaus_std = float(gaus_std)
        self.gaus_amplitude = float(gaus_amplitude)

    @property
    def params(self):
        return (self.gaus_std, self.gaus_amplitude)

    def __repr__(self):
        return f"GaussianKernel({self.gaus_std}, {self.gaus_amplitude})"


@dataclass
class MetricDescriptor:
    """Opaque metric handle (BindingModules/MetricDescriptor.hpp:17-27).

    tag names the metric; params are per-dimension weights or periods (empty means
    unit weights for the weighted metrics).
    """

    tag: str
    params: List[float] = field(default_factory=list)

    def resolve(self, ndim: int):
        """(metric id, per-dimension params) for ``ndim`` dimensions, with the
        checks of MetricDescriptor.hpp:30-78."""
        if self.tag == "weighted_euclidean":
            if not self.params:
                return _capi.METRIC_WEIGHTED_EUCLIDEAN, [1.0] * ndim
            if len(self.params) != ndim:
                raise ValueError(f"EuclideanMetric requires exactly {ndim} weight(s), "
                                 f"got {len(self.params)}")
            return _capi.METRIC_WEIGHTED_EUCLIDEAN, list(self.params)
        if self.tag == "manhattan":
            return _capi.METRIC_MANHATTAN, []
        if self.tag == "periodic_euclidean":
            if len(self.params) != ndim:
                raise ValueError(f"PeriodicEuclideanMetric requires exactly {ndim} period(s), "
                                 f"got {len(self.params)}")
            return _capi.METRIC_PERIODIC_EUCLIDEAN, list(self.params)
        if self.tag == "weighted_chebyshev":
            if not self.params:
                return _capi.METRIC_WEIGHTED_CHEBYSHEV, [1.0] * ndim
            if len(self.params) != ndim:
                raise ValueError(f"ChebyshevMetric requires exactly {ndim} weight(s), "
                                 f"got {len(self.params)}")
            return _capi.METRIC_WEIGHTED_CHEBYSHEV, list(self.params)
        if self.tag == "mahalanobis":
            return _capi.METRIC_MAHALANOBIS, []
        raise ValueError(f"unknown metric {self.tag!r}")


def EuclideanMetric(weights: Sequence[float] = ()) -> MetricDescriptor:
    """Euclidean (L2) metric; optional per-dimension weights (binding_kernels.cpp:27-34)."""
    return MetricDescriptor("weighted_euclidean", [float(w) for w in weights])


def ManhattanMetric() -> MetricDescriptor:
    """Manhattan (L1) metric (binding_kernels.cpp:36-39)."""
    return MetricDescriptor("manhattan")


def ChebyshevMetric(weights: Sequence[float] = ()) -> MetricDescriptor:
    """Chebyshev (L-infinity) metric; optional weights (binding_kernels.cpp:41-48)."""
    return MetricDescriptor("weighted_chebyshev", [float(w) for w in weights])


def PeriodicEuclideanMetric(periods: Sequence[float]) -> MetricDescriptor:
    """Periodic Euclidean metric; a period of 0 marks a non-periodic dimension
    (binding_kernels.cpp:50-55)."""
    return MetricDescriptor("periodic_euclidean", [float(p) for p in periods])


def MahalanobisMetric() -> MetricDescriptor:
    """Mahalanobis metric with per-point sigma (core/DistanceMetrics.hpp:64-87); the C++
    reference exposes it as clue::metrics::Mahalanobis."""
    return MetricDescriptor("mahalanobis")
