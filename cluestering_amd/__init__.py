"""MI355X-native CLUE (gfx950): hand-written HIP stages behind a C ABI
(include/clue_amd.h, libclue_amd.so), with the reference's Python surface on top.
"""
from . import _capi
from .CLUE_Convolutional_Kernels import (ChebyshevMetric, EuclideanMetric, ExponentialKernel,
                                         FlatKernel, GaussianKernel, MahalanobisMetric,
                                         ManhattanMetric, PeriodicEuclideanMetric)

__all__ = [
    "ChebyshevMetric", "EuclideanMetric", "ExponentialKernel", "FlatKernel", "GaussianKernel",
    "MahalanobisMetric", "ManhattanMetric", "PeriodicEuclideanMetric",
]
__version__ = "0.1.0"
