"""CPU checks of the C ABI boundary: the library loads, exports every symbol the
header declares, validates parameters like the reference, and the Python surface
raises the reference's exceptions. No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cluestering_amd import CLUE_Convolutional_Kernels as K
from cluestering_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "clue_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clue_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # The ctypes table covers exactly the declared functions.
    assert sorted(_capi.EXPORTED_SYMBOLS) == declared


def test_version_and_no_device_errors():
    lib = _capi.load()
    assert b"gfx950" in lib.clue_version()
    n = C.c_int32(-1)
    lib.clue_device_count(C.byref(n))
    assert n.value >= 0


def _p(**kw):
    base = dict(ndim=2, dc=1.5, rhoc=10.0)
    base.update(kw)
    return _capi.make_params(**base)


@pytest.mark.parametrize("kw,status", [
    (dict(), _capi.OK),
    (dict(dc=-1.0), _capi.ERR_INVALID_ARGUMENT),   # tests/test_clusterer.cpp:46-55
    (dict(rhoc=-10.0), _capi.ERR_INVALID_ARGUMENT),
    (dict(rhoc=0.0), _capi.OK),
    (dict(dm=0.0), _capi.ERR_INVALID_ARGUMENT),
    (dict(seed_dc=-2.0), _capi.ERR_INVALID_ARGUMENT),
    (dict(dc=float("nan")), _capi.ERR_INVALID_ARGUMENT),
    (dict(ndim=0), _capi.ERR_OUT_OF_RANGE),
    (dict(ndim=11), _capi.ERR_OUT_OF_RANGE),
    (dict(kparams=(0.0,)), _capi.ERR_INVALID_ARGUMENT),
    (dict(kernel=_capi.KERNEL_EXP, kparams=(1.0, 0.0)), _capi.ERR_INVALID_ARGUMENT),
    (dict(kernel=_capi.KERNEL_GAUSS, kparams=(-1.0, 1.0)), _capi.ERR_INVALID_ARGUMENT),
    (dict(kernel=7), _capi.ERR_INVALID_ARGUMENT),
    (dict(metric=9), _capi.ERR_INVALID_ARGUMENT),
])
def test_validate_params(kw, status):
    lib = _capi.load()
    p = _p(**kw)
    assert lib.clue_validate_params(C.byref(p)) == status
    if status:
        assert _capi.last_error()


def test_calls_without_context_fail_cleanly():
    lib = _capi.load()
    p = _p()
    assert lib.clue_make_clusters(None, C.byref(p), None, None) == _capi.ERR_INVALID_ARGUMENT
    assert lib.clue_get_n_clusters(None, None) == _capi.ERR_INVALID_ARGUMENT
    h = C.c_void_p()
    st = lib.clue_context_create(0, None, C.byref(h))
    if st == _capi.OK:  # a GPU is visible: release it
        lib.clue_context_destroy(h)
    else:
        assert st == _capi.ERR_NO_DEVICE and not h.value


def test_kernel_constructors_throw():
    # tests/test_convolutional_kernels.cpp:7-24
    with pytest.raises(ValueError):
        K.FlatKernel(0.0)
    with pytest.raises(ValueError):
        K.FlatKernel(-1.0)
    with pytest.raises(ValueError):
        K.ExponentialKernel(-1.0, 1.0)
    with pytest.raises(ValueError):
        K.ExponentialKernel(1.0, 0.0)
    with pytest.raises(ValueError):
        K.GaussianKernel(0.0, 1.0)
    with pytest.raises(ValueError):
        K.GaussianKernel(1.0, -1.0)
    K.FlatKernel(0.5), K.ExponentialKernel(1.0, 1.0), K.GaussianKernel(1.0, 1.0)


def test_metric_descriptors_resolve():
    # BindingModules/MetricDescriptor.hpp:30-78
    assert K.EuclideanMetric().resolve(3) == (_capi.METRIC_WEIGHTED_EUCLIDEAN, [1.0] * 3)
    assert K.EuclideanMetric([1, 2]).resolve(2) == (_capi.METRIC_WEIGHTED_EUCLIDEAN, [1.0, 2.0])
    with pytest.raises(ValueError):
        K.EuclideanMetric([1, 2]).resolve(3)
    with pytest.raises(ValueError):
        K.PeriodicEuclideanMetric([0.0]).resolve(2)
    with pytest.raises(ValueError):
        K.ChebyshevMetric([1.0]).resolve(2)
    assert K.ManhattanMetric().resolve(4)[0] == _capi.METRIC_MANHATTAN
    assert K.ChebyshevMetric().resolve(2) == (_capi.METRIC_WEIGHTED_CHEBYSHEV, [1.0, 1.0])


def test_square_root_threshold_is_exact():
    """The stages compare sum-of-squares against the largest float whose
    correctly rounded square root is within the radius; check that equivalence
    exhaustively around the radius for several radii."""
    for r in (1.5, 2.0, 2.5, 0.2, 1.0, 3.0, 0.1):
        r32 = np.float32(r)
        t = np.float32(r32 * r32)
        while np.sqrt(t) > r32:
            t = np.nextafter(t, np.float32(0))
        while np.sqrt(np.nextafter(t, np.float32(np.inf))) <= r32:
            t = np.nextafter(t, np.float32(np.inf))
        lo = t.view(np.int32) - 4000
        xs = np.arange(lo, lo + 8000, dtype=np.int32).view(np.float32)
        np.testing.assert_array_equal(np.sqrt(xs) <= r32, xs <= t)
