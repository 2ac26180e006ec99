"""ctypes wrapper of the CLUE oracle (clue_oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module as the checker; the product
package ``cluestering_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

KERNELS = {"flat": 0, "exp": 1, "gaus": 2, "gauss": 2}
METRICS = {
    "euclidean": 0,
    "weighted_euclidean": 1,
    "periodic_euclidean": 2,
    "manhattan": 3,
    "chebyshev": 4,
    "weighted_chebyshev": 5,
    "mahalanobis": 6,
}


class Params(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("kernel", C.c_int32),
        ("metric", C.c_int32),
        ("points_per_tile", C.c_int32),
        ("wrapped", C.c_int32 * 10),
        ("geometry", C.c_int32),
        ("reserved", C.c_int32),
        ("dc", C.c_double),
        ("rhoc", C.c_double),
        ("dm", C.c_double),
        ("seed_dc", C.c_double),
        ("kparam", C.c_double * 2),
        ("mparam", C.c_double * 10),
    ]


class OutF32(C.Structure):
    _fields_ = [
        ("rho", C.POINTER(C.c_float)),
        ("nh", C.POINTER(C.c_int32)),
        ("delta", C.POINTER(C.c_float)),
        ("is_seed", C.POINTER(C.c_int32)),
        ("cluster", C.POINTER(C.c_int32)),
        ("tile_offsets", C.POINTER(C.c_int32)),
        ("tile_index", C.POINTER(C.c_int32)),
        ("n_seeds", C.POINTER(C.c_int64)),
    ]


class OutF64(C.Structure):
    _fields_ = [
        ("rho", C.POINTER(C.c_double)),
        ("nh", C.POINTER(C.c_int32)),
        ("delta", C.POINTER(C.c_double)),
        ("is_seed", C.POINTER(C.c_int32)),
        ("cluster", C.POINTER(C.c_int32)),
        ("tile_offsets", C.POINTER(C.c_int32)),
        ("tile_index", C.POINTER(C.c_int32)),
        ("n_seeds", C.POINTER(C.c_int64)),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("nper", C.c_int32),
        ("ntiles", C.c_int32),
        ("mn", C.c_double * 10),
        ("mx", C.c_double * 10),
        ("side", C.c_double * 10),
    ]


_libs: dict = {}


def build() -> None:
    """Compile the oracle libraries (serial and OpenMP) in place."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    h = C.CDLL(path)
    h.oracle_expf.restype = C.c_float
    h.oracle_expf.argtypes = [C.c_float]
    for fn in (h.oracle_exp_f64, h.oracle_expl_f64):
        fn.restype = C.c_double
        fn.argtypes = [C.c_double]
    _libs[name] = h
    return h


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Config:
    """Clustering parameters in the reference's vocabulary."""

    ndim: int
    dc: float
    rhoc: float
    dm: Optional[float] = None
    seed_dc: Optional[float] = None
    kernel: str = "flat"
    kparams: Sequence[float] = (0.5,)
    metric: str = "euclidean"
    mparams: Sequence[float] = ()
    wrapped: Sequence[int] = ()
    points_per_tile: int = 128
    geometry: str = "event"  # batched: "event" or "global"

    def to_c(self) -> Params:
        p = Params()
        p.ndim = self.ndim
        p.kernel = KERNELS[self.kernel]
        p.metric = METRICS[self.metric]
        p.points_per_tile = self.points_per_tile
        for d, w in enumerate(self.wrapped):
            p.wrapped[d] = int(w)
        p.geometry = 1 if self.geometry == "global" else 0
        p.dc = self.dc
        p.rhoc = self.rhoc
        p.dm = self.dc if self.dm is None else self.dm
        p.seed_dc = self.dc if self.seed_dc is None else self.seed_dc
        for k, v in enumerate(self.kparams):
            p.kparam[k] = v
        for k, v in enumerate(self.mparams):
            p.mparam[k] = v
        return p


@dataclass
class Result:
    rho: np.ndarray
    nh: np.ndarray
    delta: np.ndarray
    is_seed: np.ndarray
    cluster: np.ndarray
    n_seeds: int
    tile_offsets: Optional[np.ndarray] = None
    tile_index: Optional[np.ndarray] = None
    event_of_cluster: Optional[np.ndarray] = field(default=None)


def _soa(coords: np.ndarray, dtype) -> np.ndarray:
    c = np.ascontiguousarray(coords, dtype=dtype)
    if c.ndim == 1:
        c = c.reshape(1, -1)
    return c


def geometry(cfg: Config, coords: np.ndarray) -> Geometry:
    c = _soa(coords, np.float32)
    g = Geometry()
    rc = lib().oracle_geometry_f32(C.byref(cfg.to_c()), C.c_int32(c.shape[1]),
                                   _ptr(c, C.c_float), C.byref(g))
    if rc != 0:
        raise ValueError(f"oracle_geometry_f32 failed ({rc})")
    return g


def run(cfg: Config, coords: np.ndarray, weight: np.ndarray, sigma=None, u=None, tags=None,
        dtype=np.float32, omp: bool = False, tiles: bool = False) -> Result:
    c = _soa(coords, dtype)
    n = c.shape[1]
    w = np.ascontiguousarray(weight, dtype=dtype)
    s = None if sigma is None else _soa(sigma, dtype)
    uu = None if u is None else np.ascontiguousarray(u, dtype=dtype)
    tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
    rho = np.zeros(n, dtype)
    delta = np.zeros(n, dtype)
    nh = np.zeros(n, np.int32)
    seed = np.zeros(n, np.int32)
    cl = np.zeros(n, np.int32)
    ns = np.zeros(1, np.int64)
    to = ti = None
    if tiles:
        g = geometry(cfg, c)
        to = np.zeros(g.ntiles + 1, np.int32)
        ti = np.zeros(n, np.int32)
    h = lib(omp)
    ft = C.c_float if dtype == np.float32 else C.c_double
    out = (OutF32 if dtype == np.float32 else OutF64)()
    out.rho = _ptr(rho, ft)
    out.nh = _ptr(nh, C.c_int32)
    out.delta = _ptr(delta, ft)
    out.is_seed = _ptr(seed, C.c_int32)
    out.cluster = _ptr(cl, C.c_int32)
    out.tile_offsets = _ptr(to, C.c_int32)
    out.tile_index = _ptr(ti, C.c_int32)
    out.n_seeds = _ptr(ns, C.c_int64)
    fn = h.oracle_run_f32 if dtype == np.float32 else h.oracle_run_f64
    rc = fn(C.byref(cfg.to_c()), C.c_int32(n), _ptr(c, ft), _ptr(w, ft), _ptr(s, ft),
            _ptr(uu, ft), _ptr(tg, C.c_uint32), C.byref(out))
    if rc != 0:
        raise ValueError(f"oracle run failed ({rc})")
    return Result(rho, nh, delta, seed, cl, int(ns[0]), to, ti)


def run_batched(cfg: Config, coords: np.ndarray, weight: np.ndarray, sizes: Sequence[int],
                sigma=None, u=None, tags=None, omp: bool = False) -> Result:
    c = _soa(coords, np.float32)
    n = c.shape[1]
    w = np.ascontiguousarray(weight, dtype=np.float32)
    s = None if sigma is None else _soa(sigma, np.float32)
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
    tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
    sz = np.ascontiguousarray(sizes, dtype=np.uint32)
    rho = np.zeros(n, np.float32)
    delta = np.zeros(n, np.float32)
    nh = np.zeros(n, np.int32)
    seed = np.zeros(n, np.int32)
    cl = np.zeros(n, np.int32)
    ns = np.zeros(1, np.int64)
    eoc = np.full(n, -1, np.int32)
    out = OutF32()
    out.rho = _ptr(rho, C.c_float)
    out.nh = _ptr(nh, C.c_int32)
    out.delta = _ptr(delta, C.c_float)
    out.is_seed = _ptr(seed, C.c_int32)
    out.cluster = _ptr(cl, C.c_int32)
    out.tile_offsets = _ptr(None, C.c_int32)
    out.tile_index = _ptr(None, C.c_int32)
    out.n_seeds = _ptr(ns, C.c_int64)
    rc = lib(omp).oracle_run_batched_f32(
        C.byref(cfg.to_c()), C.c_int32(n), _ptr(c, C.c_float), _ptr(w, C.c_float),
        _ptr(s, C.c_float), _ptr(uu, C.c_float), _ptr(tg, C.c_uint32), _ptr(sz, C.c_uint32),
        C.c_int32(len(sz)), C.byref(out), _ptr(eoc, C.c_int32))
    if rc != 0:
        raise ValueError(f"oracle batched run failed ({rc})")
    k = int(ns[0])
    return Result(rho, nh, delta, seed, cl, k, event_of_cluster=eoc[:k].copy())


def brute_density(cfg: Config, coords, weight, sigma=None, omp: bool = True):
    c = _soa(coords, np.float32)
    n = c.shape[1]
    w = np.ascontiguousarray(weight, dtype=np.float32)
    s = None if sigma is None else _soa(sigma, np.float32)
    rho = np.zeros(n, np.float64)
    a = np.zeros(n, np.float64)
    cnt = np.zeros(n, np.int32)
    rc = lib(omp).oracle_brute_density_f32(C.byref(cfg.to_c()), C.c_int32(n), _ptr(c, C.c_float),
                                           _ptr(w, C.c_float), _ptr(s, C.c_float),
                                           _ptr(rho, C.c_double), _ptr(a, C.c_double),
                                           _ptr(cnt, C.c_int32))
    if rc != 0:
        raise ValueError("brute density failed")
    return rho, a, cnt


def brute_nh(cfg: Config, coords, rho, sigma=None, u=None, tags=None, omp: bool = True):
    c = _soa(coords, np.float32)
    n = c.shape[1]
    r = np.ascontiguousarray(rho, dtype=np.float32)
    s = None if sigma is None else _soa(sigma, np.float32)
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
    tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
    nh = np.zeros(n, np.int32)
    delta = np.zeros(n, np.float32)
    rc = lib(omp).oracle_brute_nh_f32(C.byref(cfg.to_c()), C.c_int32(n), _ptr(c, C.c_float),
                                      _ptr(s, C.c_float), _ptr(uu, C.c_float),
                                      _ptr(tg, C.c_uint32), _ptr(r, C.c_float),
                                      _ptr(nh, C.c_int32), _ptr(delta, C.c_float))
    if rc != 0:
        raise ValueError("brute nh failed")
    return nh, delta


def pair_counts(cfg: Config, coords, weight, rho=None, sigma=None, u=None, tags=None,
                omp: bool = True):
    c = _soa(coords, np.float32)
    n = c.shape[1]
    w = np.ascontiguousarray(weight, dtype=np.float32)
    r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float32)
    s = None if sigma is None else _soa(sigma, np.float32)
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
    tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.uint32)
    out = np.zeros(4, np.int64)
    rc = lib(omp).oracle_pair_counts_f32(
        C.byref(cfg.to_c()), C.c_int32(n), _ptr(c, C.c_float), _ptr(w, C.c_float),
        _ptr(s, C.c_float), _ptr(uu, C.c_float), _ptr(tg, C.c_uint32), _ptr(r, C.c_float),
        _ptr(out[0:1], C.c_int64), _ptr(out[1:2], C.c_int64), _ptr(out[2:3], C.c_int64),
        _ptr(out[3:4], C.c_int64))
    if rc != 0:
        raise ValueError("pair counts failed")
    return {"visited_density": int(out[0]), "in_radius": int(out[1]),
            "visited_nh": int(out[2]), "nh_pairs": int(out[3])}


def expf(x: float) -> float:
    return float(lib().oracle_expf(C.c_float(x)))


def expf_array(xs: np.ndarray) -> np.ndarray:
    h = lib()
    return np.array([h.oracle_expf(C.c_float(float(x))) for x in xs], dtype=np.float32)


def exp_f64_array(xs: np.ndarray) -> np.ndarray:
    """The double-precision exp restatement of the oracle (and of the device)."""
    h = lib()
    return np.array([h.oracle_exp_f64(float(x)) for x in xs], dtype=np.float64)


def expl_f64_array(xs: np.ndarray) -> np.ndarray:
    """exp evaluated in long double and rounded to double: the accuracy reference."""
    h = lib()
    return np.array([h.oracle_expl_f64(float(x)) for x in xs], dtype=np.float64)


def num_threads(omp: bool = True) -> int:
    return int(lib(omp).oracle_num_threads())
