"""The Python ``clusterer`` of the reference (CLUEstering/CLUEstering.py:214-827)
on the MI355X backend: input parsing into the (Ndim+1) x N float SoA, kernel and
metric choice, ``run_clue`` / ``fit`` / ``fit_predict`` and the result accessors.
Plotting is not part of this package.

Backends: only "gpu hip" is built here; it is the default.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np

from . import CLUE_Convolutional_Kernels as clue_kernels
from . import CLUE_GPU_HIP as gpu_hip

EuclideanMetric = clue_kernels.EuclideanMetric
ManhattanMetric = clue_kernels.ManhattanMetric
ChebyshevMetric = clue_kernels.ChebyshevMetric
PeriodicEuclideanMetric = clue_kernels.PeriodicEuclideanMetric

backends = ["gpu hip"]

_MAX_DIM = 10
_DIM_ERROR = "Inadequate data. The supported dimensions are between 1 and 10."
# choice -> (constructor, number of parameters, name in messages)
_KERNELS = {
    "flat": (clue_kernels.FlatKernel, 1, "flat"),
    "exp": (clue_kernels.ExponentialKernel, 2, "exponential"),
    "gaus": (clue_kernels.GaussianKernel, 2, "gaussian"),
}
_PLAIN_METRICS = {
    "euclidean": clue_kernels.EuclideanMetric,
    "manhattan": clue_kernels.ManhattanMetric,
    "chebyshev": clue_kernels.ChebyshevMetric,
}
# choice -> (constructor taking the per-dimension list, what the list holds)
_PARAM_METRICS = {
    "weighted_euclidean": (clue_kernels.EuclideanMetric, "weights"),
    "weighted_chebyshev": (clue_kernels.ChebyshevMetric, "weights"),
    "periodic_euclidean": (clue_kernels.PeriodicEuclideanMetric, "periods"),
}


def is_hip_available() -> bool:
    from .device import device_count
    return device_count() > 0


@dataclass()
class ClusteringDataSoA:
    """Input coordinates (with the weights as last row) and the cluster ids."""

    coords: np.ndarray = None
    results: np.ndarray = None
    n_dim: int = None
    n_points: int = None


@dataclass(eq=False)
class cluster_properties:
    """Results of a run (CLUEstering.py:181-211); equality compares n_clusters."""

    n_clusters: int
    cluster_ids: np.ndarray
    cluster_points: np.ndarray
    points_per_cluster: np.ndarray
    output_df: object

    def __eq__(self, other):
        return self.n_clusters == other.n_clusters


class clusterer:
    """CLUE clusterer (CLUEstering.py:214-827).

    :param density_radius: radius of the local density (dc)
    :param min_density: density threshold separating seeds from outliers (rhoc)
    :param outlier_distance: follower search radius (dm), defaults to density_radius
    :param seeding_distance: seed search radius, defaults to density_radius
    """

    def __init__(self, density_radius: float, min_density: float,
                 outlier_distance: Optional[float] = None,
                 seeding_distance: Optional[float] = None):
        self._density_radius = density_radius
        self._min_density = min_density
        self._outlier_distance = density_radius if outlier_distance is None else outlier_distance
        self._seeding_distance = density_radius if seeding_distance is None else seeding_distance
        self.clust_data: Optional[ClusteringDataSoA] = None
        self._kernel = clue_kernels.FlatKernel(0.5)
        self._metric = clue_kernels.EuclideanMetric()
        self.wrapped = None
        self.clust_prop: Optional[cluster_properties] = None
        self._elapsed_time = 0.0

    def set_params(self, density_radius: float, min_density: float,
                   outlier_distance: Optional[float] = None,
                   seeding_distance: Optional[float] = None) -> None:
        self._density_radius = density_radius
        self._min_density = min_density
        self._outlier_distance = density_radius if outlier_distance is None else outlier_distance
        self._seeding_distance = density_radius if seeding_distance is None else seeding_distance

    # -- input (CLUEstering.py:291-404) -------------------------------------------------
    # Every input form ends as one (Ndim+1) x N C-contiguous array: the coordinate
    # rows, then the weight row. Files, dicts and DataFrames are read as float32 by
    # default; row lists and arrays keep the floating type of their first value.
    def _store_soa(self, rows, dtype) -> None:
        n_points = len(rows[-1])
        soa = np.empty((len(rows), n_points), dtype=dtype)
        for k, row in enumerate(rows):
            row = np.asarray(row)
            if row.shape != (n_points,):
                raise ValueError(f"row {k} has {row.size} values where {n_points} are expected")
            soa[k] = row
        self.clust_data = ClusteringDataSoA(soa, np.zeros(n_points, dtype=np.int32),
                                            len(rows) - 1, n_points)

    def _read_array(self, input_data) -> None:
        if not 2 <= len(input_data) <= _MAX_DIM + 1:
            raise ValueError(_DIM_ERROR)
        first = np.asarray(input_data[0]).reshape(-1)[:1]
        dtype = first.dtype if first.size and np.issubdtype(first.dtype, np.floating) \
            else np.dtype(np.float64)
        self._store_soa(list(input_data), dtype)

    def _read_string(self, input_data: str, dtype=np.float32):
        if not input_data.endswith(".csv"):
            raise ValueError("Wrong type of file. The file is not a csv file.")
        import pandas as pd
        return pd.read_csv(input_data, dtype=dtype)

    def _read_dict_df(self, input_data, dtype=np.float32):
        import pandas as pd
        frame = input_data if isinstance(input_data, pd.DataFrame) else pd.DataFrame(input_data)
        return frame.astype(dtype, copy=False)

    def _handle_dataframe(self, df_) -> None:
        n_cols = df_.shape[1]
        if n_cols < 2:
            raise ValueError("Inadequate data. The data must contain at least one coordinate "
                             "and the weight.")
        if n_cols > _MAX_DIM + 1:
            raise ValueError("Inadequate data. The maximum number of dimensions supported is 10.")
        rows = [df_.iloc[:, k].to_numpy() for k in range(n_cols)]
        self._store_soa(rows, rows[0].dtype)

    def read_data(self, input_data, wrapped_coords=None) -> None:
        """Reads a DataFrame, a CSV path, a dict, or a list/array of coordinate rows
        followed by the weight row."""
        import pandas as pd
        if isinstance(input_data, str):
            self._handle_dataframe(self._read_string(input_data))
        elif isinstance(input_data, (dict, pd.DataFrame)):
            self._handle_dataframe(self._read_dict_df(input_data))
        elif isinstance(input_data, (list, np.ndarray)):
            self._read_array(input_data)
        else:
            raise ValueError("unsupported input data type")
        self.wrapped = list(wrapped_coords) if wrapped_coords is not None else \
            [0] * self.clust_data.n_dim

    def set_wrapped(self, wrapped_coords) -> None:
        self.wrapped = list(wrapped_coords)

    # -- kernel and metric (CLUEstering.py:418-496) -------------------------------------
    def choose_kernel(self, choice: str, parameters=None) -> None:
        if choice not in _KERNELS:
            raise ValueError("Invalid kernel. Allowed choices are: flat, exp, gaus.")
        make, n_params, name = _KERNELS[choice]
        if parameters is None or len(parameters) != n_params:
            plural = "parameter" if n_params == 1 else "parameters"
            raise ValueError(f"Wrong number of parameters. The {name} kernel requires "
                             f"{n_params} {plural}.")
        self._kernel = make(*parameters)

    def choose_metric(self, choice: str, parameters=None) -> None:
        if choice in _PLAIN_METRICS:
            self._metric = _PLAIN_METRICS[choice]()
        elif choice in _PARAM_METRICS:
            make, what = _PARAM_METRICS[choice]
            if not parameters:
                raise ValueError(f"{choice} requires a 'parameters' list of per-dimension {what}.")
            self._metric = make(list(parameters))
        else:
            raise ValueError("Invalid metric. Allowed choices are: " +
                             ", ".join(list(_PLAIN_METRICS) + list(_PARAM_METRICS)) + ".")

    # -- accessors -------------------------------------------------------------------------
    @property
    def coords(self) -> np.ndarray:
        return self.clust_data.coords[:-1]

    @property
    def weight(self) -> np.ndarray:
        return self.clust_data.coords[-1]

    @property
    def n_dim(self) -> int:
        return self.clust_data.n_dim

    @property
    def n_points(self) -> int:
        return self.clust_data.n_points

    def list_devices(self, backend: str = "all") -> None:
        if backend in ("all", "gpu hip"):
            gpu_hip.listDevices("gpu hip")
        else:
            raise ValueError(f"Invalid backend. This package provides: {', '.join(backends)}.")

    def _partial_dimension_dataset(self, dimensions) -> np.ndarray:
        # The chosen coordinate rows and the weight row, as a new contiguous SoA.
        return np.ascontiguousarray(self.clust_data.coords[list(dimensions) + [-1]])

    # -- run (CLUEstering.py:609-690) ------------------------------------------------------
    def run_clue(self, backend: str = "gpu hip", batch_sample_sizes=None, device_id: int = 0,
                 verbose: bool = False, dimensions=None) -> None:
        if backend != "gpu hip":
            raise ValueError(f"Invalid backend '{backend}'. This package provides: "
                             f"{', '.join(backends)}.")
        if dimensions is None:
            data = self.clust_data
        else:
            data = ClusteringDataSoA(self._partial_dimension_dataset(dimensions),
                                     np.copy(self.clust_data.results), len(dimensions),
                                     self.clust_data.n_points)
        wrapped = self.wrapped if self.wrapped is not None else [0] * data.n_dim
        if dimensions is not None:
            wrapped = [wrapped[d] for d in dimensions]
        start = time.time_ns()
        gpu_hip.mainRun(self._density_radius, self._min_density, self._outlier_distance,
                        self._seeding_distance, wrapped, data.coords, data.results, self._kernel,
                        data.n_dim, batch_sample_sizes, data.n_points, device_id, self._metric)
        finish = time.time_ns()
        self._set_results(data.results)
        self._elapsed_time = (finish - start) / 1e6
        if verbose:
            print(f"CLUE executed in {self._elapsed_time} ms")
            print(f"Number of clusters found: {self.clust_prop.n_clusters}")

    def _set_results(self, cluster_ids: np.ndarray) -> None:
        """cluster_properties of a label array: n_clusters = max + 1, the member lists
        in ascending point index, and the sizes."""
        cluster_ids = np.asarray(cluster_ids)
        n_clusters = int(np.max(cluster_ids)) + 1 if cluster_ids.size else 0
        valid = np.nonzero(cluster_ids >= 0)[0]
        order = valid[np.argsort(cluster_ids[valid], kind="stable")]
        counts = np.bincount(cluster_ids[valid], minlength=max(n_clusters, 0))
        splits = np.split(order, np.cumsum(counts)[:-1]) if n_clusters > 0 else []
        cluster_points = np.empty(len(splits), dtype=object)
        for k, s in enumerate(splits):
            cluster_points[k] = list(s)
        try:
            import pandas as pd
            output_df = pd.DataFrame({"cluster_ids": cluster_ids})
        except ImportError:  # pragma: no cover
            output_df = None
        self.clust_prop = cluster_properties(n_clusters, cluster_ids, cluster_points,
                                             np.asarray(counts), output_df)

    # -- output (CLUEstering.py:1155-1215) --------------------------------------------------
    def to_csv(self, output_folder: str, file_name: str) -> None:
        """Writes the coordinates, the weight and the cluster id of every point
        (columns x0..x{n-1}, weight, cluster_ids)."""
        import csv
        import os
        if not file_name.endswith(".csv"):
            file_name += ".csv"
        path = os.path.join(output_folder, file_name)
        d = self.clust_data
        cols = [d.coords[k] for k in range(d.n_dim)] + [d.coords[-1], self.clust_prop.cluster_ids]
        with open(path, "w", newline="") as f:
            wr = csv.writer(f)
            wr.writerow([f"x{k}" for k in range(d.n_dim)] + ["weight", "cluster_ids"])
            for i in range(d.n_points):
                wr.writerow([repr(float(c[i])) for c in cols[:-1]] + [int(cols[-1][i])])

    def import_clusterer(self, input_folder: str, file_name: str) -> None:
        """Reads back a file written by to_csv: the points and their cluster ids."""
        import os
        path = os.path.join(input_folder, file_name)
        if not os.path.exists(path):
            raise ValueError(f"The file {path} does not exist.")
        with open(path) as f:
            header = f.readline().strip().split(",")
        if len(header) < 3 or header[-1] != "cluster_ids" or header[-2] != "weight":
            raise ValueError(f"{path} is not a clusterer output (x..., weight, cluster_ids).")
        table = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2, dtype=np.float64)
        ndim = len(header) - 2
        coords = np.ascontiguousarray(table[:, :ndim + 1].T.astype(np.float32))
        ids = table[:, -1].astype(np.int32)
        self.clust_data = ClusteringDataSoA(coords, ids.copy(), ndim, coords.shape[1])
        self._set_results(ids)

    def fit(self, data, backend: str = "gpu hip", batch_sample_sizes=None, device_id: int = 0,
            verbose: bool = False, dimensions=None) -> "clusterer":
        self.read_data(data)
        self.run_clue(backend, batch_sample_sizes, device_id, verbose, dimensions)
        return self

    def fit_predict(self, data, backend: str = "gpu hip", batch_sample_sizes=None,
                    device_id: int = 0, verbose: bool = False, dimensions=None) -> np.ndarray:
        self.read_data(data)
        self.run_clue(backend, batch_sample_sizes, device_id, verbose, dimensions)
        return self.cluster_ids

    @property
    def n_clusters(self) -> int:
        return self.clust_prop.n_clusters

    @property
    def cluster_ids(self) -> np.ndarray:
        return self.clust_prop.cluster_ids

    @property
    def labels(self) -> np.ndarray:
        return self.clust_prop.cluster_ids

    @property
    def cluster_points(self) -> np.ndarray:
        return self.clust_prop.cluster_points

    @property
    def points_per_cluster(self) -> np.ndarray:
        return self.clust_prop.points_per_cluster

    @property
    def output_df(self):
        return self.clust_prop.output_df

    @property
    def elapsed_time(self) -> float:
        return self._elapsed_time
