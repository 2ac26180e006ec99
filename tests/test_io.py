"""Result files of the Python clusterer (CLUEstering.py:1155-1215): to_csv writes
x0..x{n-1}, weight, cluster_ids; import_clusterer reads them back into the same
points and cluster properties."""
import numpy as np

from cluestering_amd.CLUEstering import clusterer


def test_to_csv_import_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    x, y, w = (rng.uniform(0, 10, 50).astype(np.float32) for _ in range(3))
    ids = rng.integers(-1, 4, 50).astype(np.int32)
    ids[:4] = [0, 1, 2, 3]
    c = clusterer(1.0, 2.0)
    c.read_data([x, y, w])
    c._set_results(ids)
    c.to_csv(str(tmp_path), "result")
    head = (tmp_path / "result.csv").read_text().splitlines()[0]
    assert head == "x0,x1,weight,cluster_ids"
    back = clusterer(1.0, 2.0)
    back.import_clusterer(str(tmp_path), "result.csv")
    assert back.n_clusters == 4
    np.testing.assert_array_equal(back.cluster_ids, ids)
    np.testing.assert_array_equal(back.clust_data.coords, np.vstack([x, y, w]))
    for k in range(4):
        assert list(back.cluster_points[k]) == list(np.nonzero(ids == k)[0])
    np.testing.assert_array_equal(back.points_per_cluster, np.bincount(ids[ids >= 0]))
