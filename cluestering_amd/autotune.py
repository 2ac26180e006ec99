"""Launch-tuning sweep behind the built-in table (csrc/clue_tuning_table.h).

For each ndim (2, 3, 4 and 6 by default) and mean-event-size bucket it builds a
synthetic batch, replays it as a captured graph under every candidate tuning
(query grouping on/off x XCD chunk sizes x waves per workgroup x persistent
waves pulling from the device work-item queues) and keeps the fastest. Tuning never changes results, and the sweep
checks that too: every candidate's labels must equal the first candidate's.

    python -m cluestering_amd.autotune --out autotune.json [--write-table]

Needs a GPU; --write-table regenerates clue_tuning_table.h from the sweep (rebuild
the library afterwards).
"""
from __future__ import annotations

import argparse
import json
import os
import time
from typing import Dict, List, Tuple

import numpy as np

from . import CLUE_Convolutional_Kernels as K
from . import synthetic
from .device import Clusterer, Context, Plan, PointsDevice

# Candidate tunings: (group_queries, xcd_chunk_log2, block_waves, persistent waves per CU).
# Persistent waves (device work-item queues) exist for 1-3 dimensions only.
def candidates(ndim: int) -> List[Tuple[int, int, int, int]]:
    pers = (0, 16, 32) if ndim <= 3 else (0,)
    return [(g, x, b, p) for g in (0, 1) for x in (-1, 2, 4, 6, 8) for b in (1, 2) for p in pers]

# Buckets by mean event size: (upper bound of the bucket, event size range of the
# synthetic batch). The last bucket is one large event.
BUCKETS = [(2048, (256, 2048)), (16384, (2048, 16384)), (131072, (16384, 131072)),
           (2 ** 62, (1 << 20, 1 << 20))]

BUCKETS_ND = [(2048, (256, 2048)), (16384, (2048, 16384)), (2 ** 62, (65536, 262144))]

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "clue_tuning_table.h")


def _batch(ndim: int, lo: int, hi: int, target_points: int, rng):
    """Events of the bucket (2-D: gaussian blobs as the batched workload; 3-D:
    detector-like showers; more dimensions: gaussian blobs in that many), about
    target_points in total."""
    if lo >= 1 << 20 and ndim <= 3:  # one large event: the 2^20-point single-event workloads
        w = (synthetic.blobs2d_1m(seed=int(rng.integers(1 << 30))) if ndim == 2
             else synthetic.detector3d_1m(seed=int(rng.integers(1 << 30))))
        return w.coords, w.weight, np.array([w.coords.shape[1]], np.uint32)
    coords, weights, sizes = [], [], []
    total = 0
    while total < target_points:
        n = int(np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi))))) if hi > lo else lo
        if ndim == 2:
            c, w = synthetic._event2d(rng, n)
        elif ndim > 3:
            c, w = synthetic.blobs_nd(rng, n, ndim)
        else:
            showers = max(1, int(round(2000 * n / (1 << 20))))
            c, w = synthetic._detector3d(rng, n, showers, half=150.0 * np.sqrt(showers / 2000.0))
        coords.append(np.asarray(c, np.float32))
        weights.append(np.asarray(w, np.float32))
        sizes.append(n)
        total += n
    return np.concatenate(coords, axis=1), np.concatenate(weights), np.array(sizes, np.uint32)


def _setup(ctx: Context, ndim: int, coords, weight):
    pts = PointsDevice.from_host(ctx, coords, weight)
    if ndim == 2:
        algo = Clusterer(ctx, 1.5, 5.0, 2.0)
        return pts, algo, None, K.GaussianKernel(1.0, 1.0)
    if ndim > 3:
        algo = Clusterer(ctx, 1.5, 3.0, 2.0)
        return pts, algo, None, K.GaussianKernel(1.0, 1.0)
    algo = Clusterer(ctx, 1.0, 4.0, 1.5)
    algo.set_wrapped_coordinates((0, 0, 1))
    metric = K.PeriodicEuclideanMetric([0.0, 0.0, float(np.float32(20 * np.pi))])
    return pts, algo, metric, K.ExponentialKernel(0.5, 1.0)


def sweep_bucket(ctx: Context, ndim: int, bound: int, lo: int, hi: int, target_points: int,
                 reps: int, seed: int) -> Dict:
    rng = np.random.default_rng(seed)
    coords, weight, sizes = _batch(ndim, lo, hi, target_points, rng)
    pts, algo, metric, kernel = _setup(ctx, ndim, coords, weight)
    results, ref_labels = [], None
    for group, xcd, waves, pers in candidates(ndim):
        ctx.set_tuning(group_queries=bool(group), xcd_chunk_log2=xcd, block_waves=waves,
                       persistent=pers)
        plan = Plan(algo, pts, sizes, metric, kernel)
        plan.launch()
        ctx.synchronize()
        labels = pts.cluster_index.to_host()
        if ref_labels is None:
            ref_labels = labels
        elif not np.array_equal(labels, ref_labels):
            raise AssertionError(f"tuning {(group, xcd, waves, pers)} changed the labels")
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            plan.launch()
            ctx.synchronize()
            times.append(time.perf_counter() - t0)
        plan.close()
        results.append({"group_queries": group, "xcd_chunk_log2": xcd, "block_waves": waves,
                        "persistent": pers, "ms": float(np.median(times) * 1e3)})
    ctx.set_tuning()
    best = min(results, key=lambda r: r["ms"])
    return {"ndim": ndim, "max_mean_event": bound, "events": int(sizes.size),
            "points": int(sizes.sum()), "candidates": results, "best": best}


def write_table(entries: List[Dict], path: str = TABLE) -> None:
    lines = []
    for e in sorted(entries, key=lambda e: (e["ndim"], e["max_mean_event"])):
        b = e["best"]
        bound = "INT64_MAX" if e["max_mean_event"] >= 2 ** 62 else str(e["max_mean_event"])
        lines.append(f"    {{{e['ndim']}, {bound}, {b['group_queries']}, {b['xcd_chunk_log2']}, "
                     f"{b['block_waves']}, {b.get('persistent', 0)}}},")
    # Dimensions without entries of their own take the highest swept dimension's
    # last entry.
    top = max(entries, key=lambda e: (e["ndim"], e["max_mean_event"]))["best"]
    lines.append(f"    {{0, INT64_MAX, {top['group_queries']}, {top['xcd_chunk_log2']}, "
                 f"{top['block_waves']}, 0}},  // other ndim")
    text = (
        "// Built-in launch tuning, by ndim (0 = any) and mean event size (first match).\n"
        "// Regenerated from MI355X measurements by cluestering_amd/autotune.py.\n"
        "#pragma once\n\n#include <cstdint>\n\nnamespace clue_rt {\n\n"
        "struct TuningEntry {\n  int32_t ndim;\n  int64_t max_mean_event;\n"
        "  int32_t group_queries;\n  int32_t xcd_chunk_log2;\n  int32_t block_waves;\n"
        "  int32_t persistent;  // persistent waves per CU (0: hardware dispatch)\n};\n\n"
        "inline constexpr TuningEntry kTuningTable[] = {\n" + "\n".join(lines) + "\n};\n\n"
        "}  // namespace clue_rt\n")
    with open(path, "w") as f:
        f.write(text)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default="autotune.json")
    ap.add_argument("--ndims", default="2,3,4,6")
    ap.add_argument("--points", type=int, default=8 << 20, help="points per bucket batch")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--write-table", action="store_true")
    ap.add_argument("--table", default=TABLE, help="path of the generated table header")
    a = ap.parse_args(argv)
    ctx = Context(0)
    entries = []
    for ndim in (int(x) for x in a.ndims.split(",")):
        # Four and more dimensions: the two smaller buckets and one of large events.
        buckets = BUCKETS if ndim <= 3 else BUCKETS_ND
        for k, (bound, (lo, hi)) in enumerate(buckets):
            e = sweep_bucket(ctx, ndim, bound, lo, hi, a.points, a.reps, seed=1000 * ndim + k)
            print(f"ndim {ndim} mean <= {bound}: best {e['best']} "
                  f"({e['events']} events, {e['points']} points)", flush=True)
            entries.append(e)
    with open(a.out, "w") as f:
        json.dump(entries, f, indent=1)
    if a.write_table:
        write_table(entries, a.table)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
