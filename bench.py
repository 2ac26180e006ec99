#!/usr/bin/env python3
"""CLUE benchmark on one MI355X: events/s of the batched graph path.

Workload (default): 4096 2-D events, sizes log-uniform in [256, 65536], each with
5-50 Gaussian blobs (sigma 1) in [0,100]^2 plus 20% uniform noise, weights uniform
in [0.5,1.5], gaussian kernel (sigma 1, amplitude 1), dc=1.5, rhoc=5, dm=2; fp32.

One step = clustering all events of the batch with one hipGraph launch (inputs
resident in HBM). Printed as one JSON line, with:
  - the plain per-event GPU path (one make_clusters call per event) on the same
    events, and the batched/per-event speed-up;
  - per-stage device times (hipEvents between the stages) and the roofline of
    the dominant stage (algorithmic bytes / average duration vs 8 TB/s);
  - the CPU baseline: the oracle (a restatement of the reference's serial
    backend), OpenMP over all host threads and serial, timed on a bounded sample
    of the same events on this host;
  - the latency of one batch (p50/p99 over launches waited for one at a time);
  - the neighbour stages' work units (pairs within the radius, counted once by the
    oracle) per second, and the VALU issue rate of the dominant kernel;
  - the served stream swept over offered rates (p50/p99 from arrival to labels);
  - bit-exactness of the GPU labels against that oracle sample, and of the
    per-event path against the batched one.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
N_CUS = 256
CLOCK_HZ = 2.4e9
VALU_PEAK_PER_CU_CYCLE = 2.0
METRIC = "events/s + p50/p99 latency, batched vs per-event GPU CLUE, bit-exact labels"


def _workload(name):
    from cluestering_amd import synthetic
    if name == "batch2d_4096":
        return synthetic.batch2d_4096()
    if name == "blobs2d_1m":
        return synthetic.blobs2d_1m()
    if name == "detector3d_1m":
        return synthetic.detector3d_1m()
    if name == "blobs2d_16k":
        return synthetic.blobs2d_16k()
    raise SystemExit(f"unknown workload {name}")


def _describe(w):
    if w.name == "batch2d_4096":
        return ("4096 2-D events, sizes log-uniform [256,65536] (total %d points), 5-50 blobs "
                "sigma 1 in [0,100]^2 + 20%% noise, weights U[0.5,1.5], gaussian kernel "
                "(1,1), dc=1.5 rhoc=5 dm=2, fp32" % w.n)
    return f"{w.name}: {w.n} points, {w.ndim}-D"


def _tiles_per_event(sizes, ndim, ppt=128):
    out = []
    for n in sizes:
        nt0 = (int(n) + ppt - 1) // ppt
        k = 1
        while k ** ndim < nt0:
            k += 1
        out.append(k ** ndim)
    return np.asarray(out, np.int64)


def _per_event_run(ctx, clus, pts, w, offsets):
    """Plain per-event path: one synchronous make_clusters per event (the drop-in
    make_clusters(queue, d_points) of the reference, core/detail/Clusterer.hpp:143-151)."""
    from cluestering_amd import _capi
    lib = _capi.load()
    p = clus.params(w.ndim, w.metric, w.kernel)
    lat = []
    labels = np.empty(w.n, np.int32)
    scratch = ctx.empty(w.n, np.int32)
    t0 = time.perf_counter()
    for e, n in enumerate(w.sizes):
        o = int(offsets[e])
        cp = _capi.Points()
        cp.n = int(n)
        for d in range(w.ndim):
            cp.coords[d] = pts.coords.address(d * w.n + o)
        cp.weight = pts.weight.address(o)
        r = _capi.Results()
        r.cluster_index = scratch.address(o)
        ts = time.perf_counter()
        _capi.check(lib.clue_make_clusters(ctx.handle, C.byref(p), C.byref(cp), C.byref(r)))
        lat.append(time.perf_counter() - ts)
    total = time.perf_counter() - t0
    labels[:] = scratch.to_host()
    return total, np.asarray(lat), labels


def _oracle_cfg(oracle, w):
    kernel = {0: "flat", 1: "exp", 2: "gaus"}[w.kernel.kind]
    if w.metric is None:
        metric, mp = "euclidean", ()
    else:
        mid, mp = w.metric.resolve(w.ndim)
        metric = {0: "euclidean", 1: "weighted_euclidean", 2: "periodic_euclidean",
                  3: "manhattan", 4: "chebyshev", 5: "weighted_chebyshev",
                  6: "mahalanobis"}[mid]
    return oracle.Config(ndim=w.ndim, dc=w.dc, rhoc=w.rhoc, dm=w.dm, kernel=kernel,
                         kparams=w.kernel.params, metric=metric, mparams=mp, wrapped=w.wrapped)


def cpu_baseline(w, offsets, gpu_labels, seconds):
    """Serial oracle over the events in order until `seconds` of CPU work, then the
    OpenMP oracle over the same events; returns both rates, the sample, and whether
    the GPU labels of the sample are bit-exact (batched ids = per-event ids + seeds
    of the previous events; the OpenMP labels must equal the serial ones)."""
    from oracle import oracle
    cfg = _oracle_cfg(oracle, w)
    done, exact, t_cpu, t_omp = 0, True, 0.0, 0.0
    first_id = 0
    points = 0
    serial_labels = []
    for e, n in enumerate(w.sizes):
        o, n = int(offsets[e]), int(n)
        c = np.ascontiguousarray(w.coords[:, o:o + n])
        t0 = time.perf_counter()
        r = oracle.run(cfg, c, w.weight[o:o + n], omp=False)
        t_cpu += time.perf_counter() - t0
        want = np.where(r.cluster >= 0, r.cluster + first_id, -1)
        exact = exact and bool(np.array_equal(gpu_labels[o:o + n], want))
        serial_labels.append(r.cluster)
        first_id += r.n_seeds
        done += 1
        points += n
        if t_cpu >= seconds:
            break
    omp_same = True
    for e in range(done):
        o, n = int(offsets[e]), int(w.sizes[e])
        c = np.ascontiguousarray(w.coords[:, o:o + n])
        t0 = time.perf_counter()
        r = oracle.run(cfg, c, w.weight[o:o + n], omp=True)
        t_omp += time.perf_counter() - t0
        omp_same = omp_same and bool(np.array_equal(r.cluster, serial_labels[e]))
    return {"serial_eps": done / t_cpu, "omp_eps": done / t_omp, "events": done, "points": points,
            "serial_s": t_cpu, "omp_s": t_omp, "exact": exact and omp_same,
            "omp_threads": oracle.num_threads(True)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch2d_4096")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-event", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rates", default="1000,2000,5000,10000,20000,50000",
                    help="offered rates (events/s) of the served stream, comma separated")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; multi-GPU runs (--gpus > 1) are not "
              "supported (events would shard across GPUs by whole events, one graph each).",
              file=sys.stderr)
        sys.exit(2)

    from cluestering_amd.device import Clusterer, Context, Plan, PointsDevice

    if args.calibrate_copy:
        from cluestering_amd import _capi
        ctx = Context(0)
        nbytes = args.calibrate_copy << 20
        src, dst = ctx.empty(nbytes, np.uint8), ctx.empty(nbytes, np.uint8)
        _capi.check(_capi.load().clue_debug_copy(ctx.handle, C.c_void_p(dst.ptr),
                                                 C.c_void_p(src.ptr), nbytes))
        print(json.dumps({"calibration_copy_bytes": nbytes}))
        return

    w = _workload(args.workload)
    sizes = w.sizes if w.sizes is not None else np.array([w.n], np.uint32)
    w.sizes = sizes
    offsets = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])[:-1]
    n_events = len(sizes)
    print(f"[bench] workload {w.name}: {n_events} events, {w.n} points", file=sys.stderr,
          flush=True)

    ctx = Context(0)
    pts = PointsDevice.from_host(ctx, w.coords, w.weight)
    clus = Clusterer(ctx, w.dc, w.rhoc, w.dm)
    clus.set_wrapped_coordinates(w.wrapped)
    plan = Plan(clus, pts, sizes, w.metric, w.kernel)

    for _ in range(args.warmup):
        plan.launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.launch()
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n_events / (ms_per_step / 1e3)
    print(f"[bench] batched graph: {ms_per_step:.3f} ms/step, {value:.1f} events/s",
          file=sys.stderr, flush=True)
    batched_labels = pts.cluster_index.to_host()

    # Latency of one batch: each launch waited for on its own (launch to completion
    # on the host clock), over as many launches as the timed loop.
    lat_ms = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        plan.launch()
        ctx.synchronize()
        lat_ms.append((time.perf_counter() - ts) * 1e3)
    lat_ms = np.asarray(lat_ms)

    # Per-stage device times: hipEvents between the stages on the context stream, on
    # the same kernels run eagerly (same batch, same launches as the graph).
    acc = {}
    reps = max(3, min(args.steps, 10))
    ctx.set_profiling(True)
    for _ in range(reps):
        clus.make_clusters_batched(pts, sizes, w.metric, w.kernel)
        for k, v in ctx.stage_times().items():
            acc[k] = acc.get(k, 0.0) + v / reps
    ctx.set_profiling(False)
    stage_us = {k: round(v, 2) for k, v in acc.items()}

    # Roofline of the dominant neighbour stage (algorithmic bytes per launch).
    tiles = _tiles_per_event(sizes, w.ndim)
    D = w.ndim
    alg = {
        "density": w.n * (4 * (D + 1) + 4) + 4 * int((tiles + 1).sum()),
        "nearest_higher": w.n * (4 * (D + 1) + 8) + 4 * int((tiles + 1).sum()),
    }
    dom = max(alg, key=lambda k: acc.get(k, 0.0))
    dom_us = acc[dom]
    achieved = alg[dom] / (dom_us * 1e-6) / 1e9
    # HBM bytes per launch and VALU instructions per launch of this kernel, from the
    # committed rocprofv3 counter runs of the same workload.
    prof = {}
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{w.name}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            prof = json.load(f)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                "traffic": prof.get("hbm_bytes_per_launch"),
                "kernel": f"k_{dom}", "alg_bytes_per_launch": alg[dom],
                "avg_launch_us": round(dom_us, 2)}
    valu = prof.get("valu_insts_per_launch")
    if valu:
        # Issue rate against 2 wave64 FP32 VALU instructions per CU per cycle
        # (157.3 TFLOP/s = 256 CUs x 2.4 GHz x 2 x 64 lanes x 2 flops).
        per_cu_cycle = valu / (dom_us * 1e-6) / (N_CUS * CLOCK_HZ)
        roofline["valu"] = {"insts_per_launch": valu, "issue_per_cu_cycle": round(per_cu_cycle, 3),
                            "peak_per_cu_cycle": VALU_PEAK_PER_CU_CYCLE,
                            "frac": round(per_cu_cycle / VALU_PEAK_PER_CU_CYCLE, 4),
                            "source": prof.get("source")}
        roofline["binding"] = "valu" if per_cu_cycle / VALU_PEAK_PER_CU_CYCLE > \
            achieved / PEAK_HBM_GBS else "hbm"
    pairs = None
    pfile = os.path.join(ROOT, "tests", "golden", f"pair_counts_{w.name}.json")
    if os.path.exists(pfile):
        with open(pfile) as f:
            pc = json.load(f)
        pairs = {"W_b": pc["W_b"], "W_c": pc["W_c"],
                 "density_pairs_per_s": round(pc["W_b"] / (acc["density"] * 1e-6), 1),
                 "nh_pairs_per_s": round(pc["W_c"] / (acc["nearest_higher"] * 1e-6), 1),
                 "density_visited": pc["visited_density"], "nh_visited": pc["visited_nh"]}
        if valu and dom == "density":
            pairs["density_lane_insts_per_pair"] = round(valu * 64 / pc["W_b"], 1)
    pipeline_bytes = w.n * (88 if D == 2 else 104)
    pipeline_us = sum(acc.values())

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "events/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": _describe(w), "events": int(n_events), "points": int(w.n),
                   "path": "batched, one hipGraph launch per step"},
        "roofline": roofline,
        "stage_us": stage_us,
        "pipeline_hbm_frac": round(pipeline_bytes / (pipeline_us * 1e-6) / 1e9 / PEAK_HBM_GBS, 5),
        "batched_latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4),
                               "p99": round(float(np.percentile(lat_ms, 99)), 4),
                               "launches": int(lat_ms.size)},
    }
    if pairs:
        out["pairs"] = pairs

    if not args.no_per_event:
        total, lat, labels = _per_event_run(ctx, clus, pts, w, offsets)
        # Per-event ids are local; the batched ids add the seeds of previous events.
        first = 0
        same = True
        for e, n in enumerate(sizes):
            o, n = int(offsets[e]), int(n)
            loc = labels[o:o + n]
            same = same and bool(np.array_equal(batched_labels[o:o + n],
                                                np.where(loc >= 0, loc + first, -1)))
            first += int(loc.max()) + 1 if n and loc.max() >= 0 else 0
        pe = n_events / total
        out["per_event"] = {"events_per_s": round(pe, 2),
                            "latency_us": {"p50": round(float(np.percentile(lat, 50)) * 1e6, 1),
                                           "p99": round(float(np.percentile(lat, 99)) * 1e6, 1)},
                            "labels_equal_batched": same}
        out["speedup_batched_vs_per_event"] = round(value / pe, 2)
        print(f"[bench] per-event path: {pe:.1f} events/s", file=sys.stderr, flush=True)

    if not args.no_cpu:
        cb = cpu_baseline(w, offsets, batched_labels, args.cpu_seconds)
        sample = (f"first {cb['events']} events ({cb['points']} points) of the same batch; "
                  f"oracle (restatement of the reference's serial backend; the OpenMP build "
                  f"runs the points of each stage in parallel)")
        out["cpu_baseline"] = {
            "value": round(cb["omp_eps"], 3), "unit": "events/s", "cores": cb["omp_threads"],
            "kind": "port", "sample": sample + f", OpenMP {cb['omp_s']:.2f} s",
            "serial": {"value": round(cb["serial_eps"], 3), "cores": 1,
                       "seconds": round(cb["serial_s"], 2)}}
        out["labels_bit_exact"] = {"vs_oracle_sample_events": cb["events"], "equal": cb["exact"]}
        print(f"[bench] cpu baseline: {cb['omp_eps']:.2f} events/s on {cb['omp_threads']} threads, "
              f"{cb['serial_eps']:.2f} events/s serial, over {cb['events']} events",
              file=sys.stderr, flush=True)

    plan.close()
    ctx.close()

    if not args.no_serve:
        # Served stream: Poisson arrivals, ndim 2 or 3, sizes log-uniform [256, 262144],
        # continuous batching under a 64 GB HBM budget; latency from arrival to labels.
        from cluestering_amd.serving import Server, make_pool, run_stream
        pool = make_pool(256)
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        served = []
        with Server(n_lanes=2, hbm_budget_bytes=64 << 30) as srv:
            # Warm-up at the top rate: the lanes grow once to the largest batches the
            # sweep forms, so no rate below pays for buffer growth.
            run_stream(srv, pool, rate=max(rates), n_events=3000, seed=99)
            for rate in rates:
                n_ev = int(min(max(rate * 2.0, 2000), 8000))
                r = run_stream(srv, pool, rate=rate, n_events=n_ev)
                row = {k: (round(v, 3) if isinstance(v, float) else v)
                       for k, v in r.items() if k != "latency_ms"}
                row["latency_ms"] = {k: round(v, 3) for k, v in r["latency_ms"].items()}
                served.append(row)
                print(f"[bench] served {rate:.0f}/s: {r['events_per_s']:.1f} events/s, "
                      f"p50 {r['latency_ms']['p50']:.2f} ms, p99 {r['latency_ms']['p99']:.2f} ms",
                      file=sys.stderr, flush=True)
        out["served"] = {"workload": "Poisson arrivals, ndim 2|3, sizes log-uniform [256,262144], "
                                     "2 lanes, 64 GB budget, latency arrival->labels on host",
                         "rates": served}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
