"""Writes tests/golden/pair_counts_batch2d_4096.json: the exact work units of the
neighbour stages on the 4096-event batch (sizes log-uniform in [256, 65536], 2-D,
gaussian kernel (1, 1), dc 1.5, rhoc 5, dm 2), counted by the oracle per event:

  W_b = #{(i, j) : d_ij <= dc}, i = j included (density terms),
  W_c = #{(i, j) : j higher than i and d_ij <= eff_dist_i} (nearest-higher candidates),

plus the candidates each stage visits in its search boxes. bench.py reports the
neighbour stages' pairs per second from these counts. Run once on the CPU:
    python tests/golden/make_pair_counts.py      (OpenMP over all host cores)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cluestering_amd import synthetic  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    w = synthetic.batch2d_4096()
    cfg = oracle.Config(ndim=2, dc=w.dc, rhoc=w.rhoc, dm=w.dm, kernel="gaus",
                        kparams=w.kernel.params)
    tot = {"visited_density": 0, "in_radius": 0, "visited_nh": 0, "nh_pairs": 0}
    off = 0
    t0 = time.time()
    for e, n in enumerate(w.sizes):
        n = int(n)
        c = np.ascontiguousarray(w.coords[:, off:off + n])
        wt = w.weight[off:off + n]
        r = oracle.run(cfg, c, wt, omp=True)
        pc = oracle.pair_counts(cfg, c, wt, rho=r.rho, omp=True)
        for k in tot:
            tot[k] += pc[k]
        off += n
        if e % 256 == 255:
            print(f"{e + 1} events, {time.time() - t0:.0f} s", flush=True)
    out = {"workload": "batch2d_4096", "events": int(len(w.sizes)), "points": int(w.n),
           "W_b": tot["in_radius"], "W_c": tot["nh_pairs"],
           "visited_density": tot["visited_density"], "visited_nh": tot["visited_nh"]}
    path = os.path.join(ROOT, "tests", "golden", "pair_counts_batch2d_4096.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
