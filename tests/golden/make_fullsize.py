"""Writes tests/golden/fullsize.json: the pair count and pair hash of every
self-join workload at its full size, from the CPU oracle.

    python tests/golden/make_fullsize.py [--configs 2d 6d 18d 90d]

For each workload the points come from the oracle's copy of the reference
generator (mt19937 seed 0, exponential rate 40, redraw above 1; pinned
byte-for-byte against datasets/dataset_gen_fixed_len_expo_dist.cpp built from
source) and eps is the workload's (bench.py CONFIGS).  The algorithm-0 pair
set (grid stencil, fma chain, sqrt(acc) <= eps, cell-0 wrap) is counted and
hashed by og_join_stats (hash = sum of mix64(query << 32 | neighbour) mod
2^64, the library's pair_hash).

For the 2-D workload it also runs the reference's own algorithm 3 (Util.cpp /
MultiThreadJoin.cpp Super-EGO, compiled from source into oracle/_ref) and
explains the symmetric difference between the two pair sets pair by pair:
  - "cell0": the query's cell has coordinate 0 in an indexed dimension, so
    algorithm 0 emits nothing for it (kernel.cu:780 max(0, id - 1) on
    unsigned operands);
  - "boundary": |distance - eps| <= 1e-12 * eps, where the two algorithms'
    predicates (sqrt of an fma chain vs. an early-exit sum compared with
    eps^2 over reordered dimensions) may round differently.
Anything else is recorded as "unexplained" (the committed file has none).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bench import CONFIGS  # noqa: E402  (eps and sizes of the workloads)
from oracle import pyoracle as O  # noqa: E402

OUT = os.path.join(HERE, "fullsize.json")


def explain_alg3(pts, k, eps, a0_cnt, a0_hash):
    """Symmetric difference of algorithm 3 against algorithm 0, attributed."""
    t0 = time.time()
    c3, h3 = O.ref_superego_rowstats(pts, eps)
    t_alg3 = time.time() - t0
    differ = np.flatnonzero((c3 != a0_cnt) | (h3 != a0_hash)).astype(np.uint32)
    # per-query cell coordinates: a row is silenced by the cell-0 wrap when
    # any indexed coordinate is 0
    mn, _, _, _ = O.grid_dims(pts, k, eps)
    cell0 = np.zeros(len(pts), bool)
    for j in range(k):
        cell0 |= ((pts[:, j] - mn[j]) / eps).astype(np.int64) == 0
    o0, n0 = O.join_rows(pts, k, eps, differ)
    o3, n3 = O.ref_superego_rows(pts, eps, differ)
    stats = {"cell0_rows": 0, "cell0_pairs": 0, "boundary_pairs": 0, "unexplained_pairs": 0,
             "only_alg0": 0, "only_alg3": 0}
    worst = 0.0
    for i, q in enumerate(differ):
        r0 = set(n0[o0[i]:o0[i + 1]].tolist())
        r3 = set(n3[o3[i]:o3[i + 1]].tolist())
        diff = r0 ^ r3
        stats["only_alg0"] += len(r0 - r3)
        stats["only_alg3"] += len(r3 - r0)
        if cell0[q] and not r0:
            stats["cell0_rows"] += 1
            stats["cell0_pairs"] += len(diff)
            continue
        for c in diff:
            d = float(np.sqrt(np.sum((np.float128(1) * pts[c] - pts[q]) ** 2)))
            rel = abs(d - eps) / eps
            if rel <= 1e-12:
                stats["boundary_pairs"] += 1
                worst = max(worst, rel)
            else:
                stats["unexplained_pairs"] += 1
    stats["max_boundary_rel"] = worst
    return {"pairs": int(c3.sum(dtype=np.uint64)),
            "pair_hash": "%016x" % (int(h3.sum(dtype=np.uint64)) & (2 ** 64 - 1)),
            "differing_rows": int(len(differ)), "attribution": stats,
            "seconds": round(t_alg3, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", nargs="+", default=["2d", "6d"])
    a = ap.parse_args()
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for key in a.configs:
        c = CONFIGS[key]
        pts = O.generate_expo(c["n"], c["dim"], seed=0)
        t0 = time.time()
        rows = key == "2d"
        r = O.join_stats(pts, c["k"], c["eps"], rows=rows)
        entry = {"workload": c["name"], "n": c["n"], "dim": c["dim"], "indexed_dims": c["k"],
                 "eps": c["eps"], "seed": 0, "pairs": r[0], "pair_hash": "%016x" % r[1],
                 "cell0_silenced": int(O.count_cell0(pts, c["k"], c["eps"])),
                 "oracle_seconds": round(time.time() - t0, 1)}
        print(key, entry, flush=True)
        if rows:
            entry["alg3"] = explain_alg3(pts, c["k"], c["eps"], r[2], r[3])
            print(key, "alg3", entry["alg3"], flush=True)
        data[key] = entry
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main()
