"""bench.py -- result pairs/s of the MI355X epsilon-grid self-join (algorithm 0).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 6d]

A step is one self-join pass (hegj_selfjoin: batch estimate, persistent join
kernel, neighbour table on the device) over the resident grid index of the
configured dataset.  The default workload is the 6-D one: 5,000,000 points
from the reference's exponential generator (mt19937 seed 0, rate 40, values
in [0, 1]), FP64, NUMINDEXEDDIM = 6, eps calibrated from a seeded 10k sample
for a mean of 64 neighbours per point.  Inputs are resident in HBM before the
timed region; the timed region is K steps between two device syncs.

Prints one JSON line with value = result pairs/s (self pairs included, as the
reference's [RESULT] count), the join kernel's roofline (live HIP events on
the library's stream) and a CPU baseline: the reference's own Super-EGO
(algorithm 3, Util.cpp / MultiThreadJoin.cpp compiled from source into
oracle/_ref) timed on a seeded query sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# eps values: m-th smallest exact distance of a 10k-point sample
# (np.random.RandomState(1).choice without replacement) of the seed-0
# dataset, m = round(target * ns * (ns - 1) / 2 / (N - 1)); `--calibrate`
# recomputes them.
CONFIGS = {
    "2d": dict(dim=2, n=1_000_000, k=2, eps=0.001, target=None,
               name="2-D 1M expo FP64 k=2 eps=0.001"),
    "6d": dict(dim=6, n=5_000_000, k=6, eps=0.0063131373103199606, target=64,
               name="6-D 5M expo FP64 k=6 eps calibrated for 64 neighbours"),
    "18d": dict(dim=18, n=2_000_000, k=6, eps=0.041121185939337571, target=32,
                name="18-D 2M expo FP64 k=6 eps calibrated for 32 neighbours"),
    "90d": dict(dim=90, n=1_000_000, k=6, eps=0.19773135137186568, target=16,
                name="90-D 1M expo FP64 k=6 eps calibrated for 16 neighbours"),
}
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def calibrate(dim, n, target, ns=10000):
    import hegjoin_amd as H
    pts = H.generate_expo(n, dim, seed=0)
    s = pts[np.sort(np.random.RandomState(1).choice(n, ns, replace=False))]
    parts = []
    for i in range(0, ns, 500):
        d2 = ((s[i:i + 500, None, :] - s[None, :, :]) ** 2).sum(-1)
        for a, row in zip(range(i, min(i + 500, ns)), d2):
            parts.append(np.sqrt(row[a + 1:]))
    d = np.concatenate(parts)
    m = int(round(target * ns * (ns - 1) / 2 / (n - 1)))
    return float(np.partition(d, m)[m])


def cpu_baseline(pts, eps, dim, seconds_budget=20.0):
    """The reference's Super-EGO on a seeded query sample (oracle/_ref), or the
    oracle's restatement when that binary is absent."""
    from oracle import pyoracle as O
    n = len(pts)
    order = np.random.RandomState(7).permutation(n).astype(np.uint32)
    # probe 2000 queries to size a sample of about seconds_budget of join time
    if O.ref_available(dim):
        probe = O.ref_superego_time(pts, eps, order, 2000)
        per_q = max(probe["join_s"], 1e-6) / 2000
        nq = int(min(n, max(2000, seconds_budget / per_q)))
        r = O.ref_superego_time(pts, eps, order, nq)
        return {"value": r["pairs"] / r["join_s"], "unit": "result pairs/s",
                "cores": r["threads"], "kind": "reference",
                "sample": f"Super-EGO (reference Util.cpp/MultiThreadJoin.cpp built from source) "
                          f"join of {nq} seeded random queries of {n} points; reorder "
                          f"{r['reorder_s']:.2f} s + EGO sort {r['sort_s']:.2f} s not in the rate; "
                          f"join {r['join_s']:.2f} s, {r['pairs']} pairs",
                "join_s": r["join_s"], "queries": nq}
    threads = os.cpu_count() or 1
    qids = order[:2000]
    tot, t = O.superego_sample(pts, eps, qids, threads)
    return {"value": tot / max(t[2], 1e-9), "unit": "result pairs/s", "cores": threads,
            "kind": "port", "sample": f"Super-EGO restatement, {len(qids)} queries"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="6d", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--queries-per-unit", type=int, default=0)
    ap.add_argument("--unit-order", default="morton")
    ap.add_argument("--xcd-parts", type=int, default=0)
    ap.add_argument("--parts-by-block", type=int, default=0)
    ap.add_argument("--grab", type=int, default=0)
    ap.add_argument("--refine", default="auto", choices=["auto", "fp64", "mfma"])
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    cfg = CONFIGS[a.config]
    if a.calibrate:
        for key, c in CONFIGS.items():
            if c["target"]:
                print(key, "%.17g" % calibrate(c["dim"], c["n"], c["target"]), flush=True)
        return

    import hegjoin_amd as H
    dim, n, k, eps = cfg["dim"], cfg["n"], cfg["k"], cfg["eps"]
    # a plain 1 GiB device copy: the attainable HBM roof, and the reference
    # kernel that scales rocprof FETCH_SIZE / WRITE_SIZE
    copy_gbs = H.stream_copy_bandwidth(1 << 30)
    t0 = time.time()
    pts = H.generate_expo(n, dim, seed=0)
    print(f"# generated {n}x{dim} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ix = H.GridIndex(pts, eps, indexed_dims=k)
    iinfo = ix.info
    cells, _, total_cand = ix.sort_by_workload()
    print(f"# index: {iinfo['nNonEmptyCells']} cells, build {iinfo['build_ms']:.1f} ms, "
          f"candidates {total_cand}", file=sys.stderr, flush=True)
    kw = dict(queries_per_unit=a.queries_per_unit, unit_order=a.unit_order,
              xcd_parts=a.xcd_parts, units_per_grab=a.grab, part_by_block=a.parts_by_block,
              refine=a.refine)
    for _ in range(a.warmup):
        r = ix.selfjoin(**kw)
        r.close()
    infos = []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        r = ix.selfjoin(**kw)
        infos.append(r.info)
        r.close()
    elapsed = time.perf_counter() - t_start
    print(f"# timed {a.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
    pairs = infos[-1]["total_neighbors"]
    assert all(i["total_neighbors"] == pairs and i["pair_hash"] == infos[0]["pair_hash"]
               for i in infos), "steps disagree"
    ms_step = elapsed * 1000 / a.steps
    # dominant kernel: the join; one launch per step when one batch suffices
    launches = sum(i["join_launches"] for i in infos)
    kernel_ms = sum(i["join_ms"] for i in infos) / launches
    flops_alg = 3.0 * dim * total_cand / (launches / a.steps)  # sub + fma per dim per eval
    achieved_tf = flops_alg / (kernel_ms * 1e-3) / 1e12
    bytes_alg = 8.0 * n * dim + 4.0 * n + 24.0 * iinfo["nNonEmptyCells"] + 4.0 * pairs + 12.0 * n
    out = {
        "metric": "result pairs/s (1 GPU)",
        "value": pairs * a.steps / elapsed,
        "unit": "pairs/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (reference exponential generator, seed 0)",
        "config": {"workload": cfg["name"], "n": n, "dim": dim, "indexed_dims": k,
                   "eps": eps, "pairs": pairs, "total_candidates": total_cand,
                   "cells": iinfo["nNonEmptyCells"], "cell0_silenced": iinfo["n_cell0_queries"],
                   "parallelism": "1 GPU"},
        "distance_evals_per_s": total_cand * a.steps / elapsed,
        "pair_hash": "%016x" % infos[0]["pair_hash"],
        "join_kernel_ms": kernel_ms,
        "estimate_ms": infos[-1]["estimate_ms"],
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tf / PEAK_FP64_TFLOPS, "traffic": None,
                     "note": "FP64 compute bound: 3*dim flops (sub + fma per dim) per "
                             "reference candidate; FP64 vector and matrix peaks are equal"},
        "roofline_hbm": {"bound": "hbm", "achieved": bytes_alg / (kernel_ms * 1e-3) / 1e9,
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": bytes_alg / (kernel_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                         "attainable_copy": copy_gbs},
    }
    ix.close()
    if not a.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(pts, eps, dim, a.cpu_seconds)
        print(f"# cpu baseline {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
