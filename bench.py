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
oracle/_ref) timed on a seeded query sample on this host.  The default run
also reports the other workloads (2-D, 18-D, 90-D, serving at 4 QPS), each
run by this script in a child process, under `other_configs`
(--no-others skips them).
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
SERVE = dict(dim=6, n=10_000_000, k=6, nq_lo=1_000, nq_hi=500_000, eps_lo=0.002, eps_hi=0.02,
             name="serving: 6-D 10M expo FP64 resident index; Poisson arrivals; request queries "
                  "log-uniform 1k-500k, eps log-uniform [0.002, 0.02]")
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector with packed (2-wide) FP32 instructions
# The library compiles without packed FP32 (v_pk_*_f32), so an FP32 VALU
# instruction does one operation per lane: the roof it can reach is half the
# packed peak.
PEAK_FP32_UNPACKED_TFLOPS = PEAK_FP32_TFLOPS / 2
PEAK_BF16_TFLOPS = 2500.0 # MI355X dense bf16 MFMA
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


def host_cpu():
    """CPU model, CPUs visible to this process, and the threads the baseline
    uses: OMP_NUM_THREADS when set (a shared host sets it to the CPU share of
    one GPU), else every CPU the process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else visible
    return model, visible, min(threads, visible)


def cpu_baseline(pts, eps, dim, seconds_budget=20.0):
    """The reference's Super-EGO on a seeded query sample (oracle/_ref), or the
    oracle's restatement when that binary is absent."""
    from oracle import pyoracle as O
    n = len(pts)
    order = np.random.RandomState(7).permutation(n).astype(np.uint32)
    model, visible, threads = host_cpu()
    # probe 20000 queries (small probes overstate the per-query cost: thread
    # start-up) to size a sample of about seconds_budget of join time
    if O.ref_available(dim):
        probe_n = min(n, 20000)
        probe = O.ref_superego_time(pts, eps, order, probe_n, threads=threads)
        per_q = max(probe["join_s"], 1e-6) / probe_n
        nq = int(min(n, max(probe_n, seconds_budget / per_q)))
        r = O.ref_superego_time(pts, eps, order, nq, threads=threads)
        return {"value": r["pairs"] / r["join_s"], "unit": "result pairs/s",
                "cores": r["threads"], "kind": "reference", "cpu_model": model,
                "host_cpus_visible": visible,
                "sample": f"Super-EGO (reference Util.cpp/MultiThreadJoin.cpp built from source) "
                          f"join of {nq} seeded random queries of {n} points; reorder "
                          f"{r['reorder_s']:.2f} s + EGO sort {r['sort_s']:.2f} s not in the rate; "
                          f"join {r['join_s']:.2f} s, {r['pairs']} pairs",
                "join_s": r["join_s"], "queries": nq}
    qids = order[:2000]
    tot, t = O.superego_sample(pts, eps, qids, threads)
    return {"value": tot / max(t[2], 1e-9), "unit": "result pairs/s", "cores": threads,
            "kind": "port", "cpu_model": model, "host_cpus_visible": visible,
            "sample": f"Super-EGO restatement, {len(qids)} queries"}


def rooflines(n, dim, cells, pairs, evals, kernel_ms, refine):
    """The dominant join kernel against its two roofs, per launch.

    Algorithmic work (per unit, times the units one launch processes):
      HBM bytes   = 8 n D (coordinates once) + 4 n (indexLookupArr)
                    + 24 C (grid + gridCellLookup per cell) + 8 P (query and
                    neighbour id per pair)
      FP64 flops  = 3 D per distance evaluation E the kernel performs
                    (subtract + fma per dim; E is the reference's candidate
                    count less the candidates cell-box pruning leaves out)
      MFMA flops  = 2 D per candidate pair, against the bf16 peak divided by
                    the 3 products of a split-BF16 dot (hi.hi + hi.lo + lo.hi)
    Returns (binding, other): the binding roof is the one whose roof time is
    longer (the roof the kernel would hit first)."""
    t = kernel_ms * 1e-3
    bytes_alg = 8.0 * n * dim + 4.0 * n + 24.0 * cells + 8.0 * pairs
    hbm = {"bound": "hbm", "achieved": bytes_alg / t / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": bytes_alg / t / 1e9 / PEAK_HBM_GBS, "traffic": None,
           "work": bytes_alg, "roof_ms": bytes_alg / (PEAK_HBM_GBS * 1e9) * 1e3,
           "note": "8nD + 4n + 24C + 8P bytes per launch"}
    if refine == "mfma":
        flops = 2.0 * dim * evals
        peak = PEAK_BF16_TFLOPS / 3.0
        note = ("2 D flops per candidate pair; peak = bf16 dense MFMA / 3 (split-BF16 "
                "hi.hi + hi.lo + lo.hi); kernel k_gb_join")
        bound = "mfma"
    elif refine == "f32":
        flops = 3.0 * dim * evals
        peak = PEAK_FP32_UNPACKED_TFLOPS
        bound = "valu-fp32"
        note = ("3 D FP32 flops (sub + fma per dim) per distance evaluation against the FP32 "
                "vector peak without packed FP32 (157.3 / 2 TFLOP/s; the build does not emit "
                "packed FP32); kernel k_join with the FP32 filter")
    else:
        flops = 3.0 * dim * evals
        peak = PEAK_FP64_TFLOPS
        bound = "valu-fp64"
        note = "3 D FP64 flops (sub + fma per dim) per distance evaluation; kernel k_join"
    comp = {"bound": bound, "achieved": flops / t / 1e12, "peak": peak, "unit": "TFLOP/s",
            "frac": flops / t / 1e12 / peak, "traffic": None, "work": flops,
            "roof_ms": flops / (peak * 1e12) * 1e3, "note": note}
    return (comp, hbm) if comp["roof_ms"] >= hbm["roof_ms"] else (hbm, comp)


def attach_traffic(roof, config, refine, kernel_ms):
    """HBM traffic per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload (each
    counter scaled by its reading of the bench's 1 GiB copy kernel)."""
    tag = config + {"mfma": "_mfma", "f32": "_f32"}.get(refine, "")
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return
    with open(path) as f:
        t = json.load(f).get(tag)
    if t:
        roof["traffic"] = t["traffic_bytes"]
        roof["traffic_source"] = (f"{t['profile']} (profiled launch {t['kernel_avg_ms']:.2f} ms; "
                                  f"this run {kernel_ms:.2f} ms)")


def serve_levels(eps_lo, eps_hi, ratio=2 ** 0.125, max_levels=32):
    """Grid levels for the serving index: cell sides from just above eps_hi
    down by `ratio` while they still exceed eps_lo / ratio, so a request's
    cell side is at most `ratio` times its eps."""
    top = eps_hi * (1 + 1e-6)
    sides = [top]
    while sides[-1] / ratio >= eps_lo * 0.999 and len(sides) < max_levels:
        sides.append(sides[-1] / ratio)
    return sorted(sides)


def serve_main(a):
    """Serving: requests arrive as a Poisson stream at --qps against the
    resident index; latency = submit -> result ready in device memory."""
    import threading
    import hegjoin_amd as H
    c = SERVE
    t0 = time.time()
    pts = H.generate_expo(c["n"], c["dim"], seed=0)
    sides = serve_levels(c["eps_lo"], c["eps_hi"])
    srv = H.Server(pts, sides, indexed_dims=c["k"], slots=a.slots,
                   result_bytes=int(a.pool_gb * (1 << 30)), max_queries=c["nq_hi"])
    del pts
    print(f"# index {c['n']}x{c['dim']} at {len(sides)} levels in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(a.serve_seed)
    total = a.warmup + a.requests
    nqs = np.exp(rng.uniform(np.log(c["nq_lo"]), np.log(c["nq_hi"]), total)).astype(np.int64)
    epss = np.exp(rng.uniform(np.log(c["eps_lo"]), np.log(c["eps_hi"]), total))
    gaps = rng.exponential(1.0 / a.qps, total)
    queries = [H.generate_expo(int(nqs[i]), c["dim"], seed=100_000 + i) for i in range(total)]

    def run(lo, hi, timed):
        tickets = [None] * (hi - lo)
        ready = [threading.Event() for _ in range(hi - lo)]
        start = time.perf_counter()
        arrivals = start + np.cumsum(gaps[lo:hi]) - gaps[lo]

        def submitter():
            for j in range(hi - lo):
                d = arrivals[j] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[j] = srv.submit(queries[lo + j], epss[lo + j])
                ready[j].set()
        th = threading.Thread(target=submitter)
        th.start()
        # collect in completion order and release at once (a client that held
        # finished results while waiting for another request could starve
        # admission of the pool)
        infos, left = [], list(range(hi - lo))
        while left:
            for j in list(left):
                if not ready[j].is_set():
                    continue
                try:
                    info = srv.poll(tickets[j])
                    if info is None:
                        continue
                    info["ok"] = True
                except H.HegjError as e:
                    info = {"ok": False, "error": str(e), "n_queries": int(nqs[lo + j])}
                info["eps"] = float(epss[lo + j])
                srv.release(tickets[j])
                infos.append(info)
                left.remove(j)
                if timed:
                    print(f"# request {j}: nq {info['n_queries']} eps {info['eps']:.4f} "
                          + (f"pairs {info['pairs']} latency {info['latency_ms']:.1f} ms"
                             if info["ok"] else "rejected: " + info["error"]),
                          file=sys.stderr, flush=True)
            time.sleep(0.0005)
        th.join()
        return infos, time.perf_counter() - start

    if a.warmup:
        run(0, a.warmup, False)
    srv.stats(reset=True)
    infos, wall = run(a.warmup, total, True)
    st = srv.stats()
    srv.close()
    ok = [i for i in infos if i["ok"]]
    pairs = sum(i["pairs"] for i in ok)
    evals = sum(i["evals"] for i in ok)
    out = {
        "metric": "serving p50/p99 latency at fixed QPS, result pairs/s (1 GPU)",
        "value": pairs / wall, "unit": "pairs/s",
        "n_gpus": 1, "steps": a.requests, "warmup": a.warmup,
        "ms_per_step": wall * 1000 / max(a.requests, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (reference exponential generator)",
        "config": {"workload": c["name"], "n": c["n"], "dim": c["dim"], "qps": a.qps,
                   "requests": a.requests, "slots": a.slots, "pool_gb": a.pool_gb,
                   "levels": [round(x, 6) for x in sides], "seed": a.serve_seed},
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
        "mean_ms": st["mean_ms"], "completed": len(ok), "rejected": len(infos) - len(ok),
        "distance_evals_per_s": evals / wall, "pool_peak_bytes": st["pool_peak"],
        "wall_s": wall,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="6d", choices=sorted(CONFIGS) + ["serve"])
    ap.add_argument("--qps", type=float, default=0.5, help="serving: request arrival rate")
    ap.add_argument("--requests", type=int, default=20, help="serving: timed requests")
    ap.add_argument("--slots", type=int, default=4, help="serving: requests on the GPU at once")
    ap.add_argument("--pool-gb", type=float, default=96.0, help="serving: result-buffer pool")
    ap.add_argument("--serve-seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--queries-per-unit", type=int, default=None)
    ap.add_argument("--unit-order", default=None, choices=["spatial", "workload", "morton"])
    ap.add_argument("--xcd-parts", type=int, default=None)
    ap.add_argument("--parts-by-block", type=int, default=0)
    ap.add_argument("--grab", type=int, default=None)
    ap.add_argument("--refine", default=None, choices=["auto", "fp64", "mfma", "f32"])
    ap.add_argument("--graph", action="store_true",
                    help="replay the captured batch graph (FP64 refine)")
    ap.add_argument("--n", type=int, default=0, help="override the point count (experiments)")
    ap.add_argument("--no-prune", action="store_true",
                    help="evaluate every point of every stencil cell (no cell-box pruning)")
    ap.add_argument("--no-others", action="store_true",
                    help="default run: skip the other workloads' summary (other_configs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.config == "serve":
        return serve_main(a)
    cfg = dict(CONFIGS[a.config])
    if a.n:
        cfg["n"] = a.n
        cfg["name"] += f" (n overridden to {a.n})"
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
    # options not given on the command line take the tuned defaults (tuning.tsv)
    kw = dict(part_by_block=a.parts_by_block, use_graph=a.graph, prune=not a.no_prune)
    for key, val in (("queries_per_unit", a.queries_per_unit), ("unit_order", a.unit_order),
                     ("xcd_parts", a.xcd_parts), ("units_per_grab", a.grab), ("refine", a.refine)):
        if val is not None:
            kw[key] = val
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
    refine_used = infos[-1]["refine_used"]
    per_launch = launches / a.steps
    # compute work: the distance evaluations the kernel performed (cell-box
    # pruning leaves out candidates the reference evaluates)
    evals_launch = sum(i["evals_executed"] for i in infos) / launches
    roof, roof_alt = rooflines(n, dim, iinfo["nNonEmptyCells"], pairs, evals_launch,
                               kernel_ms, refine_used)
    roof_alt["attainable_copy_gbs"] = copy_gbs
    attach_traffic(roof, a.config, refine_used, kernel_ms)
    out = {
        "metric": "result pairs/s (1 GPU)",
        "value": pairs * a.steps / elapsed,
        "unit": "pairs/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mfma": "bf16 hi/lo (f32 acc) + f64 band recheck",
                  "f32": "f32 filter + f64 band recheck"}.get(refine_used, "f64"), "data": "synthetic (reference exponential generator, seed 0)",
        "config": {"workload": cfg["name"], "n": n, "dim": dim, "indexed_dims": k,
                   "eps": eps, "pairs": pairs, "total_candidates": total_cand,
                   "cells": iinfo["nNonEmptyCells"], "cell0_silenced": iinfo["n_cell0_queries"],
                   "parallelism": "1 GPU"},
        "distance_evals_per_s": sum(i["evals_executed"] for i in infos) / elapsed,
        "evals_vs_reference_candidates": infos[-1]["evals_executed"] / max(total_cand, 1),
        "prune": not a.no_prune,
        "pair_hash": "%016x" % infos[0]["pair_hash"],
        "join_kernel_ms": kernel_ms,
        "estimate_ms": infos[-1]["estimate_ms"],
        "refine": refine_used,
        "band_rechecks": infos[-1]["band_rechecks"],
        "roofline": roof,
        "roofline_other": roof_alt,
    }
    ix.close()
    if not a.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(pts, eps, dim, a.cpu_seconds)
        print(f"# cpu baseline {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if a.config == "6d" and not a.no_others and not a.n:
        out["other_configs"] = other_configs()
    print(json.dumps(out))


OTHER_RUNS = [
    ("2d", ["--config", "2d"]),
    ("18d", ["--config", "18d"]),
    ("90d", ["--config", "90d"]),
    ("serve_4qps", ["--config", "serve", "--qps", "4", "--requests", "30"]),
]
OTHER_KEYS = ("value", "unit", "ms_per_step", "join_kernel_ms", "refine", "pair_hash", "band_rechecks",
              "distance_evals_per_s", "p50_ms", "p99_ms", "max_ms", "mean_ms", "completed", "rejected")


def other_configs(timeout_s=400):
    """The other workloads of the benchmark, each in a child process with its
    own defaults (no CPU baseline): their key figures and roofline fraction,
    so one default run reports every workload."""
    import subprocess
    res = {}
    for key, args in OTHER_RUNS:
        t0 = time.time()
        cmd = [sys.executable, os.path.abspath(__file__), *args, "--no-cpu-baseline", "--no-others"]
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
            lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
            if p.returncode != 0 or not lines:
                res[key] = {"error": f"exit {p.returncode}", "stderr_tail": p.stderr[-400:]}
            else:
                d = json.loads(lines[-1])
                r = {k: d[k] for k in OTHER_KEYS if k in d}
                r["workload"] = d.get("config", {}).get("workload")
                if d.get("roofline"):
                    r["roofline"] = {k: d["roofline"].get(k) for k in ("bound", "achieved", "peak", "unit", "frac")}
                res[key] = r
        except subprocess.TimeoutExpired:
            res[key] = {"error": f"timed out after {timeout_s} s"}
        print(f"# {key}: {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    return res


if __name__ == "__main__":
    main()
