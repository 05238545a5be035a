"""Autotuner of the join options, per dimensionality, on the GPU.

    python -m hegjoin_amd.autotune --dim 6 --n 5000000 --eps 0.0063131373103199606
    python -m hegjoin_amd.autotune --dim 18 --n 2000000 --eps 0.041121185939337571

For one workload (the reference's exponential generator, seed 0) it sweeps
the runtime choices of the join -- queries per work unit (the points that
share one streamed candidate tile), unit order, per-XCD unit ranges, units
per queue grab, and at 16 dimensions and above the refine (FP64 chain or the
split-BF16 MFMA guard band) -- times each configuration, and checks that all
of them produce the same pair count and pair hash.  A configuration whose
result differs is an error, not a candidate.  A second stage sweeps, around
the first stage's best grid-join configuration, the workgroup size (1, 2 or
4 waves: the reference's BLOCKSIZE, params.h) and the number of device
result batches (1, 2, 4, 8: the reference's batch count, GPU.cu).

The fastest configuration (host wall time of hegj_selfjoin, mean over
--steps runs; among those within 1% of it, the one with the fastest join
kernel) is written to tuning.tsv (one line per dimensionality: dim
queries_per_unit unit_order xcd_parts units_per_grab refine block_waves
batches), which the library, the Python package and the CLI load at start;
every measurement goes to tuning.json.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

ORDERS = {"spatial": 0, "workload": 1, "morton": 2}
REFINES = {"auto": 0, "fp64": 1, "mfma": 2, "f32": 3}
HERE = os.path.dirname(os.path.abspath(__file__))


def candidates(dim, with_fp64=False):
    """Option sets to try for one dimensionality.  At 16 dims and above the
    MFMA refine is swept; the FP64 kernel only with with_fp64 (at the full
    18-D/90-D sizes one FP64 join takes minutes)."""
    out = []
    if dim >= 16:
        for order in ("morton", "workload", "spatial"):
            out.append(dict(refine="mfma", unit_order=order, queries_per_unit=0, xcd_parts=1,
                            units_per_grab=0))
        if with_fp64:
            for qb in (4, 8):
                out.append(dict(refine="fp64", unit_order="morton", queries_per_unit=qb,
                                xcd_parts=8, units_per_grab=0))
        return out
    refines = ("fp64", "f32") if dim in (2, 6) else ("fp64",)
    for refine, qb, order, parts, grab in itertools.product(
            refines, (4, 8, 16), ("morton", "workload", "spatial"), (1, 8), (0, 1, 4)):
        if refine == "f32" and qb * dim > 48:
            continue  # not instantiated
        out.append(dict(refine=refine, unit_order=order, queries_per_unit=qb, xcd_parts=parts,
                        units_per_grab=grab))
    return out


WAVES = (4, 2, 1)
BATCHES = (1, 2, 4, 8)


def second_stage(best):
    """Workgroup sizes and batch counts around a grid-join configuration (the
    MFMA refine has its own fixed tiling and one pair list)."""
    if best.get("refine") == "mfma":
        return []
    base = {k: best[k] for k in ("refine", "unit_order", "queries_per_unit", "xcd_parts",
                                 "units_per_grab")}
    return [dict(base, block_waves=w, batches=b) for w in WAVES for b in BATCHES
            if (w, b) != (4, 1)]


def measure(ix, opts, steps, limit_s):
    """Mean wall ms of hegj_selfjoin over `steps` runs (after one warm-up),
    with the result's pair count and hash.  Stops early past limit_s."""
    r = ix.selfjoin(**opts)
    info = r.info
    r.close()
    if info["total_ms"] > limit_s * 1000:
        return info["total_ms"], info, 1
    times = []
    for _ in range(steps):
        r = ix.selfjoin(**opts)
        info = r.info
        r.close()
        times.append(info["total_ms"])
    return sum(times) / len(times), info, len(times)


def tune(dim, n, eps, k=None, steps=2, limit_s=20.0, log=print, with_fp64=False):
    import hegjoin_amd as H
    pts = H.generate_expo(n, dim, seed=0)
    k = k or min(dim, 6)
    results = []
    ref = None
    def run(configs):
        nonlocal ref
        for opts in configs:
            try:
                ms, info, runs = measure(ix, opts, steps, limit_s)
            except H.HegjError as e:
                log(f"# skip {opts}: {e}")
                continue
            key = (info["total_neighbors"], info["pair_hash"])
            if ref is None:
                ref = key
            if key != ref:
                raise RuntimeError(f"configuration {opts} changed the result: {key} != {ref}")
            row = dict({"block_waves": 4, "batches": 1}, **opts, ms=ms, runs=runs,
                       kernel_ms=info["join_kernel_avg_ms"], batches_run=info["n_batches"],
                       pairs=info["total_neighbors"], pair_hash="%016x" % info["pair_hash"])
            results.append(row)
            log(json.dumps(row))

    with H.GridIndex(pts, eps, indexed_dims=k) as ix:
        run(candidates(dim, with_fp64))
        run(second_stage(pick_best(results)))
    return pick_best(results), results


def pick_best(results, tie=0.01):
    """Fastest wall time; among configurations within `tie` (relative) of it,
    the one with the fastest join kernel (wall-time differences that small
    are run-to-run noise, and the kernel time is the steadier figure)."""
    fastest = min(r["ms"] for r in results)
    close = [r for r in results if r["ms"] <= fastest * (1 + tie)]
    return min(close, key=lambda r: r["kernel_ms"])


def write_tables(dim, workload, best, results, tsv=None, js=None):
    tsv = tsv or os.path.join(HERE, "tuning.tsv")
    js = js or os.path.join(HERE, "tuning.json")
    table = {}
    if os.path.exists(js):
        with open(js) as f:
            table = json.load(f)
    table[str(dim)] = dict(workload=workload, best=best, results=results)
    with open(js, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    lines = ["# dim queries_per_unit unit_order xcd_parts units_per_grab refine block_waves batches",
             "# written by hegjoin_amd/autotune.py; workloads and timings in tuning.json"]
    for d in sorted(table, key=int):
        b = table[d]["best"]
        lines.append(f"{d} {b['queries_per_unit']} {ORDERS[b['unit_order']]} {b['xcd_parts']} "
                     f"{b['units_per_grab']} {REFINES[b['refine']]} {b.get('block_waves', 4)} "
                     f"{b.get('batches', 1)}")
    with open(tsv, "w") as f:
        f.write("\n".join(lines) + "\n")
    return tsv, js


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--eps", type=float, required=True)
    ap.add_argument("--k", type=int, default=None)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--limit-s", type=float, default=20.0)
    ap.add_argument("--with-fp64", action="store_true", help="also time the FP64 kernel at dim >= 16")
    ap.add_argument("--tsv", default=None)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    t0 = time.time()
    best, results = tune(a.dim, a.n, a.eps, a.k, a.steps, a.limit_s,
                         log=lambda s: print(s, file=sys.stderr, flush=True), with_fp64=a.with_fp64)
    workload = dict(dim=a.dim, n=a.n, eps=a.eps, k=a.k or min(a.dim, 6), generator="expo seed 0")
    tsv, js = write_tables(a.dim, workload, best, results, a.tsv, a.json)
    print(json.dumps(dict(best=best, configs=len(results), seconds=time.time() - t0, tsv=tsv)))


if __name__ == "__main__":
    main()
