"""Writes the golden fixtures in tests/golden/ (run here, where the reference
sources are; the fixtures are data and travel with the tests).

    python tests/golden/make_golden.py

Each fixture holds the points, eps, k and the expected outputs:
  * alg0_*.npz   -- algorithm-0 neighbour table (CSR, rows ascending), per
                    query counts, multiset hash, grid index and workload,
                    computed by the oracle (oracle/hegj_oracle.cpp) and
                    cross-checked here against the O(n^2) brute force before
                    being written;
  * cell0_*.npz  -- a dataset crafted so that its minimum lands in grid cell 0
                    (the reference's cell-0 wrap silences that query);
  * ref_alg3_*.npz -- the reference's own Super-EGO (Util::egoJoinV2 from
                    Util.cpp, compiled from source into oracle/_ref) pair set,
                    with the points it was run on.
  * ref_gen_*.txt  -- text written by the reference's dataset generator
                    (datasets/dataset_gen_fixed_len_expo_dist.cpp compiled
                    from source), first 200 points of `genDataset 2 2000`.
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import pyoracle as O  # noqa: E402

CASES = [  # name, dim, k, n, seed, mean neighbours (eps calibrated)
    ("alg0_d2_n3000", 2, 2, 3000, 3, 40),
    ("alg0_d6_n3000", 6, 6, 3000, 5, 30),
    ("alg0_d6k3_n2500", 6, 3, 2500, 6, 30),
    ("alg0_d18_n1500", 18, 6, 1500, 7, 20),
    ("alg0_d90_n800", 90, 6, 800, 8, 12),
    ("alg0_d3k1_n2000", 3, 1, 2000, 9, 25),
]


def calibrate_eps(pts, target):
    """eps = the distance at which the mean neighbour count (self excluded)
    is `target`, rounded to 4 significant digits."""
    n = len(pts)
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    d = np.sqrt(d2[np.triu_indices(n, 1)])
    m = int(round(target * n / 2))
    eps = float(np.partition(d, m)[m])
    return float(f"{eps:.4g}")


def craft_cell0(pts, eps, dim_j=0):
    """Shifts coordinate dim_j by a constant v and puts its minimum exactly at
    v, choosing v so that (v - (v - eps)) / eps < 1 in FP64: that point then
    lands in grid cell 0.  (The rounding error of v - eps depends only on eps
    and the binade of v, so v is searched over a grid of shifts.)"""
    for v in np.linspace(0.02, 0.7, 2000):
        v = float(v)
        if np.floor((v - (v - eps)) / eps) == 0:
            pts = pts.copy()
            pts[:, dim_j] += v
            i = int(np.argmin(pts[:, dim_j]))
            pts[i, dim_j] = v
            return pts, i
    raise RuntimeError("no cell-0 value found")


def alg0_fixture(name, pts, k, eps):
    off, nb = O.join(pts, k, eps)
    boff, bnb = O.brute(pts, k, eps)
    assert np.array_equal(off, boff) and np.array_equal(nb, bnb), name
    ila, lookup, grid = O.index_build(pts, k, eps)
    work, sorted_cells, opi, tot = O.workload(pts, k, eps)
    mn, mx, nc, tc = O.grid_dims(pts, k, eps)
    noff, nnb = O.join(pts, k, eps, emulate_cell0=False)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"), pts=pts, eps=eps, k=k, offsets=off, nbrs=nb,
        hash=np.uint64(O.csr_hash(off, nb)), cell_ids=O.point_cells(pts, k, eps),
        indexLookupArr=ila, cell_keys=lookup["gridLinearID"], grid_min=grid["indexmin"],
        grid_max=grid["indexmax"], work=work, sorted_cells=sorted_cells, originPointIndex=opi,
        totalCandidates=np.uint64(tot), minArr=mn, maxArr=mx, nCells=nc, totalCells=np.uint64(tc),
        nowrap_offsets=noff, nowrap_nbrs=nnb, n_cell0=np.uint32(O.count_cell0(pts, k, eps)),
        threshold=O.threshold(eps))
    print(name, "pairs", len(nb), "cells", len(lookup), "dead", O.count_cell0(pts, k, eps))


def main():
    for name, dim, k, n, seed, target in CASES:
        pts = O.generate_expo(n, dim, seed=seed)
        alg0_fixture(name, pts, k, calibrate_eps(pts, target))
    # crafted cell-0 datasets
    pts = O.generate_expo(3000, 2, seed=21)
    pts, i = craft_cell0(pts, 0.004)
    assert O.count_cell0(pts, 2, 0.004) >= 1
    alg0_fixture("cell0_d2_n3000", pts, 2, 0.004)
    pts = O.generate_expo(2000, 6, seed=22)
    pts, i = craft_cell0(pts, 0.03, dim_j=4)
    assert O.count_cell0(pts, 6, 0.03) >= 1
    alg0_fixture("cell0_d6_n2000", pts, 6, 0.03)
    # the reference's own Super-EGO pair sets
    if O.ref_available(6):
        for dim, n, seed, target in [(2, 3000, 31, 40), (6, 2500, 32, 30), (18, 1200, 33, 20),
                                     (90, 600, 34, 12)]:
            pts = O.generate_expo(n, dim, seed=seed)
            eps = calibrate_eps(pts, target)
            off, nb = O.ref_superego_pairs(pts, eps)
            np.savez_compressed(os.path.join(HERE, f"ref_alg3_d{dim}_n{n}.npz"), pts=pts, eps=eps,
                                offsets=off, nbrs=nb)
            print(f"ref_alg3_d{dim}_n{n}", "pairs", len(nb))
    if O.ref_available():
        with tempfile.TemporaryDirectory() as td:
            p = O.ref_generate_text(2, 2000, td)
            lines = open(p).read().splitlines(True)[:200]
        with open(os.path.join(HERE, "ref_gen_d2_head200.txt"), "w") as f:
            f.writelines(lines)


if __name__ == "__main__":
    main()
