"""Pins the CPU oracle (CPU only, no GPU).

The reference ships no tests or golden outputs, so every oracle function is
pinned here by something independent of its own code:
  * the reference's own C++ compiled from source into oracle/_ref
    (generator, loader, Super-EGO) -- exact equality;
  * O(n^2) brute force with the literal sqrt predicate and no grid;
  * closed forms on lattices whose distances are exact in FP64;
  * the mathematical definition of the sqrt threshold;
  * invariants (symmetry, self pairs, layout of the grid index).
Planted-mistake tests mutate the oracle's inputs/outputs on the CPU and check
that the pins catch the change.
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import golden_files, load_golden, GOLDEN


def rows(off, nb):
    return [nb[off[i]:off[i + 1]] for i in range(len(off) - 1)]


# --------------------------------------------------------------------------
# data I/O vs the reference's own code
# --------------------------------------------------------------------------
def test_generator_matches_reference_text(oracle):
    """og_generate_expo + 6-digit text == datasets/dataset_gen...cpp output."""
    ref = open(os.path.join(GOLDEN, "ref_gen_d2_head200.txt")).read()
    pts = oracle.generate_expo(2000, 2, seed=2137834274)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.txt")
        oracle.write_text(p, pts[:200], 6)
        assert open(p).read() == ref


@pytest.mark.skipif(not os.path.exists("/root/reference"), reason="reference sources absent")
def test_generator_full_file_matches_reference(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    with tempfile.TemporaryDirectory() as td:
        p = oracle.ref_generate_text(3, 5000, td)
        pts = oracle.generate_expo(5000, 3, seed=2137834274)
        q = os.path.join(td, "m.txt")
        oracle.write_text(q, pts, 6)
        assert open(p).read() == open(q).read()


def test_loader_matches_reference_importNDDataset(oracle):
    if not oracle.ref_available(2):
        pytest.skip("oracle/_ref not built")
    pts = oracle.generate_expo(3000, 2, seed=4)
    with tempfile.TemporaryDirectory() as td:
        for prec in (6, 17):
            p = os.path.join(td, f"d{prec}.txt")
            oracle.write_text(p, pts, prec)
            a = oracle.load_text(p, 2)
            b = oracle.ref_load_text(p, 2)
            assert a.shape == b.shape == (3000, 2)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
            if prec == 17:  # 17 significant digits round-trip exactly
                assert np.array_equal(a.view(np.uint64), pts.view(np.uint64))


def test_loader_ragged_and_separators(oracle):
    """Values separated by commas and/or whitespace; a trailing partial point
    is dropped (count = values / dim, import_dataset.cpp:28)."""
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "r.txt")
        open(p, "w").write("0.1, 0.2\n0.3,0.4\n 0.5 ,\t0.6\n0.7\n")
        a = oracle.load_text(p, 2)
        assert a.tolist() == [[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]]
        if oracle.ref_available(2):
            b = oracle.ref_load_text(p, 2)
            assert np.array_equal(a, b)


# --------------------------------------------------------------------------
# S1-S3, S11: grid geometry and index layout
# --------------------------------------------------------------------------
def test_grid_dims_closed_form(oracle):
    pts = np.array([[0.5, 0.25], [0.75, 0.125], [0.625, 0.5]])
    mn, mx, nc, tc = oracle.grid_dims(pts, 2, 0.125)
    assert mn.tolist() == [0.375, 0.0]  # min - eps, exact in binary
    assert mx.tolist() == [0.875, 0.625]
    assert nc.tolist() == [4, 5]        # ceil((max - min) / eps)
    assert tc == 20


def test_cell_ids_floor_definition(oracle):
    """(x - min)/eps truncated, dim 0 fastest: checked on a dyadic lattice
    where every operation is exact, so floor() is unambiguous."""
    h = 1.0 / 64
    g = np.array([[i * h, j * h] for i in range(1, 20) for j in range(1, 15)])
    eps = 2 * h
    ids = oracle.point_cells(g, 2, eps)
    mn, mx, nc, _ = oracle.grid_dims(g, 2, eps)
    c = np.floor((g - mn) / eps).astype(np.uint64)
    assert np.array_equal(ids, c[:, 0] + c[:, 1] * np.uint64(nc[0]))


@pytest.mark.parametrize("path", golden_files("alg0_"))
def test_index_layout_invariants(oracle, path):
    g = load_golden(path)
    ila, lookup, grid = oracle.index_build(g["pts"], int(g["k"]), float(g["eps"]))
    ids = oracle.point_cells(g["pts"], int(g["k"]), float(g["eps"]))
    # indexLookupArr: stable sort by (cell id, point id)
    assert np.array_equal(ila, np.lexsort((np.arange(len(ids)), ids)).astype(np.uint32))
    # gridCellLookupArr: unique ids ascending with idx = i
    assert np.array_equal(lookup["gridLinearID"], np.unique(ids))
    assert np.array_equal(lookup["idx"], np.arange(len(lookup)))
    # grid: inclusive runs covering all points in order
    assert grid["indexmin"][0] == 0 and grid["indexmax"][-1] == len(ids) - 1
    assert np.array_equal(grid["indexmin"][1:], grid["indexmax"][:-1] + 1)
    assert np.all(ids[ila[grid["indexmin"]]] == lookup["gridLinearID"])
    # the golden file agrees
    assert np.array_equal(ila, g["indexLookupArr"])


# --------------------------------------------------------------------------
# S7: the sqrt threshold
# --------------------------------------------------------------------------
def test_threshold_definition(oracle):
    rng = np.random.default_rng(0)
    eps_list = list(rng.random(2000) * 0.5 + 1e-6) + [0.001, 0.02, 0.1, 0.5, 1.0, 2.0 ** -10]
    bumped = 0
    for e in eps_list:
        t = oracle.threshold(e)
        assert np.sqrt(t) <= e
        assert np.sqrt(np.nextafter(t, np.inf)) > e
        bumped += t != e * e
    assert bumped > 0  # T is not always fl(eps^2): the rule matters


def test_threshold_f32_definition(oracle):
    rng = np.random.default_rng(1)
    for e in rng.random(500).astype(np.float32):
        e = np.float32(e) + np.float32(1e-6)
        t = np.float32(oracle.threshold_f(float(e)))
        assert np.sqrt(t) <= e and np.sqrt(np.nextafter(t, np.float32(np.inf))) > e


# --------------------------------------------------------------------------
# S4-S10: the join
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,k,n,eps,seed", [
    (2, 2, 2500, 0.004, 1), (3, 1, 1500, 0.01, 2), (6, 6, 2000, 0.03, 3),
    (6, 3, 2000, 0.03, 4), (18, 6, 1200, 0.12, 5), (90, 6, 500, 0.45, 6)])
def test_join_equals_brute_force(oracle, dim, k, n, eps, seed):
    pts = oracle.generate_expo(n, dim, seed=seed)
    o1, n1 = oracle.join(pts, k, eps)
    o2, n2 = oracle.brute(pts, k, eps)
    assert np.array_equal(o1, o2) and np.array_equal(n1, n2)
    assert len(n1) > n  # more than the self pairs


@pytest.mark.parametrize("path", golden_files("cell0_"))
def test_cell0_wrap(oracle, path):
    """S5: a query whose cell has coordinate 0 emits nothing (not even its
    self pair) but is still found by other queries."""
    g = load_golden(path)
    pts, k, eps = g["pts"], int(g["k"]), float(g["eps"])
    assert int(g["n_cell0"]) >= 1
    off, nb = oracle.join(pts, k, eps)
    boff, bnb = oracle.brute(pts, k, eps)
    assert np.array_equal(off, boff) and np.array_equal(nb, bnb)
    cnt = np.diff(off)
    dead = np.where(cnt == 0)[0]
    assert len(dead) == int(g["n_cell0"])
    for q in dead:
        assert (nb == q).any()  # still a neighbour of others
    noff, nnb = oracle.join(pts, k, eps, emulate_cell0=False)
    assert np.all(np.diff(noff)[dead] >= 1)  # intended semantics keep it
    assert len(nnb) > len(nb)


def test_lattice_closed_form(oracle):
    """Integer lattice scaled by h = 1/64 (exact in FP64): with eps = h each
    interior point has 5 neighbours (self + 4), with eps just below h only
    itself; with eps = sqrt(2) h rounded up, 9.  The cell-0 wrap is off so the
    boundary is symmetric."""
    h = 1.0 / 64
    m = 12
    pts = np.array([[(i + 8) * h, (j + 8) * h] for i in range(m) for j in range(m)])
    off, nb = oracle.join(pts, 2, h, emulate_cell0=False)
    cnt = np.diff(off).reshape(m, m)
    assert np.all(cnt[1:-1, 1:-1] == 5) and cnt[0, 0] == 3 and cnt[0, 5] == 4
    off, nb = oracle.join(pts, 2, np.nextafter(h, 0), emulate_cell0=False)
    assert np.all(np.diff(off) == 1)
    e = np.sqrt(2) * h
    if np.sqrt(2 * h * h) > e:
        e = np.nextafter(e, 1)
    off, nb = oracle.join(pts, 2, e, emulate_cell0=False)
    cnt = np.diff(off).reshape(m, m)
    assert np.all(cnt[1:-1, 1:-1] == 9)


def test_planted_strict_predicate_is_caught(oracle):
    """A mutated predicate (acc < T instead of acc <= T) changes the lattice
    answer: the closed form above would catch it."""
    h = 1.0 / 64
    pts = np.array([[(i + 8) * h, (j + 8) * h] for i in range(6) for j in range(6)])
    off, nb = oracle.join(pts, 2, h, emulate_cell0=False)
    strict = 0
    for q in range(len(pts)):
        d = np.sqrt(((pts - pts[q]) ** 2).sum(1))
        strict += int((d < h).sum())
    assert strict != len(nb)


@pytest.mark.parametrize("path", golden_files("alg0_") + golden_files("cell0_"))
def test_join_symmetry_and_self_pairs(oracle, path):
    g = load_golden(path)
    off, nb = g["offsets"], g["nbrs"]
    n = len(off) - 1
    q = np.repeat(np.arange(n, dtype=np.uint64), np.diff(off).astype(np.int64))
    live = np.diff(off) > 0
    fwd = set(zip(q.tolist(), nb.tolist()))
    for a, b in list(fwd)[:20000]:
        if live[b]:
            assert (b, a) in fwd
    assert all((i, i) in fwd for i in np.where(live)[0][:5000])
    # golden matches the oracle recomputed now
    o2, n2 = oracle.join(g["pts"], int(g["k"]), float(g["eps"]))
    assert np.array_equal(o2, off) and np.array_equal(n2, nb)
    assert oracle.csr_hash(o2, n2) == int(g["hash"])


def test_float_dtype_join_equals_f32_brute(oracle):
    """DTYPE=float: coordinates, eps, distance and sqrt in float."""
    pts = oracle.generate_expo(1500, 3, seed=12)
    eps = 0.02
    off, nb = oracle.join(pts, 3, eps, f32=True)
    p = pts.astype(np.float32)
    e = np.float32(eps)
    mn = p.min(0) - e
    dead = (((p - mn) / e).astype(np.uint32) == 0).any(1)
    exp = []
    for i in range(len(p)):
        if dead[i]:
            exp.append(np.zeros(0, np.uint32))
            continue
        acc = np.zeros(len(p), np.float32)
        for l in range(3):
            d = (p[:, l] - p[i, l]).astype(np.float32)
            acc = (d.astype(np.float64) * d + acc).astype(np.float32)  # fma in f32
        exp.append(np.where(np.sqrt(acc) <= e)[0].astype(np.uint32))
    got = rows(off, nb)
    mism = sum(not np.array_equal(a, b) for a, b in zip(got, exp))
    assert mism == 0


@pytest.mark.parametrize("dim,k,n,eps,f32,emulate", [
    (2, 2, 6000, 0.004, False, True), (6, 6, 8000, 0.03, False, True), (6, 3, 5000, 0.05, True, True),
    (18, 6, 1500, 0.2, False, True), (3, 1, 4000, 0.01, False, False)])
def test_join_stats_equals_join(oracle, dim, k, n, eps, f32, emulate):
    """The streaming count + hash (stencil once per cell, fma chain stopped
    once above the threshold) equals the per-query join's CSR and its hash,
    per query and in total; join_rows gives the same rows for a subset."""
    pts = oracle.generate_expo(n, dim, seed=n + dim)
    off, nb = oracle.join(pts, k, eps, emulate_cell0=emulate, f32=f32)
    tot, h, rc, rh = oracle.join_stats(pts, k, eps, emulate_cell0=emulate, f32=f32, rows=True)
    assert tot == len(nb) and h == oracle.csr_hash(off, nb)
    assert np.array_equal(rc, np.diff(off))
    for q in (0, n // 3, n - 1):
        row = nb[off[q]:off[q + 1]]
        assert rh[q] == sum(int(oracle.pair_mix(q, int(c))) for c in row) % 2 ** 64
    qids = np.array([n - 1, 7, 0, n // 2], np.uint32)
    o2, n2 = oracle.join_rows(pts, k, eps, qids, emulate_cell0=emulate, f32=f32)
    for i, q in enumerate(qids):
        assert np.array_equal(n2[o2[i]:o2[i + 1]], nb[off[q]:off[q + 1]])


def test_join_stats_planted_mistake_is_caught(oracle):
    """A strict < in place of <= changes the stats on a lattice with exact
    distances eps (the early exit must not hide the boundary)."""
    g = np.arange(10) / 8.0
    pts = np.array(np.meshgrid(g, g, indexing="ij")).reshape(2, -1).T.copy()
    tot, _ = oracle.join_stats(pts, 2, 0.125, emulate_cell0=False)
    # every lattice point: itself + its 2..4 axis neighbours at exactly eps
    axis = sum(min(i, 1) + min(9 - i, 1) for i in range(10)) * 10 * 2
    assert tot == 100 + axis
    assert oracle.join_stats(pts, 2, np.nextafter(0.125, 0), emulate_cell0=False)[0] == 100


def test_fullsize_fixture_is_consistent():
    """tests/golden/fullsize.json: one entry per self-join workload, and at
    2-D the reference's algorithm 3 differs from algorithm 0 only by pairs
    the cell-0 wrap or the eps boundary explains."""
    import json
    with open(os.path.join(GOLDEN, "fullsize.json")) as f:
        full = json.load(f)
    assert {"2d", "6d"} <= set(full)
    for key, e in full.items():
        assert e["pairs"] >= e["n"] - e["cell0_silenced"]  # self pairs at least
        assert len(e["pair_hash"]) == 16
    a3 = full["2d"]["alg3"]
    assert a3["attribution"]["unexplained_pairs"] == 0
    diff = a3["attribution"]["only_alg0"] + a3["attribution"]["only_alg3"]
    assert diff == a3["attribution"]["cell0_pairs"] + a3["attribution"]["boundary_pairs"]
    assert (a3["pairs"] - full["2d"]["pairs"]
            == a3["attribution"]["only_alg3"] - a3["attribution"]["only_alg0"])


# --------------------------------------------------------------------------
# workload (sortByWorkLoad)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("path", golden_files("alg0_")[:3] + golden_files("cell0_"))
def test_workload_brute_force(oracle, path):
    """Per-cell candidate counts = number of points whose cell coordinates are
    within +-1 of the cell's in every indexed dim (none for a wrapped cell)."""
    g = load_golden(path)
    pts, k, eps = g["pts"], int(g["k"]), float(g["eps"])
    work, sorted_cells, opi, tot = oracle.workload(pts, k, eps)
    mn, mx, nc, _ = oracle.grid_dims(pts, k, eps)
    coords = ((pts[:, :k] - mn) / eps).astype(np.int64)
    ids = oracle.point_cells(pts, k, eps)
    keys = np.unique(ids)
    first = {int(v): i for i, v in enumerate(ids[np.lexsort((np.arange(len(ids)), ids))])}
    for ci, key in enumerate(keys[:300]):
        pc = coords[np.where(ids == key)[0][0]]
        if (pc == 0).any():
            assert work[ci] == 0
            continue
        inside = (np.abs(coords - pc) <= 1).all(1)
        assert work[ci] == inside.sum()
    sizes = np.bincount(np.searchsorted(keys, ids), minlength=len(keys))
    assert tot == int((sizes.astype(np.uint64) * work.astype(np.uint64)).sum())
    # order: workload descending, ties by cell id; opi = member lists
    w = work[sorted_cells]
    assert np.all(w[:-1] >= w[1:])
    ties = (w[:-1] == w[1:])
    assert np.all(sorted_cells[:-1][ties] < sorted_cells[1:][ties])
    assert np.array_equal(np.sort(opi), np.arange(len(pts)))
    assert np.array_equal(opi, g["originPointIndex"])


# --------------------------------------------------------------------------
# Super-EGO (algorithm 3) vs the reference's own code
# --------------------------------------------------------------------------
@pytest.mark.parametrize("path", golden_files("ref_alg3_"))
def test_superego_equals_reference_golden(oracle, path):
    """The restated Super-EGO gives exactly the pair set of the reference's
    Util::egoJoinV2 (compiled from source) stored in the fixture."""
    g = load_golden(path)
    off, nb, _ = oracle.superego(g["pts"], float(g["eps"]))
    assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])


def test_superego_live_reference(oracle):
    if not oracle.ref_available(6):
        pytest.skip("oracle/_ref not built")
    pts = oracle.generate_expo(2000, 6, seed=40)
    off, nb, _ = oracle.superego(pts, 0.03)
    o2, n2 = oracle.ref_superego_pairs(pts, 0.03)
    assert np.array_equal(off, o2) and np.array_equal(nb, n2)


@pytest.mark.parametrize("path", golden_files("ref_alg3_"))
def test_alg3_vs_alg0_differences_explained(oracle, path):
    """Algorithms 0 and 3 agree except (a) rows of queries silenced by the
    cell-0 wrap and (b) pairs whose distance lies within rounding of eps (the
    two use different summation orders and predicates)."""
    g = load_golden(path)
    pts, eps = g["pts"], float(g["eps"])
    k = min(6, pts.shape[1])
    a0 = rows(*oracle.join(pts, k, eps))
    a3 = rows(g["offsets"], g["nbrs"])
    dead = set(np.where(np.array([len(r) for r in a0]) == 0)[0].tolist())
    for q in range(len(pts)):
        if q in dead:
            continue
        diff = set(a0[q].tolist()) ^ set(a3[q].tolist())
        for c in diff:
            d = np.sqrt(((pts[c] - pts[q]) ** 2).sum())
            assert abs(d - eps) <= 1e-12, (q, c, d, eps)


# ---- serving oracle: brute-force range join --------------------------------

def test_range_join_of_the_points_themselves_is_the_selfjoin(oracle):
    """Queries = the indexed points: the range join is the brute self-join
    without the cell-0 wrap (both brute force, written independently)."""
    pts = oracle.generate_expo(1500, 3, seed=41)
    off, nb = oracle.range_join(pts, pts, 0.02)
    eoff, enb = oracle.brute(pts, 3, 0.02, emulate_cell0=False)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_range_join_lattice_closed_form(oracle):
    """Integer lattice (spacing 1/8, exact in binary) and queries on lattice
    points: with eps = 1/8 exactly the neighbours are the query's lattice
    point and its 2*dim axis neighbours inside the box."""
    g = np.arange(8) / 8.0
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    qs = pts[[0, 73, 511]]
    off, nb = oracle.range_join(pts, qs, 0.125)
    expect = []
    for q in qs:
        d = np.abs(pts - q).sum(1)
        expect.append(np.flatnonzero((d == 0) | ((d == 0.125) & ((pts != q).sum(1) == 1))))
    assert np.array_equal(np.diff(off), [len(e) for e in expect])
    assert np.array_equal(nb, np.concatenate(expect))


def test_range_join_boundary_is_sqrt_le_eps(oracle):
    """A pair at exactly distance eps (3-4-5 triangle scaled by a power of
    two) is included; the next representable larger offset is not."""
    pts = np.array([[0.0, 0.0], [0.375, 0.5]])
    q = np.array([[0.0, 0.0]])
    off, nb = oracle.range_join(pts, q, 0.625)
    assert list(nb) == [0, 1]
    pts2 = pts.copy()
    pts2[1, 1] = np.nextafter(0.5, 1.0)
    off, nb = oracle.range_join(pts2, q, 0.625)
    assert list(nb) == [0]
