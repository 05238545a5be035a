"""GPU parity: libhegjoin (HIP, through the C ABI) against the oracle and the
golden fixtures.  Integer outputs (cell ids, index layout, workload order,
neighbour sets, counts, hashes) must be bit-identical."""
import os

import numpy as np
import pytest

from conftest import golden_files, load_golden, ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hegjoin_amd as H
    if H.device_count() == 0:
        pytest.fail("no HIP device visible for a GPU test")
    return H


def gpu_join(H, pts, eps, k, emulate=True, **kw):
    with H.GridIndex(pts, eps, indexed_dims=k, emulate_cell0=emulate) as ix:
        t = ix.selfjoin(**kw)
        off, nb = t.csr(sort_rows=True)
        info = t.info
        counts = t.counts()
        t.close()
    assert np.array_equal(np.diff(off).astype(np.uint32), counts)
    return off, nb, info


ALL = golden_files("alg0_") + golden_files("cell0_")


@pytest.mark.parametrize("path", ALL, ids=os.path.basename)
def test_index_matches_golden(H, path):
    g = load_golden(path)
    pts, k, eps = g["pts"], int(g["k"]), float(g["eps"])
    print("case", os.path.basename(path), pts.shape, flush=True)
    with H.GridIndex(pts, eps, indexed_dims=k) as ix:
        info = ix.info
        assert info["minArr"] == g["minArr"].tolist()
        assert info["maxArr"] == g["maxArr"].tolist()
        assert info["nCells"] == g["nCells"].tolist()
        assert info["totalCells"] == int(g["totalCells"])
        assert info["threshold"] == float(g["threshold"])
        assert info["n_cell0_queries"] == int(g["n_cell0"])
        assert np.array_equal(ix.cell_ids(), g["cell_ids"])
        grid, lookup, ila = ix.copy_grid()
        assert np.array_equal(ila, g["indexLookupArr"])
        assert np.array_equal(lookup["gridLinearID"], g["cell_keys"])
        assert np.array_equal(lookup["idx"], np.arange(len(lookup)))
        assert np.array_equal(grid["indexmin"], g["grid_min"])
        assert np.array_equal(grid["indexmax"], g["grid_max"])


@pytest.mark.parametrize("path", ALL, ids=os.path.basename)
def test_workload_matches_golden(H, path):
    g = load_golden(path)
    pts, k, eps = g["pts"], int(g["k"]), float(g["eps"])
    with H.GridIndex(pts, eps, indexed_dims=k) as ix:
        cells, opi, tot = ix.sort_by_workload()
    assert tot == int(g["totalCandidates"])
    assert np.array_equal(cells["cellId"], g["sorted_cells"].astype(np.int32))
    assert np.array_equal(cells["nbPoints"], g["work"][g["sorted_cells"]].astype(np.int32))
    assert np.array_equal(opi, g["originPointIndex"])


@pytest.mark.parametrize("path", ALL, ids=os.path.basename)
def test_join_matches_golden(H, path):
    g = load_golden(path)
    pts, k, eps = g["pts"], int(g["k"]), float(g["eps"])
    off, nb, info = gpu_join(H, pts, eps, k)
    assert info["total_neighbors"] == len(g["nbrs"])
    assert np.array_equal(off, g["offsets"])
    assert np.array_equal(nb, g["nbrs"])
    assert info["pair_hash"] == int(g["hash"])
    assert info["total_candidates"] == int(g["totalCandidates"])


@pytest.mark.parametrize("path", golden_files("cell0_"), ids=os.path.basename)
def test_join_without_cell0_wrap(H, path):
    g = load_golden(path)
    off, nb, info = gpu_join(H, g["pts"], float(g["eps"]), int(g["k"]), emulate=False)
    assert np.array_equal(off, g["nowrap_offsets"]) and np.array_equal(nb, g["nowrap_nbrs"])


@pytest.mark.parametrize("qb", [4, 8, 16])
@pytest.mark.parametrize("order", ["spatial", "workload", "morton"])
@pytest.mark.parametrize("parts", [1, 8])
def test_join_variants_identical(H, qb, order, parts):
    g = load_golden(golden_files("cell0_d2")[0])
    off, nb, info = gpu_join(H, g["pts"], float(g["eps"]), int(g["k"]), queries_per_unit=qb,
                             unit_order=order, xcd_parts=parts)
    assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])
    assert info["pair_hash"] == int(g["hash"])


def test_graph_replay_matches_eager(H, oracle):
    """The captured batch graph, replayed several times with eager joins and
    synchronisations in between, gives the eager result every time."""
    pts = oracle.generate_expo(30000, 6, seed=77)
    eoff, enb = oracle.join(pts, 6, 0.012)
    with H.GridIndex(pts, 0.012, indexed_dims=6) as ix:
        for i in range(4):
            t = ix.selfjoin(use_graph=True)
            info = t.info
            off, nb = t.csr(sort_rows=True)
            t.close()
            assert info["graph_replay"] == 1
            assert np.array_equal(off, eoff) and np.array_equal(nb, enb), i
            assert info["pair_hash"] == oracle.csr_hash(eoff, enb)
            e = ix.selfjoin(queries_per_unit=8)  # eager work between replays
            assert e.info["pair_hash"] == info["pair_hash"]
            e.close()
        # other options: the plan is rebuilt and captured again
        t = ix.selfjoin(use_graph=True, queries_per_unit=8, unit_order="workload")
        off, nb = t.csr(sort_rows=True)
        assert t.info["graph_replay"] == 1
        t.close()
        assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_graph_replay_with_small_capacity_falls_back(H, oracle):
    """A fixed capacity too small for one batch: the replay defers, the call
    runs eagerly in batches; the plan's capacity grows, so the next call
    replays the graph.  The result never changes."""
    g = load_golden(golden_files("alg0_d6_")[0])
    with H.GridIndex(g["pts"], float(g["eps"]), indexed_dims=int(g["k"])) as ix:
        for i in range(3):
            t = ix.selfjoin(use_graph=True, result_capacity=5000)
            off, nb = t.csr(sort_rows=True)
            info = t.info
            t.close()
            if i == 0:
                assert info["graph_replay"] == 0 and info["n_batches"] > 1
            else:
                assert info["graph_replay"] == 1 and info["n_batches"] == 1
            assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])


@pytest.mark.parametrize("cap", [5000, 40000])
def test_small_batches_defer_and_match(H, cap):
    """A result batch far smaller than the output forces many batches and
    deferred units; the neighbour table must not change."""
    g = load_golden(golden_files("alg0_d6_")[0])
    off, nb, info = gpu_join(H, g["pts"], float(g["eps"]), int(g["k"]), result_capacity=cap)
    assert info["n_batches"] > 1 and info["n_deferred"] > 0
    assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])
    assert info["pair_hash"] == int(g["hash"])


@pytest.mark.parametrize("dim,k,n,eps,seed", [
    (1, 1, 5000, 0.0005, 1), (2, 1, 5000, 0.003, 2), (3, 3, 6000, 0.01, 3), (4, 2, 5000, 0.02, 4),
    (5, 5, 5000, 0.03, 5), (7, 6, 4000, 0.05, 6), (8, 4, 4000, 0.05, 7), (6, 6, 20000, 0.012, 8),
    (18, 3, 2000, 0.1, 9), (90, 2, 800, 0.4, 10), (33, 6, 1000, 0.2, 11), (2, 2, 50000, 0.001, 12)])
def test_join_random_vs_oracle(H, oracle, dim, k, n, eps, seed):
    """Generic-D kernels (dims without a specialised template) and
    k < dim, against the oracle computed live."""
    pts = oracle.generate_expo(n, dim, seed=seed)
    print("case dim", dim, "k", k, "n", n, flush=True)
    off, nb, info = gpu_join(H, pts, eps, k, refine="fp64")
    eoff, enb = oracle.join(pts, k, eps)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)
    assert info["pair_hash"] == oracle.csr_hash(eoff, enb)


@pytest.mark.parametrize("refine", ["f32", "fp64"])
def test_block_waves_and_batches_same_table(H, oracle, refine):
    """Workgroups of 1, 2 or 4 waves and the estimated result split into up
    to 8 device batches give the same neighbour table (6-D, 400k points,
    ~20M pairs, so 8 batches are several batches of real size)."""
    pts = oracle.generate_expo(400000, 6, seed=7)
    eps = 0.0095
    base = gpu_join(H, pts, eps, 6, refine=refine)
    assert base[2]["refine_used"] == refine and base[2]["n_batches"] == 1
    for w, b in [(1, 1), (2, 1), (4, 8), (2, 8), (1, 4)]:
        print("case", refine, "block_waves", w, "batches", b, flush=True)
        off, nb, info = gpu_join(H, pts, eps, 6, refine=refine, block_waves=w, batches=b)
        if b > 1:
            assert info["n_batches"] >= b // 2 + 1
        assert info["pair_hash"] == base[2]["pair_hash"]
        assert np.array_equal(off, base[0]) and np.array_equal(nb, base[1])


def test_block_waves_and_batches_validated(H):
    pts = np.random.default_rng(1).random((1000, 2))
    with H.GridIndex(pts, 0.01, indexed_dims=2) as ix:
        for kw in (dict(block_waves=3), dict(block_waves=8), dict(batches=33)):
            with pytest.raises(H.HegjError):
                ix.selfjoin(**kw)


HIGH = golden_files("alg0_d18") + golden_files("alg0_d90")


@pytest.mark.parametrize("refine", ["fp64", "mfma"])
@pytest.mark.parametrize("path", HIGH, ids=os.path.basename)
def test_high_dim_golden_both_refines(H, path, refine):
    """18-D and 90-D fixtures through the plain FP64 kernel and through the
    split-BF16 MFMA guard band: both must give the fixture's pair set."""
    g = load_golden(path)
    off, nb, info = gpu_join(H, g["pts"], float(g["eps"]), int(g["k"]), refine=refine)
    assert info["refine_used"] == refine
    assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])
    assert info["pair_hash"] == int(g["hash"])


@pytest.mark.parametrize("dim,k,n,eps,seed", [
    (16, 6, 3000, 0.12, 21), (18, 6, 4000, 0.15, 22), (20, 4, 3000, 0.15, 23),
    (29, 6, 2000, 0.25, 24), (40, 6, 1500, 0.3, 25), (64, 3, 1200, 0.4, 26),
    (90, 6, 1500, 0.45, 27), (100, 5, 1000, 0.5, 28), (120, 2, 800, 0.5, 29)])
def test_mfma_refine_vs_oracle(H, oracle, dim, k, n, eps, seed):
    """Every instantiated K-block count of the MFMA kernel (K = 3(dim+2)
    padded to 64 .. 384) against the FP64 oracle."""
    pts = oracle.generate_expo(n, dim, seed=seed)
    print("case mfma dim", dim, "k", k, "n", n, flush=True)
    off, nb, info = gpu_join(H, pts, eps, k, refine="mfma")
    eoff, enb = oracle.join(pts, k, eps)
    assert info["refine_used"] == "mfma"
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)
    assert info["pair_hash"] == oracle.csr_hash(eoff, enb)


def test_mfma_pair_list_rebuilt_after_overflow(H, oracle):
    """result_capacity sizes the MFMA refine's first pair list; a list far
    too small for the join overflows and the refine runs again with a list
    sized from the count it reached: same pair set as the oracle."""
    pts = oracle.generate_expo(4000, 18, seed=22)
    off, nb, info = gpu_join(H, pts, 0.15, 6, refine="mfma", result_capacity=4096)
    eoff, enb = oracle.join(pts, 6, 0.15)
    assert info["refine_used"] == "mfma" and info["total_neighbors"] > 4096
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def planted_boundary(dim, n_base, eps, seed):
    """Random points plus, for each, partners at distance eps (1 + r) with
    |r| <= 1e-9 along random directions: pairs whose FP64 distance sits on
    the eps boundary, decided only by the FP64 recheck."""
    rng = np.random.default_rng(seed)
    base = 0.3 + 0.4 * rng.random((n_base, dim))
    out = [base]
    for rel in (-1e-9, -1e-15, 0.0, 1e-15, 1e-9):
        d = rng.standard_normal((n_base, dim))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out.append(base + d * eps * (1 + rel))
    return np.ascontiguousarray(np.concatenate(out))


@pytest.mark.parametrize("dim,k", [(18, 6), (90, 6), (33, 3)])
def test_mfma_boundary_pairs_rechecked(H, oracle, dim, k):
    """Pairs on the eps boundary land in the guard band; the FP64 recheck
    must decide each exactly as the oracle does (sqrt(acc) <= eps)."""
    eps = 0.05
    pts = planted_boundary(dim, 400, eps, seed=dim)
    off, nb, info = gpu_join(H, pts, eps, k, refine="mfma")
    eoff, enb = oracle.join(pts, k, eps)
    assert info["band_rechecks"] > 0
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_mfma_large_offset_coordinates(H, oracle):
    """Coordinates far from the origin (centring keeps the BF16 split
    accurate) and duplicates (distance 0)."""
    rng = np.random.default_rng(4)
    pts = 1000.0 + rng.random((2000, 24))
    pts[1000:1100] = pts[:100]
    off, nb, info = gpu_join(H, pts, 0.6, 6, refine="mfma")
    eoff, enb = oracle.join(pts, 6, 0.6)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


LOWD = [p for p in golden_files("alg0_") + golden_files("cell0_")
        if os.path.basename(p).startswith(("alg0_d2", "alg0_d6", "cell0_d2", "cell0_d6"))]


@pytest.mark.parametrize("path", LOWD, ids=os.path.basename)
def test_f32_filter_matches_golden(H, path):
    """The FP32 filter with its FP64 band recheck gives the fixture's exact
    neighbour table (2-D and 6-D fixtures, k = 2, 3, 6, cell-0 wrap)."""
    g = load_golden(path)
    off, nb, info = gpu_join(H, g["pts"], float(g["eps"]), int(g["k"]), refine="f32")
    assert info["refine_used"] == "f32"
    assert np.array_equal(off, g["offsets"]) and np.array_equal(nb, g["nbrs"])
    assert info["pair_hash"] == int(g["hash"])


@pytest.mark.parametrize("dim,qb", [(2, 4), (2, 8), (2, 16), (6, 4), (6, 8)])
@pytest.mark.parametrize("order", ["morton", "spatial"])
def test_f32_filter_variants_vs_oracle(H, oracle, dim, qb, order):
    pts = oracle.generate_expo(30000, dim, seed=40 + dim)
    eps = 0.002 if dim == 2 else 0.012
    off, nb, info = gpu_join(H, pts, eps, dim, refine="f32", queries_per_unit=qb, unit_order=order)
    eoff, enb = oracle.join(pts, dim, eps)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


@pytest.mark.parametrize("dim", [2, 6])
def test_f32_filter_boundary_pairs_rechecked(H, oracle, dim):
    """Pairs planted at eps(1 +- 1e-9, 1e-15, 0) fall in the FP32 band and
    are decided by the FP64 recheck exactly as the oracle does."""
    eps = 0.01
    pts = planted_boundary(dim, 2000, eps, seed=60 + dim)
    off, nb, info = gpu_join(H, pts, eps, dim, refine="f32")
    eoff, enb = oracle.join(pts, dim, eps)
    assert info["band_rechecks"] > 0
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_f32_filter_declines_wide_bands(H, oracle):
    """Coordinates far from their centre relative to eps: the FP32 band
    would exceed 1% of eps^2, so refine="f32" is refused and "auto" runs the
    FP64 chain."""
    rng = np.random.default_rng(8)
    pts = rng.random((3000, 6)) * 1e3
    pts[::2] = pts[1::2] + 1e-4
    eps = 1e-3
    with pytest.raises(H.HegjError) as e:
        gpu_join(H, pts, eps, 2, refine="f32")
    assert e.value.status == 2
    off, nb, info = gpu_join(H, pts, eps, 2, refine="auto")
    assert info["refine_used"] == "fp64"
    eoff, enb = oracle.join(pts, 2, eps)
    assert len(enb) > 3000
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_grid_with_too_many_cells_is_refused(H):
    """1e6 cells per dimension in 6 dimensions: the linear cell id would need
    more than 64 bits (the reference's product wraps), so the build fails."""
    pts = np.random.default_rng(2).random((1000, 6)) * 1e3
    with pytest.raises(H.HegjError) as e:
        H.GridIndex(pts, 1e-3, indexed_dims=6)
    assert e.value.status == 2


def test_f32_filter_graph_replay(H, oracle):
    pts = oracle.generate_expo(20000, 6, seed=78)
    eoff, enb = oracle.join(pts, 6, 0.012)
    with H.GridIndex(pts, 0.012, indexed_dims=6) as ix:
        for _ in range(3):
            t = ix.selfjoin(use_graph=True, refine="f32")
            off, nb = t.csr(sort_rows=True)
            info = t.info
            t.close()
            assert info["graph_replay"] == 1 and info["refine_used"] == "f32"
            assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


@pytest.mark.parametrize("block_waves", [4, 1])
@pytest.mark.parametrize("n,eps", [(40000, 0.004), (60000, 0.005)])
def test_large_units_use_global_masks(H, oracle, n, eps, block_waves):
    """Slabs (k = 1) holding thousands of points: with 16 queries per unit
    the (query, chunk) masks exceed LDS (first case: per-wave global
    scratch, one slice per wave of the launch at any workgroup size) or even
    that scratch (second case, more than 8192 masks: recomputed in the write
    phase)."""
    rng = np.random.default_rng(3)
    pts = rng.random((n, 3))
    pts[:, 0] *= 0.02
    off, nb, info = gpu_join(H, pts, eps, 1, queries_per_unit=16, block_waves=block_waves)
    eoff, enb = oracle.join(pts, 1, eps)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def dense_slab(dim, n, eps, width, seed):
    """A slab three cells wide along dim 0 (k = 1: every point of the slab is
    a candidate of every query in it), the other dims uniform over `width`,
    plus partners of 300 points at distance eps (1 + r), r in {+-1e-9,
    +-1e-15, 0}, which the FP32 filter can only settle by the FP64 recheck."""
    rng = np.random.default_rng(seed)
    pts = np.empty((n, dim))
    pts[:, 0] = 0.5 + 2.5 * eps * rng.random(n)
    pts[:, 1:] = 0.3 + width * rng.random((n, dim - 1))
    d = rng.standard_normal((300, dim))
    d[:, 0] = 0.0  # partners stay in the slab's cells
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rel = rng.choice([-1e-9, -1e-15, 0.0, 1e-15, 1e-9], 300)
    part = pts[:300] + d * (eps * (1 + rel))[:, None]
    return np.ascontiguousarray(np.concatenate([pts, part]))


# (dim, n, eps, width, queries per unit, expected mask store): with M slab
# candidates a unit keeps qb * ceil(M / 64) hit masks; above 512 they go to
# the per-wave global scratch, above 8192 the write pass re-evaluates the
# distances instead
_DENSE_CACHE = {}
DENSE = [(2, 40000, 0.004, 1.0, 8, "global"), (2, 40000, 0.004, 1.0, 16, "reevaluate"),
         (6, 70000, 0.01, 0.05, 4, "global"), (6, 70000, 0.01, 0.05, 8, "reevaluate")]


@pytest.mark.parametrize("refine", ["f32", "fp64"])
@pytest.mark.parametrize("dim,n,eps,width,qb,store", DENSE,
                         ids=[f"d{c[0]}-qb{c[4]}-{c[5]}" for c in DENSE])
def test_dense_units_mask_stores(H, oracle, dim, n, eps, width, qb, store, refine):
    """The register kernels (D = 2, 6) with the FP32 filter and the FP64
    chain on units whose masks exceed LDS: the global-scratch path (with FP32
    band rechecks OR-ed into global masks) and the re-evaluating write pass
    (its rows must hold exactly the counted pairs), both against the oracle."""
    print("case", dim, n, qb, store, refine, flush=True)
    if dim not in _DENSE_CACHE:
        pts = dense_slab(dim, n, eps, width, seed=dim * 100)
        # count, pair hash and per-query counts (the pair set has millions of
        # pairs; the hash covers every one of them)
        _DENSE_CACHE[dim] = (pts, oracle.join_stats(pts, 1, eps, rows=True))
    pts, (total, phash, rcnt, _) = _DENSE_CACHE[dim]
    with H.GridIndex(pts, eps, indexed_dims=1) as ix:
        t = ix.selfjoin(refine=refine, queries_per_unit=qb, prune=False)
        info, counts = t.info, t.counts()
        t.close()
    assert info["refine_used"] == refine
    if refine == "f32":
        assert info["band_rechecks"] > 0
    assert info["total_neighbors"] == total and info["pair_hash"] == phash
    assert np.array_equal(counts, rcnt)


def test_duplicate_points_and_tiny_inputs(H, oracle):
    """Duplicates (distance 0), a single point, and all points in one cell."""
    for pts, eps in [(np.full((50, 3), 0.5), 0.1), (np.array([[0.3, 0.7]]), 0.2),
                     (np.random.default_rng(0).random((300, 2)) * 0.01 + 0.5, 0.5)]:
        off, nb, _ = gpu_join(H, pts, eps, pts.shape[1])
        eoff, enb = oracle.join(pts, pts.shape[1], eps)
        assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_cli_end_to_end(H, oracle, tmp_path):
    import subprocess
    cli = os.path.join(ROOT, "hegjoin_amd", "main")
    pts = oracle.generate_expo(20000, 2, seed=5)
    path = str(tmp_path / "d.txt")
    oracle.write_text(path, pts, 17)
    dump = str(tmp_path / "csr.bin")
    r = subprocess.run([cli, path, "0.002", "2", "0", f"--dump-csr={dump}"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    eoff, enb = oracle.join(pts, 2, 0.002)
    line = [l for l in r.stdout.splitlines() if l.startswith("[RESULT] ~ Total result set size:")]
    assert int(line[0].split(":")[1]) == len(enb)
    raw = open(dump, "rb").read()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    off = np.frombuffer(raw[4:4 + 8 * (n + 1)], np.uint64)
    nb = np.frombuffer(raw[4 + 8 * (n + 1):], np.uint32)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def lattice_boundary(dim, eps, seed, offset=0.0, cmax=30):
    """Points on the cell lattice and a hair either side of it (cell
    boundaries at min - eps + c eps), and pairs at eps (1 +- 1e-15, 0)
    straddling cells along one or several dims."""
    rng = np.random.default_rng(seed)
    c = rng.integers(2, cmax, size=(1500, dim)).astype(np.float64)
    lat = offset + eps * c
    jit = lat * (1 + rng.choice([-1e-16, 0.0, 1e-16, -2e-16, 2e-16], size=lat.shape))
    d = np.zeros((1500, dim))
    axes = rng.integers(0, dim, size=(1500, 2))
    d[np.arange(1500), axes[:, 0]] = 1.0
    d[np.arange(1500), axes[:, 1]] += 1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rel = rng.choice([-1e-15, 0.0, 1e-15], 1500)
    part = jit + d * (eps * (1 + rel))[:, None]
    return np.ascontiguousarray(np.concatenate([lat, jit, part]))


@pytest.mark.parametrize("dim,k,offset,cmax", [
    (2, 2, 0.0, 30), (3, 3, 0.0, 30), (6, 6, 0.0, 30), (6, 4, 0.0, 30), (8, 6, 0.0, 30),
    (6, 6, 3.0e5, 30), (2, 2, 1.0e7, 30), (2, 2, 0.0, 10 ** 7), (3, 3, 0.0, 10 ** 6)])
def test_pruning_is_exact(H, oracle, dim, k, offset, cmax):
    """Cell-box pruning leaves out stencil cells farther than eps from every
    query of a unit; on points on and next to cell boundaries (also far from
    the origin, and at cell coordinates up to 1e7) the pair set is the
    oracle's, with and without pruning, for the FP64 and FP32 paths."""
    eps = 0.01
    pts = lattice_boundary(dim, eps, seed=70 + dim + k, offset=offset, cmax=cmax)
    eoff, enb = oracle.join(pts, k, eps)
    refines = ["fp64", "f32"] if (dim in (2, 6) and offset == 0.0 and cmax == 30) else ["fp64"]
    for refine in refines:
        for prune in (True, False):
            for qb in (4, 8):
                print("case dim", dim, "k", k, "offset", offset, refine, "prune", prune, "qb", qb, flush=True)
                off, nb, info = gpu_join(H, pts, eps, k, refine=refine, prune=prune, queries_per_unit=qb)
                assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_pruning_reduces_evaluations(H, oracle):
    """On uniform data with a few points per cell, pruning evaluates fewer
    candidates than the full stencil and finds the same pairs."""
    pts = oracle.generate_expo(40000, 6, seed=77)
    eps = 0.02
    a = gpu_join(H, pts, eps, 6, refine="fp64", queries_per_unit=4, prune=True)
    b = gpu_join(H, pts, eps, 6, refine="fp64", queries_per_unit=4, prune=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2]["evals_executed"] < b[2]["evals_executed"]
    assert b[2]["evals_executed"] == b[2]["total_candidates"]


@pytest.mark.parametrize("qb", [4, 8])
def test_f32_band_list_overflow(H, oracle, qb):
    """A cluster of near-identical points with partners at eps (1 +- 1e-15):
    nearly every (query, chunk) pair of a unit has lanes in the FP32 band,
    more than the per-wave list holds, so every pair is rechecked in FP64."""
    rng = np.random.default_rng(81)
    eps = 0.01
    base = 0.5 + 1e-13 * rng.random((1200, 6))
    d = rng.standard_normal((1200, 6))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rel = rng.choice([-1e-15, 0.0, 1e-15], 1200)
    pts = np.ascontiguousarray(np.concatenate([base, base + d * (eps * (1 + rel))[:, None]]))
    off, nb, info = gpu_join(H, pts, eps, 6, refine="f32", queries_per_unit=qb)
    eoff, enb = oracle.join(pts, 6, eps)
    assert info["refine_used"] == "f32" and info["band_rechecks"] > 64 * 64
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_cached_buffers_released_between_joins(H, oracle):
    """Joins reuse the buffers of finished calls; releasing them between
    calls changes nothing in the results."""
    pts = oracle.generate_expo(20000, 6, seed=91)
    eoff, enb = oracle.join(pts, 6, 0.015)
    for release in (False, True, False):
        if release:
            H.release_cached_memory()
        off, nb, _ = gpu_join(H, pts, 0.015, 6)
        assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def planted_pairs(pts, eps, seed, m=400):
    """m partners of the first m points at distance eps (1 + r), r in
    {+-3e-8, +-1e-8, 0}: near the threshold, where FP32 and FP64 decide
    differently, so a float join that ran in FP64 (or the reverse) fails."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((m, pts.shape[1]))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rel = rng.choice([-3e-8, -1e-8, 0.0, 1e-8, 3e-8], m)
    part = pts[:m] + d * (eps * (1 + rel))[:, None]
    return np.ascontiguousarray(np.concatenate([pts, part]))


# (dim, k, n, eps, refine, queries per unit): the FP32-filter instances with
# the float data itself (dim 2 and 6) and the generic FP32-chain kernel
# (refine fp64 = the plain kernel, dim 3 / 18, 6-D with 16 queries per unit)
FLOAT_CASES = [(2, 2, 20000, 0.01, None, None), (2, 2, 20000, 0.01, "fp64", None),
               (2, 2, 20000, 0.01, "f32", 16), (6, 6, 20000, 0.06, None, None),
               (6, 6, 20000, 0.06, "fp64", 4), (6, 6, 20000, 0.06, "auto", 16),
               (3, 3, 20000, 0.03, None, None), (18, 6, 4000, 0.25, None, None)]


@pytest.mark.parametrize("dim,k,n,eps,refine,qb", FLOAT_CASES,
                         ids=[f"d{c[0]}-{c[4] or 'auto'}-qb{c[5] or 0}" for c in FLOAT_CASES])
def test_dtype_float_vs_oracle(H, oracle, dim, k, n, eps, refine, qb):
    """DTYPE=float (params.h:24): database and epsilon rounded to float,
    minArr / nCells / cell ids and the distance chain in FP32 -- the oracle's
    Grid<float> restatement.  The grid is also checked against numpy float32
    arithmetic (IEEE, correctly rounded) and the planted near-eps pairs make
    the FP32 and FP64 pair sets differ, so the test tells the two apart."""
    print("case float", dim, k, n, eps, refine, qb, flush=True)
    pts = planted_pairs(H.generate_expo(n, dim, seed=11 + dim), eps, seed=dim)
    e64 = oracle.join_stats(pts, k, eps)
    e32 = oracle.join_stats(pts, k, eps, f32=True)
    assert e32 != e64  # the planted pairs separate the two precisions
    kw = {}
    if refine:
        kw["refine"] = refine
    if qb:
        kw["queries_per_unit"] = qb
    with H.GridIndex(pts, eps, indexed_dims=k, dtype="f32") as ix:
        info = ix.info
        ids = ix.cell_ids()
        t = ix.selfjoin(**kw)
        off, nb = t.csr(sort_rows=True)
        ri = t.info
        t.close()
    # grid in float32 arithmetic (main.cu:463-510, kernel.cu:54-57)
    p32 = pts.astype(np.float32)
    e = np.float32(eps)
    mn = p32[:, :k].min(0) - e
    mx = p32[:, :k].max(0) + e
    nc = np.ceil((mx - mn) / e).astype(np.uint32)
    assert info["dtype"] == 1 and info["epsilon"] == float(e)
    assert info["minArr"] == mn.astype(np.float64).tolist()
    assert info["nCells"] == nc.tolist()
    c = ((p32[:, :k] - mn) / e).astype(np.uint64)
    mult = np.concatenate([[1], np.cumprod(nc.astype(np.uint64))[:-1]]).astype(np.uint64)
    assert np.array_equal(ids, (c * mult).sum(1).astype(np.uint64))
    assert info["threshold"] == float(oracle.threshold_f(eps))
    eoff, enb = oracle.join(pts, k, eps, f32=True)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)
    assert (ri["total_neighbors"], ri["pair_hash"]) == e32
    # refine left open: the tuned default of this dimensionality decides
    eff = refine
    if eff is None:
        import ctypes
        o = H._JoinOpts()
        H.library().hegj_join_opts_tuned(ctypes.byref(o), dim)
        eff = {1: "fp64", 3: "f32"}.get(o.refine, "auto")
    assert ri["refine_used"] == ("f32" if (dim in (2, 6) and eff != "fp64"
                                           and not (dim == 6 and qb == 16)) else "fp64")
    assert ri["band_rechecks"] == 0


def test_dtype_float_refuses_mfma(H):
    pts = H.generate_expo(2000, 18, seed=1)
    with H.GridIndex(pts, 0.3, dtype="f32") as ix:
        with pytest.raises(H.HegjError) as e:
            ix.selfjoin(refine="mfma")
    assert e.value.status == 2
