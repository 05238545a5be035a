"""The product's host Super-EGO (hegjoin_amd/csrc/hegj_ego.cpp, the CPU
share of the hybrid modes) against the reference's own Util::egoJoinV2
compiled from source (oracle/_ref) and its recorded fixtures: the row of
every query must be the same set.  No GPU needed."""
import os

import numpy as np
import pytest

from conftest import golden_files, load_golden


def sorted_rows(off, nb):
    out = nb.copy()
    for i in range(len(off) - 1):
        out[off[i]:off[i + 1]].sort()
    return out


@pytest.mark.parametrize("path", golden_files("ref_alg3_"), ids=os.path.basename)
def test_superego_matches_reference_fixtures(path):
    """Fixtures written by the reference's Util::egoJoinV2 (D = 2, 6, 18, 90)."""
    import hegjoin_amd as H
    g = load_golden(path)
    off, nb = H.superego_rows(g["pts"], float(g["eps"]))
    assert np.array_equal(off, g["offsets"])
    assert np.array_equal(sorted_rows(off, nb), g["nbrs"])


@pytest.mark.parametrize("dim,n,eps,seed", [(2, 20000, 0.01, 1), (6, 20000, 0.08, 2),
                                            (18, 4000, 0.35, 3), (90, 1500, 0.9, 4)])
def test_superego_matches_reference_build(oracle, dim, n, eps, seed):
    """Random exponential data and, at 2-D, a lattice whose pairs sit exactly
    at eps; the reference's compiled egoJoinV2 gives the expected rows."""
    import hegjoin_amd as H
    if not oracle.ref_available(dim):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    pts = oracle.generate_expo(n, dim, seed=seed)
    eoff, enb = oracle.ref_superego_pairs(pts, eps)
    for threads in (1, 8):
        off, nb = H.superego_rows(pts, eps, threads=threads)
        assert np.array_equal(off, eoff)
        assert np.array_equal(sorted_rows(off, nb), enb)


def test_superego_lattice_and_query_subset(oracle):
    """A 3-4-5 lattice with spacing eps / 5 (many pairs exactly at eps) and
    a query subset in a shuffled order: rows follow the subset."""
    import hegjoin_amd as H
    if not oracle.ref_available(2):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    eps = 0.05
    g = np.arange(40) * (eps / 5)
    pts = np.ascontiguousarray(np.stack(np.meshgrid(g, g), -1).reshape(-1, 2) + 0.1)
    eoff, enb = oracle.ref_superego_pairs(pts, eps)
    qids = np.random.default_rng(0).permutation(len(pts))[:500].astype(np.uint32)
    off, nb = H.superego_rows(pts, eps, qids=qids, threads=4)
    nb = sorted_rows(off, nb)
    for i, q in enumerate(qids):
        assert np.array_equal(nb[off[i]:off[i + 1]], enb[eoff[q]:eoff[q + 1]]), q


def test_superego_differs_from_grid_join_only_where_explained(oracle):
    """Against the algorithm-0 oracle: rows differ only for queries the cell-0
    wrap silences or by pairs within 1e-12 relative of eps."""
    import hegjoin_amd as H
    g = load_golden(golden_files("cell0_d2")[0])
    pts, eps, k = g["pts"], float(g["eps"]), int(g["k"])
    off, nb = H.superego_rows(pts, eps)
    nb = sorted_rows(off, nb)
    aoff, anb = oracle.join(pts, k, eps)
    mn, _, _, _ = oracle.grid_dims(pts, k, eps)
    cell0 = np.zeros(len(pts), bool)
    for j in range(k):
        cell0 |= ((pts[:, j] - mn[j]) / eps).astype(np.int64) == 0
    assert cell0.any()
    for q in range(len(pts)):
        a = set(anb[aoff[q]:aoff[q + 1]].tolist())
        b = set(nb[off[q]:off[q + 1]].tolist())
        if a == b:
            continue
        if cell0[q]:
            assert not a  # silenced on the grid
            continue
        for p in a ^ b:
            d = np.sqrt(((pts[p] - pts[q]) ** 2).sum())
            assert abs(d - eps) <= 1e-12 * eps, (q, p, d)


def test_superego_argument_checks():
    import hegjoin_amd as H
    with pytest.raises(H.HegjError):
        H.superego_rows(np.zeros((4, 2)), 0.1, qids=[7])
    with pytest.raises(H.HegjError):
        H.superego_rows(np.zeros((4, 2)), 0.0)


def test_static_partition_model_formula():
    """t = exp(c0 + c1 ln dim + c2 ln n + c3 ln eps) per side, share =
    gpu_qps / (gpu_qps + cpu_qps) with integer rates (main.cu:145-156): with
    t_gpu = 2 s and t_cpu = 8 s at n = 1000 the rates are 500 and 125.  The
    default coefficients are the reference's (StaticPartition.cpp:42-59)."""
    import math
    import hegjoin_amd as H
    base = H.static_partition_model(5_000_000, 6, 0.0063)
    try:
        H.set_static_partition_model([math.log(2), 0, 0, 0], [math.log(8), 0, 0, 0])
        assert H.static_partition_model(1000, 6, 0.01) == 500 / 625
        # rates truncate: n / 3 s = 333, n / 7 s = 142
        H.set_static_partition_model([math.log(3), 0, 0, 0], [math.log(7), 0, 0, 0])
        assert H.static_partition_model(1000, 2, 0.5) == 333 / (333 + 142)
        # dimension, size and epsilon enter through their logarithms
        H.set_static_partition_model([0, 1, 0, 0], [0, 0, 0.5, 0])
        # t_gpu = dim = 3 -> 333 queries/s; t_cpu = sqrt(n) = 31.6 -> 31
        assert H.static_partition_model(1000, 3, 0.5) == 333 / (333 + 31)
        H.set_static_partition_model([0, 0, 0, -1], [0, 0, 0, 0])
        # t_gpu = 1 / eps = 8 -> 125; t_cpu = 1 -> 1000
        assert H.static_partition_model(1000, 3, 0.125) == 125 / 1125
    finally:
        H.set_static_partition_model([-0.9589, 0.4342, 0.3532, 0.2686],
                                     [-7.2411, 0.9499, 0.7669, 0.4269])
    assert H.static_partition_model(5_000_000, 6, 0.0063) == base
    assert 0.0 < base < 1.0
