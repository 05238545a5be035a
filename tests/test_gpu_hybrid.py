"""Hybrid modes (algorithms 1-3, hegj_selfjoin_hybrid) on the GPU.

The queue is the workload order (originPointIndex); the GPU takes it from
the head, Super-EGO CPU threads from the tail, so the GPU's queries are a
prefix of originPointIndex and the CPU's the rest.  Every query's row must
then be its algorithm-0 row (the grid oracle) when the GPU answered it and
the reference's Super-EGO row (Util::egoJoinV2 compiled from source in
oracle/_ref, or the oracle's restatement of it) when the CPU did."""
import os

import numpy as np
import pytest

from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hegjoin_amd as H
    if H.device_count() == 0:
        pytest.fail("no HIP device visible for a GPU test")
    return H


def alg3_rows(oracle, pts, eps):
    if oracle.ref_available(pts.shape[1]):
        return oracle.ref_superego_pairs(pts, eps)
    off, nb, _ = oracle.superego(pts, eps)
    return off, nb


def expected_rows(oracle, pts, k, eps, gpu_q, f32=False):
    """CSR with the algorithm-0 row for the queries in gpu_q and the
    Super-EGO row for the others (sorted rows)."""
    aoff, anb = oracle.join(pts, k, eps, f32=f32)
    if f32:  # DTYPE=float: Super-EGO on the float values, eps as a float
        soff, snb = alg3_rows(oracle, pts.astype(np.float32).astype(np.float64),
                              float(np.float32(eps)))
    else:
        soff, snb = alg3_rows(oracle, pts, eps)
    n = len(pts)
    on_gpu = np.zeros(n, bool)
    on_gpu[gpu_q] = True
    cnt = np.where(on_gpu, np.diff(aoff), np.diff(soff)).astype(np.uint64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(cnt, out=off[1:])
    nb = np.empty(int(off[-1]), np.uint32)
    for q in range(n):
        src = anb[aoff[q]:aoff[q + 1]] if on_gpu[q] else snb[soff[q]:soff[q + 1]]
        nb[off[q]:off[q + 1]] = src
    return off, nb, (aoff, anb), (soff, snb)


def run_and_check(H, oracle, pts, k, eps, mode, dtype="f64", **kw):
    print("case", mode, pts.shape, k, eps, dtype, kw, flush=True)
    with H.GridIndex(pts, eps, indexed_dims=k, dtype=dtype) as ix:
        _, opi, _ = ix.sort_by_workload()
        t = ix.selfjoin_hybrid(mode=mode, **kw)
        off, nb = t.csr(sort_rows=True)
        info, hi = t.info, t.hybrid_info
        t.close()
    n = len(pts)
    ng = hi["n_queries_gpu"]
    assert ng + hi["n_queries_cpu"] == n
    eoff, enb, (aoff, _), (soff, _) = expected_rows(oracle, pts, k, eps, opi[:ng], f32=dtype == "f32")
    assert np.array_equal(off, eoff)
    assert np.array_equal(nb, enb)
    assert info["total_neighbors"] == len(enb)
    assert info["pair_hash"] == oracle.csr_hash(eoff, enb)
    assert hi["pairs_gpu"] == int(np.diff(aoff)[opi[:ng]].sum())
    assert hi["pairs_cpu"] == int(np.diff(soff)[opi[ng:]].sum())
    return hi, info


CELL0 = load_golden(golden_files("cell0_d2")[0])


def test_cpu_mode_is_superego(H, oracle):
    """Algorithm 3: every row is Super-EGO's, including the queries the
    grid's cell-0 wrap would silence."""
    hi, _ = run_and_check(H, oracle, CELL0["pts"], int(CELL0["k"]), float(CELL0["eps"]), "cpu",
                          cpu_threads=4)
    assert hi["n_queries_gpu"] == 0 and hi["pairs_gpu"] == 0


@pytest.mark.parametrize("frac", [0.3, 0.5, 0.9])
def test_static_split(H, oracle, frac):
    """Algorithm 2: the GPU answers the first `frac` of the queue, moved up
    to a unit boundary (at most queries_per_unit - 1 more queries)."""
    pts = CELL0["pts"]
    n = len(pts)
    hi, _ = run_and_check(H, oracle, pts, int(CELL0["k"]), float(CELL0["eps"]), "static",
                          gpu_fraction=frac, cpu_threads=4, queries_per_unit=4)
    target = int(np.float32(n) * np.float32(frac))
    assert target <= hi["n_queries_gpu"] < target + 4
    assert hi["gpu_fraction"] == frac


@pytest.mark.parametrize("chunk,cap", [(0, 0), (16, 0), (16, 3000)])
def test_dynamic_split(H, oracle, chunk, cap):
    """Algorithm 1: any split point is allowed, each side's rows must be its
    own; small GPU chunks (many launches into one open batch) and a small
    result capacity (batches fill and close between chunks, units defer)."""
    pts = CELL0["pts"]
    run_and_check(H, oracle, pts, int(CELL0["k"]), float(CELL0["eps"]), "hybrid",
                  gpu_chunk=chunk, result_capacity=cap, cpu_threads=2, cpu_batch=64)


@pytest.mark.parametrize("dim,k,n,eps", [(2, 2, 40000, 0.01), (6, 6, 30000, 0.07)])
def test_dynamic_equals_gpu_only_where_algorithms_agree(H, oracle, dim, k, n, eps):
    """On data where algorithms 0 and 3 give every query the same row, the
    hybrid pair set is the GPU-only one, whatever the split."""
    pts = H.generate_expo(n, dim, seed=31)
    a = oracle.join_stats(pts, k, eps, rows=True)
    soff, snb = alg3_rows(oracle, pts, eps)
    s_rows = np.diff(soff).astype(np.uint32)
    if not np.array_equal(a[2], s_rows):
        pytest.skip("algorithms 0 and 3 differ on this data")
    with H.GridIndex(pts, eps, indexed_dims=k) as ix:
        t0 = ix.selfjoin()
        ref = (t0.info["total_neighbors"], t0.info["pair_hash"])
        t0.close()
        for mode, kw in [("hybrid", {}), ("hybrid", dict(gpu_chunk=8)),
                         ("static", dict(gpu_fraction=0.6)), ("cpu", {})]:
            t = ix.selfjoin_hybrid(mode=mode, cpu_threads=4, **kw)
            assert (t.info["total_neighbors"], t.info["pair_hash"]) == ref, mode
            t.close()
    assert ref == (a[0], a[1])


def test_float_dtype_hybrid(H, oracle):
    """DTYPE=float: the GPU's rows are the FP32 grid join's, the CPU's are
    Super-EGO's over the float values (REAL stays double, params.h:22-23)."""
    pts = H.generate_expo(8000, 3, seed=5)
    run_and_check(H, oracle, pts, 3, 0.03, "static", dtype="f32", gpu_fraction=0.5, cpu_threads=4)


def test_hybrid_rejects_bad_options(H):
    pts = H.generate_expo(1000, 2, seed=1)
    with H.GridIndex(pts, 0.02) as ix:
        with pytest.raises(H.HegjError):
            ix.selfjoin_hybrid(mode="static", gpu_fraction=1.5)
        with pytest.raises(H.HegjError):
            ix.selfjoin_hybrid(mode="hybrid", refine="mfma")
        with pytest.raises(H.HegjError):
            ix.selfjoin().hybrid_info


def test_cli_hybrid_modes(H, oracle, tmp_path):
    """`./main data eps dim 1|2|3 [partition]`: the total result size and the
    CPU/GPU lines add up; algorithm 3's total is Super-EGO's."""
    import subprocess
    cli = os.path.join(os.path.dirname(H.LIB_PATH), "main")
    pts = H.generate_expo(20000, 2, seed=8)
    p = str(tmp_path / "d.txt")
    oracle.write_text(p, pts, 17)
    eps = 0.01
    soff, _ = alg3_rows(oracle, pts, eps)
    for args in (["3"], ["2", "0.4"], ["2"], ["1"]):
        r = subprocess.run([cli, p, repr(eps), "2", *args, "--indexed-dims=2", "--cpu-threads=4"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        res = {}
        for line in r.stdout.splitlines():
            for key in ("Total result set size on the GPU", "Total result set size on the CPU",
                        "Total result set size"):
                if f"[RESULT] ~ {key}:" in line:
                    res.setdefault(key, int(line.split(":")[-1]))
        total = res["Total result set size"]
        assert total == res["Total result set size on the GPU"] + res["Total result set size on the CPU"]
        if args == ["3"]:
            assert total == int(soff[-1])
            assert "[EGO | RESULT] ~ Query points computed by Super-EGO: 20000" in r.stdout
        if args[0] == "2" and len(args) == 1:
            assert "[MODEL | RESULT] ~ Modeled GPU partition" in r.stdout
        assert ("The GPU ended before the CPU" in r.stdout) or ("The CPU ended before the GPU" in r.stdout)
