"""GPU tests of the serving loop (hegj_serve_*): each request's neighbour
table against the brute-force range-join oracle, the same result alone and
in a mix of concurrent requests, and admission within the result-buffer
pool.  No reference counterpart: the oracle is oracle.range_join."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDES = [0.02, 0.04, 0.08]


@pytest.fixture(scope="module")
def H():
    import hegjoin_amd as H
    if H.device_count() == 0:
        pytest.fail("no HIP device visible for a GPU test")
    return H


@pytest.fixture(scope="module")
def base(oracle):
    return oracle.generate_expo(30000, 6, seed=101)


def sorted_rows(off, nb):
    nb = nb.copy()
    for q in range(len(off) - 1):
        nb[off[q]:off[q + 1]].sort()
    return nb


REQS = [(1, 0.03, 1), (37, 0.0101, 2), (700, 0.02, 3), (2500, 0.05, 4), (1200, 0.079, 5),
        (3000, 0.0005, 6)]


def make_queries(oracle, nq, seed):
    return oracle.generate_expo(nq, 6, seed=1000 + seed)


def test_requests_match_bruteforce(H, oracle, base):
    with H.Server(base, SIDES, slots=3, result_bytes=256 << 20, max_queries=5000) as S:
        for nq, eps, seed in REQS:
            print("request nq", nq, "eps", eps, flush=True)
            q = make_queries(oracle, nq, seed)
            t = S.submit(q, eps)
            info = S.wait(t)
            off, nb = S.result(t)
            eoff, enb = oracle.range_join(base, q, eps)
            assert np.array_equal(off, eoff)
            assert np.array_equal(sorted_rows(off, nb), enb)
            assert info["pairs"] == len(enb)
            assert info["pair_hash"] == oracle.csr_hash(eoff, enb)
            assert SIDES[info["level"]] >= eps
            assert info["refine_used"] == "f32"  # 6-D, data in [0, 1]: narrow band


def test_alone_and_in_a_mix_identical(H, oracle, base):
    """Results (including row order) do not depend on what else runs."""
    qs = [(make_queries(oracle, nq, seed), eps) for nq, eps, seed in REQS]
    with H.Server(base, SIDES, slots=4, result_bytes=256 << 20, max_queries=5000) as S:
        alone = []
        for q, eps in qs:
            alone.append(S.result(S.submit(q, eps)))
        tickets = [S.submit(q, eps) for q, eps in qs * 3]
        mixed = [S.result(t) for t in tickets]
    for i, (off, nb) in enumerate(mixed):
        a_off, a_nb = alone[i % len(qs)]
        assert np.array_equal(off, a_off) and np.array_equal(nb, a_nb)


def collect_as_finished(S, tickets, timeout_s=120):
    """Release results in completion order (the serving contract: a client
    does not hold finished results while it waits for another request)."""
    import time
    out, left, t0 = {}, list(tickets), time.time()
    while left:
        assert time.time() - t0 < timeout_s, "requests did not finish"
        for t in list(left):
            info = S.poll(t)
            if info is not None:
                out[t] = S.result(t)
                left.remove(t)
        time.sleep(0.001)
    return out


def test_admission_stays_within_the_pool(H, oracle, base):
    """A pool that holds about two results: every request completes (the
    client releases results as they finish), the pool's high-water mark never
    exceeds it, and a request larger than the pool fails alone."""
    q = make_queries(oracle, 2500, 7)
    eoff, enb = oracle.range_join(base, q, 0.03)
    one = 8 * (len(q) + 1) + 4 * len(enb) + 512
    cap = int(one * 2.2)
    with H.Server(base, SIDES, slots=4, result_bytes=cap, max_queries=5000) as S:
        big = S.submit(make_queries(oracle, 3000, 8), 0.079)
        tickets = [S.submit(q, 0.03) for _ in range(8)]
        with pytest.raises(H.HegjError) as e:
            S.wait(big)
        assert e.value.status == 6  # HEGJ_ERR_CAPACITY
        S.release(big)
        for off, nb in collect_as_finished(S, tickets).values():
            assert np.array_equal(off, eoff)
            assert np.array_equal(sorted_rows(off, nb), enb)
        st = S.stats()
        assert st["pool_peak"] <= cap
        assert st["completed"] == 8 and st["failed"] == 1


def test_poll_reports_finished_requests(H, oracle, base):
    """poll: the info of a finished request (same as wait), an error for an
    unknown ticket; results collected in completion order are complete."""
    qs = [(make_queries(oracle, nq, seed), eps) for nq, eps, seed in REQS]
    with H.Server(base, SIDES, slots=4, result_bytes=256 << 20, max_queries=5000) as S:
        tickets = [S.submit(q, eps) for q, eps in qs]
        got = collect_as_finished(S, tickets)
        assert sorted(got) == sorted(tickets)
        t = S.submit(qs[0][0], qs[0][1])
        w = S.wait(t)
        p = S.poll(t)
        assert p["pairs"] == w["pairs"] and p["pair_hash"] == w["pair_hash"]
        S.release(t)
        with pytest.raises(H.HegjError):
            S.poll(t)
    for (q, eps), t in zip(qs, tickets):
        off, nb = got[t]
        eoff, enb = oracle.range_join(base, q, eps)
        assert np.array_equal(off, eoff) and np.array_equal(sorted_rows(off, nb), enb)


def test_concurrent_clients(H, oracle, base):
    """Several host threads submitting and collecting at once."""
    work = [(make_queries(oracle, 400 + 50 * i, 20 + i), 0.01 + 0.005 * i) for i in range(8)]
    expect = [oracle.range_join(base, q, e) for q, e in work]
    out = [None] * len(work)
    with H.Server(base, SIDES, slots=4, result_bytes=256 << 20, max_queries=5000) as S:
        def client(i):
            q, e = work[i]
            out[i] = S.result(S.submit(q, e))
        th = [threading.Thread(target=client, args=(i,)) for i in range(len(work))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = S.stats()
    assert st["completed"] == len(work) and st["p99_ms"] >= st["p50_ms"] > 0
    for (off, nb), (eoff, enb) in zip(out, expect):
        assert np.array_equal(off, eoff) and np.array_equal(sorted_rows(off, nb), enb)


def test_boundary_pairs_decided_by_fp64_recheck(H, oracle):
    """Queries planted at eps(1 +- 1e-9, 1e-15, 0) from index points: they
    land in the FP32 band and the FP64 recheck decides them as the oracle."""
    rng = np.random.default_rng(12)
    base = 0.3 + 0.4 * rng.random((3000, 6))
    d = rng.standard_normal((3000, 6))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    eps = 0.015
    rel = rng.choice([-1e-9, -1e-15, 0.0, 1e-15, 1e-9], 3000)
    q = base + d * (eps * (1 + rel))[:, None]
    with H.Server(base, SIDES, slots=2, result_bytes=64 << 20, max_queries=5000) as S:
        t = S.submit(q, eps)
        info = S.wait(t)
        off, nb = S.result(t)
    eoff, enb = oracle.range_join(base, q, eps)
    assert info["refine_used"] == "f32" and info["band_rechecks"] > 0
    assert np.array_equal(off, eoff) and np.array_equal(sorted_rows(off, nb), enb)


def test_queries_outside_the_grid_and_empty_requests(H, oracle, base):
    """Queries beyond the data's bounding box (far: no cell in reach; near:
    neighbours in the boundary cells), duplicates of index points, and a
    request with no queries."""
    q = np.array([[5.0] * 6, [-3.0] * 6, base[0] - 0.01, base[1], base[1],
                  np.full(6, base.max(0).max() + 0.005)])
    with H.Server(base, SIDES, slots=2, result_bytes=64 << 20, max_queries=100) as S:
        off, nb = S.result(S.submit(q, 0.03))
        eoff, enb = oracle.range_join(base, q, 0.03)
        assert np.array_equal(off, eoff) and np.array_equal(sorted_rows(off, nb), enb)
        assert off[1] == 0 and off[2] == 0
        off, nb = S.result(S.submit(np.zeros((0, 6)), 0.03))
        assert list(off) == [0] and len(nb) == 0


def test_large_groups_use_coarse_hit_bitmaps(H, oracle):
    """Requests whose query groups see far more candidate chunks than their
    hit bitmap has bits (4096 per query of the group): 1.2M 2-D points in a
    0.03 square under cells of 0.01, so one query's stencil holds ~18,750
    chunks (a bit covers 5 of them) and a group of 4 ~2 each.  The fill pass
    must write exactly the rows the count pass sized."""
    rng = np.random.default_rng(5)
    pts = rng.random((1_200_000, 2)) * 0.03
    eps = 0.0095
    with H.Server(pts, [0.01], indexed_dims=2, slots=2, result_bytes=256 << 20, max_queries=16) as S:
        for nq in (1, 4, 9):
            print("request nq", nq, flush=True)
            q = 0.0125 + rng.random((nq, 2)) * 0.005
            t = S.submit(q, eps)
            info = S.wait(t)
            off, nb = S.result(t)
            eoff, enb = oracle.range_join(pts, q, eps)
            assert np.array_equal(off, eoff)
            assert np.array_equal(sorted_rows(off, nb), enb)
            assert info["pair_hash"] == oracle.csr_hash(eoff, enb)
            assert info["evals"] >= nq * 900_000  # the stencils did hold the dense block

