"""hegj_selfjoin_noalloc (the self-join with no device allocation, its
temporaries in a fixed device array) against hegj_selfjoin on the same
index: the same rows, count and pair hash, on two calls in a row, the second
on a smaller input; a unit whose masks exceed LDS (re-evaluated in the write
pass, since this entry point has no mask scratch); a result that does not fit
the caller's buffer.  Device buffers come straight from the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hegjoin_amd as H
    if H.device_count() == 0:
        pytest.fail("no HIP device visible for a GPU test")
    return H


@pytest.fixture(scope="module")
def hip():
    rt = C.CDLL("libamdhip64.so.7")  # the runtime libhegjoin.so already loaded
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 16)) == 0
        self.nbytes = nbytes

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        if count:
            assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.p)


def reference(ix, **kw):
    t = ix.selfjoin(**kw)
    off, nb = t.csr(sort_rows=True)
    info = t.info
    t.close()
    return off, nb, info


def noalloc_csr(H, hip, ix, capacity, **kw):
    n = ix.n
    nbuf, obuf, cbuf = DevBuf(hip, 4 * capacity), DevBuf(hip, 8 * n), DevBuf(hip, 4 * n)
    try:
        info = ix.selfjoin_noalloc(nbuf.p.value, capacity, obuf.p.value, cbuf.p.value, **kw)
        rows = nbuf.get(np.uint32, int(info["total_neighbors"]))
        roff = obuf.get(np.uint64, n)
        rcnt = cbuf.get(np.uint32, n)
    finally:
        for b in (nbuf, obuf, cbuf):
            b.free()
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(rcnt, out=off[1:])
    nb = np.empty(int(off[-1]), np.uint32)
    for q in range(n):
        r = rows[int(roff[q]):int(roff[q]) + int(rcnt[q])]
        nb[off[q]:off[q + 1]] = np.sort(r)
    return off, nb, info


@pytest.mark.parametrize("kw", [{}, dict(refine="fp64", queries_per_unit=8, unit_order="morton")],
                         ids=["tuned", "fp64-qb8-morton"])
def test_noalloc_equals_selfjoin_twice(H, hip, kw):
    """Two calls in a row on one device, the second on a smaller input: each
    equals hegj_selfjoin on its own index (the scratch keeps the previous
    call's contents; whatever must start at zero is cleared per call)."""
    for n, seed in [(60000, 1), (15000, 2)]:
        print("case noalloc", n, seed, kw, flush=True)
        pts = H.generate_expo(n, 6, seed=seed)
        with H.GridIndex(pts, 0.03) as ix:
            eoff, enb, einfo = reference(ix, **kw)
            off, nb, info = noalloc_csr(H, hip, ix, int(einfo["total_neighbors"]) + 100, **kw)
        assert info["total_neighbors"] == einfo["total_neighbors"]
        assert info["pair_hash"] == einfo["pair_hash"]
        assert info["refine_used"] == einfo["refine_used"]
        assert np.array_equal(off, eoff) and np.array_equal(nb, enb)


def test_noalloc_large_units_reevaluate(H, hip, oracle):
    """A dense 2-D slab with k = 1: units of thousands of candidates, whose
    masks the no-allocation path (no per-wave mask scratch) re-evaluates in
    the write pass; against hegj_selfjoin and the oracle's count and hash."""
    rng = np.random.default_rng(4)
    pts = rng.random((30000, 2))
    pts[:, 0] = 0.5 + 0.01 * pts[:, 0]
    eps = 0.004
    with H.GridIndex(pts, eps, indexed_dims=1) as ix:
        eoff, enb, einfo = reference(ix, queries_per_unit=8)
        off, nb, info = noalloc_csr(H, hip, ix, int(einfo["total_neighbors"]), queries_per_unit=8)
    assert np.array_equal(off, eoff) and np.array_equal(nb, enb)
    assert (info["total_neighbors"], info["pair_hash"]) == oracle.join_stats(pts, 1, eps)


def test_noalloc_capacity_and_preparation_errors(H, hip):
    pts = H.generate_expo(20000, 6, seed=3)
    with H.GridIndex(pts, 0.03) as ix:
        buf = [DevBuf(hip, 4 * 1000), DevBuf(hip, 8 * ix.n), DevBuf(hip, 4 * ix.n)]
        try:
            # nothing built yet on this index: refused
            with pytest.raises(H.HegjError) as e:
                ix.selfjoin_noalloc(buf[0].p.value, 1000, buf[1].p.value, buf[2].p.value)
            assert e.value.status == 1
            _, _, einfo = reference(ix)
            # a buffer too small: refused, with the count needed
            with pytest.raises(H.HegjError) as e:
                ix.selfjoin_noalloc(buf[0].p.value, 1000, buf[1].p.value, buf[2].p.value)
            assert e.value.status == 6
            assert e.value.info["total_neighbors"] == einfo["total_neighbors"]
            with pytest.raises(H.HegjError):
                ix.selfjoin_noalloc(buf[0].p.value, 1000, buf[1].p.value, buf[2].p.value,
                                    refine="mfma")
        finally:
            for b in buf:
                b.free()
