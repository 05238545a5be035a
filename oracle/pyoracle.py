"""ctypes wrapper for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() for
the expected result and by bench.py's cpu_baseline leg.  The product never
imports it.

Also drives the reference's own CPU code built from source into oracle/_ref
(ref_harness_d<D>, gen_dataset) when those binaries exist.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_lib = None


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, u64, f64, i32 = C.c_uint32, C.c_uint64, C.c_double, C.c_int
        L.og_generate_expo.argtypes = [u32, u32, u32, f64, _f64p]
        L.og_write_text.argtypes = [C.c_char_p, _f64p, u32, u32, i32]
        L.og_load_text.argtypes = [C.c_char_p, u32, C.POINTER(C.POINTER(C.c_double))]
        L.og_load_text.restype = C.c_int64
        L.og_free.argtypes = [C.c_void_p]
        L.og_grid_dims.argtypes = [_f64p, u32, u32, u32, f64, _f64p, _f64p, _u32p, C.POINTER(u64)]
        L.og_point_cells.argtypes = [_f64p, u32, u32, u32, f64, _f64p, _u32p, _u64p]
        L.og_index_build.argtypes = [_f64p, u32, u32, u32, f64, _f64p, _u32p, _u32p,
                                     C.c_void_p, C.c_void_p]
        L.og_index_build.restype = u32
        L.og_threshold.argtypes = [f64]
        L.og_threshold.restype = f64
        L.og_chain.argtypes = [_f64p, _f64p, u32, u32, _f64p]
        L.og_chain.restype = None
        L.og_threshold_f.argtypes = [C.c_float]
        L.og_threshold_f.restype = C.c_float
        L.og_workload.argtypes = [_f64p, u32, u32, u32, f64, _f64p, _u32p, _u32p, C.c_void_p,
                                  C.c_void_p, u32, i32, _u32p, _u32p, _u32p]
        L.og_workload.restype = u64
        L.og_join_counts.argtypes = [_f64p, u32, u32, u32, f64, i32, i32, _u32p]
        L.og_join_counts.restype = u64
        L.og_join_fill.argtypes = [_f64p, u32, u32, u32, f64, i32, i32, _u64p, _u32p]
        L.og_join_stats.argtypes = [_f64p, u32, u32, u32, f64, i32, i32, C.POINTER(u64),
                                    C.c_void_p, C.c_void_p]
        L.og_join_stats.restype = u64
        L.og_join_rows.argtypes = [_f64p, u32, u32, u32, f64, i32, i32, _u32p, u32, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
        L.og_join_rows.restype = u64
        L.og_brute_counts.argtypes = [_f64p, u32, u32, u32, f64, i32, _u32p]
        L.og_brute_counts.restype = u64
        L.og_brute_fill.argtypes = [_f64p, u32, u32, u32, f64, i32, _u64p, _u32p]
        L.og_range_counts.argtypes = [_f64p, u32, _f64p, u32, u32, f64, _u32p]
        L.og_range_counts.restype = u64
        L.og_range_fill.argtypes = [_f64p, u32, _f64p, u32, u32, f64, _u64p, _u32p]
        L.og_count_cell0.argtypes = [_f64p, u32, u32, u32, f64]
        L.og_count_cell0.restype = u32
        L.og_pair_mix.argtypes = [u32, u32]
        L.og_pair_mix.restype = u64
        L.og_csr_hash.argtypes = [u32, _u64p, _u32p]
        L.og_csr_hash.restype = u64
        L.og_superego_join.argtypes = [_f64p, u32, u32, f64, i32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.og_superego_join.restype = u64
        L.og_superego_sample.argtypes = [_f64p, u32, u32, f64, i32, _u32p, u32, _f64p]
        L.og_superego_sample.restype = u64
        _lib = L
    return _lib


def _pts(pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    assert pts.ndim == 2
    return pts, pts.shape[0], pts.shape[1]


def generate_expo(n, dim, seed=0, lam=40.0):
    out = np.empty((n, dim), np.float64)
    lib().og_generate_expo(seed, n, dim, lam, out.reshape(-1))
    return out


def write_text(path, pts, precision=17):
    pts, n, d = _pts(pts)
    rc = lib().og_write_text(path.encode(), pts.reshape(-1), n, d, precision)
    assert rc == 0


def load_text(path, dim):
    p = C.POINTER(C.c_double)()
    n = lib().og_load_text(path.encode(), dim, C.byref(p))
    if n < 0:
        raise IOError(path)
    arr = np.ctypeslib.as_array(p, shape=(n * dim,)).copy().reshape(n, dim)
    lib().og_free(p)
    return arr


def grid_dims(pts, k, eps):
    pts, n, d = _pts(pts)
    mn = np.zeros(k); mx = np.zeros(k); nc = np.zeros(k, np.uint32); tc = C.c_uint64()
    lib().og_grid_dims(pts.reshape(-1), n, d, k, eps, mn, mx, nc, C.byref(tc))
    return mn, mx, nc, tc.value


def point_cells(pts, k, eps):
    pts, n, d = _pts(pts)
    mn, mx, nc, _ = grid_dims(pts, k, eps)
    ids = np.zeros(n, np.uint64)
    lib().og_point_cells(pts.reshape(-1), n, d, k, eps, mn, nc, ids)
    return ids


LOOKUP_DTYPE = np.dtype([("idx", np.uint32), ("pad", np.uint32), ("gridLinearID", np.uint64)])
GRID_DTYPE = np.dtype([("indexmin", np.int32), ("indexmax", np.int32)])


def index_build(pts, k, eps):
    """Returns (indexLookupArr, gridCellLookupArr, grid) in the reference layout."""
    pts, n, d = _pts(pts)
    mn, mx, nc, _ = grid_dims(pts, k, eps)
    ila = np.zeros(n, np.uint32)
    lookup = np.zeros(n, LOOKUP_DTYPE)
    grid = np.zeros(n, GRID_DTYPE)
    c = lib().og_index_build(pts.reshape(-1), n, d, k, eps, mn, nc, ila,
                             lookup.ctypes.data, grid.ctypes.data)
    return ila, lookup[:c].copy(), grid[:c].copy()


def threshold(eps):
    return lib().og_threshold(eps)


def chain(q, c):
    """The reference's FP64 distance chain for pairs (rows of q and c)."""
    q = np.ascontiguousarray(q, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    out = np.zeros(len(q))
    lib().og_chain(q, c, len(q), q.shape[1], out)
    return out


def threshold_f(eps):
    return lib().og_threshold_f(eps)


def workload(pts, k, eps, emulate_cell0=True):
    """Returns (nbPoints per cell, sortedCells, originPointIndex, totalCandidates)."""
    pts, n, d = _pts(pts)
    mn, mx, nc, _ = grid_dims(pts, k, eps)
    ila, lookup, grid = index_build(pts, k, eps)
    c = len(lookup)
    nb = np.zeros(c, np.uint32); sc = np.zeros(c, np.uint32); opi = np.zeros(n, np.uint32)
    tot = lib().og_workload(pts.reshape(-1), n, d, k, eps, mn, nc, ila, lookup.ctypes.data,
                            grid.ctypes.data, c, int(emulate_cell0), nb, sc, opi)
    return nb, sc, opi, tot


def _csr(counts_fn, fill_fn, n):
    counts = np.zeros(n, np.uint32)
    total = counts_fn(counts)
    offsets = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=offsets[1:])
    nbrs = np.zeros(max(int(total), 1), np.uint32)
    fill_fn(offsets, nbrs)
    return offsets, nbrs[:total]


def join(pts, k, eps, emulate_cell0=True, f32=False):
    """Algorithm-0 CSR (rows by original id, neighbours ascending)."""
    pts, n, d = _pts(pts)
    flat = pts.reshape(-1)
    L = lib()
    return _csr(lambda c: L.og_join_counts(flat, n, d, k, eps, int(emulate_cell0), int(f32), c),
                lambda o, b: L.og_join_fill(flat, n, d, k, eps, int(emulate_cell0), int(f32), o, b),
                n)


def join_stats(pts, k, eps, emulate_cell0=True, f32=False, rows=False):
    """Algorithm-0 (count, hash) without the pairs; with rows=True also the
    per-query counts and hashes."""
    pts, n, d = _pts(pts)
    h = C.c_uint64()
    rc = np.zeros(n, np.uint32) if rows else None
    rh = np.zeros(n, np.uint64) if rows else None
    total = lib().og_join_stats(pts.reshape(-1), n, d, k, eps, int(emulate_cell0), int(f32),
                                C.byref(h), rc.ctypes.data if rows else None,
                                rh.ctypes.data if rows else None)
    if rows:
        return int(total), h.value, rc, rh
    return int(total), h.value


def join_rows(pts, k, eps, qids, emulate_cell0=True, f32=False):
    """Algorithm-0 rows of the queries qids only: CSR over qids, rows ascending."""
    pts, n, d = _pts(pts)
    qids = np.ascontiguousarray(qids, np.uint32)
    flat = pts.reshape(-1)
    L = lib()
    counts = np.zeros(max(len(qids), 1), np.uint32)
    total = L.og_join_rows(flat, n, d, k, eps, int(emulate_cell0), int(f32), qids, len(qids),
                           counts.ctypes.data, None, None)
    offsets = np.zeros(len(qids) + 1, np.uint64)
    np.cumsum(counts[:len(qids)], out=offsets[1:])
    nbrs = np.zeros(max(int(total), 1), np.uint32)
    L.og_join_rows(flat, n, d, k, eps, int(emulate_cell0), int(f32), qids, len(qids), None,
                   offsets.ctypes.data, nbrs.ctypes.data)
    return offsets, nbrs[:total]


def ref_superego_rowstats(pts, eps):
    """Reference Util::egoJoinV2 per query: (counts[n], per-row hashes[n])."""
    pts, n, d = _pts(pts)
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "p.bin")
        out = os.path.join(td, "o.bin")
        pts.tofile(pin)
        subprocess.run([os.path.join(REF_DIR, f"ref_harness_d{d}"), "rowstats", pin, repr(eps),
                        out], check=True, capture_output=True)
        raw = np.fromfile(out, np.uint8)
    return raw[:4 * n].view(np.uint32).copy(), raw[4 * n:].view(np.uint64).copy()


def ref_superego_rows(pts, eps, qids):
    """Reference Util::egoJoinV2 rows of the queries qids: CSR over qids."""
    pts, n, d = _pts(pts)
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "p.bin")
        qin = os.path.join(td, "q.bin")
        out = os.path.join(td, "o.bin")
        pts.tofile(pin)
        np.ascontiguousarray(qids, np.uint32).tofile(qin)
        subprocess.run([os.path.join(REF_DIR, f"ref_harness_d{d}"), "rows", pin, repr(eps), qin,
                        out], check=True, capture_output=True)
        raw = np.fromfile(out, np.uint32)
    nq = len(qids)
    offsets = np.zeros(nq + 1, np.uint64)
    np.cumsum(raw[:nq], out=offsets[1:])
    return offsets, raw[nq:].copy()


def brute(pts, k, eps, emulate_cell0=True):
    pts, n, d = _pts(pts)
    flat = pts.reshape(-1)
    L = lib()
    return _csr(lambda c: L.og_brute_counts(flat, n, d, k, eps, int(emulate_cell0), c),
                lambda o, b: L.og_brute_fill(flat, n, d, k, eps, int(emulate_cell0), o, b), n)


def range_join(pts, queries, eps):
    """Brute-force range join of `queries` against `pts` (serving oracle):
    CSR over queries, neighbour ids ascending."""
    pts, n, d = _pts(pts)
    qs, nq, dq = _pts(queries)
    assert dq == d
    fp, fq = pts.reshape(-1), qs.reshape(-1)
    L = lib()
    return _csr(lambda c: L.og_range_counts(fp, n, fq, nq, d, eps, c),
                lambda o, b: L.og_range_fill(fp, n, fq, nq, d, eps, o, b), nq)


def count_cell0(pts, k, eps):
    pts, n, d = _pts(pts)
    return lib().og_count_cell0(pts.reshape(-1), n, d, k, eps)


def pair_mix(key, value):
    return lib().og_pair_mix(key, value)


def csr_hash(offsets, nbrs):
    offsets = np.ascontiguousarray(offsets, np.uint64)
    nbrs = np.ascontiguousarray(nbrs, np.uint32)
    if nbrs.size == 0:
        nbrs = np.zeros(1, np.uint32)
    return lib().og_csr_hash(len(offsets) - 1, offsets, nbrs)


def superego(pts, eps, nthreads=0):
    """Algorithm-3 CSR from the Super-EGO restatement + (reorder, sort, join) s."""
    pts, n, d = _pts(pts)
    flat = pts.reshape(-1)
    L = lib()
    counts = np.zeros(n, np.uint32)
    t = np.zeros(3)
    total = L.og_superego_join(flat, n, d, eps, nthreads, counts.ctypes.data, None, None,
                               t.ctypes.data)
    offsets = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=offsets[1:])
    nbrs = np.zeros(max(int(total), 1), np.uint32)
    L.og_superego_join(flat, n, d, eps, nthreads, None, offsets.ctypes.data, nbrs.ctypes.data,
                       None)
    return offsets, nbrs[:total], t


def superego_sample(pts, eps, qids, nthreads=0):
    pts, n, d = _pts(pts)
    qids = np.ascontiguousarray(qids, np.uint32)
    t = np.zeros(3)
    tot = lib().og_superego_sample(pts.reshape(-1), n, d, eps, nthreads, qids, len(qids), t)
    return tot, t


# ---------------------------------------------------------------------------
# The reference's own CPU code, built from source into oracle/_ref.
# ---------------------------------------------------------------------------

def ref_available(dim=None):
    names = ["gen_dataset"] if dim is None else [f"ref_harness_d{dim}"]
    return all(os.path.exists(os.path.join(REF_DIR, x)) for x in names)


def ref_generate_text(dim, n, workdir):
    """Runs the reference generator; returns the path of the text it wrote."""
    subprocess.run([os.path.join(REF_DIR, "gen_dataset"), str(dim), str(n)], cwd=workdir,
                   check=True, capture_output=True)
    return os.path.join(workdir, f"dataset_fixed_len_pts_expo_NDIM_{dim}_pts_{n}.txt")


def ref_load_text(path, dim):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "v.bin")
        subprocess.run([os.path.join(REF_DIR, f"ref_harness_d{dim}"), "load", path, out],
                       check=True, capture_output=True)
        return np.fromfile(out, np.float64).reshape(-1, dim)


def ref_superego_pairs(pts, eps):
    """Reference Util::egoJoinV2 per query -> CSR (rows sorted ascending)."""
    pts, n, d = _pts(pts)
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "p.bin")
        out = os.path.join(td, "o.bin")
        pts.tofile(pin)
        subprocess.run([os.path.join(REF_DIR, f"ref_harness_d{d}"), "pairs", pin, repr(eps), out],
                       check=True, capture_output=True)
        raw = np.fromfile(out, np.uint32)
    counts = raw[:n]
    offsets = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=offsets[1:])
    return offsets, raw[n:].copy()


def ref_superego_time(pts, eps, order, nq, workdir=None, threads=None):
    """Times the reference's multiThreadJoinWorkQueue over order[:nq] on
    `threads` threads (its CPU_THREADS; default: the harness's own choice,
    OMP_NUM_THREADS or the process's CPUs).  Returns the parsed REF_TIME dict."""
    import json
    pts, n, d = _pts(pts)
    env = dict(os.environ)
    if threads:
        env["HEGJ_REF_THREADS"] = str(int(threads))
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        pin = os.path.join(td, "p.bin")
        oin = os.path.join(td, "order.bin")
        pts.tofile(pin)
        np.ascontiguousarray(order, np.uint32).tofile(oin)
        r = subprocess.run([os.path.join(REF_DIR, f"ref_harness_d{d}"), "time", pin, repr(eps),
                            oin, str(nq)], check=True, capture_output=True, text=True, env=env)
    for line in r.stdout.splitlines():
        if line.startswith("REF_TIME "):
            return json.loads(line[len("REF_TIME "):])
    raise RuntimeError("no REF_TIME line")
