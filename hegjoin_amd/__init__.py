"""hegjoin_amd -- MI355X-native HEGJoin epsilon-grid self-join (algorithm 0).

Python host mirror of the C ABI in include/hegjoin.h (libhegjoin.so, built
in-tree by hegjoin_amd/build.py).  The names follow the reference's host
entry points:

    GridIndex(database, epsilon, ...)  <- generateNDGridDimensions +
                                          gridIndexingGPU (main.cu:449-520, GPU.cu:74-398)
    GridIndex.sort_by_workload()       <- sortByWorkLoad (SortByWorkload.cu:19-247)
    GridIndex.selfjoin()               <- distanceTableNDGridBatches (GPU.cu:787-1702)
    NeighborTable.csr()/counts()       <- constructNeighborTableKeyValueWithPtrs
                                          (GPU.cu:1731-1776), as CSR by point id

The library is required: importing this package raises ImportError when
libhegjoin.so has not been built, and every call raises HegjError on a
non-zero hegj_status.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhegjoin.so")
MAX_INDEXED_DIMS = 6

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python hegjoin_amd/build.py`")
_lib = C.CDLL(LIB_PATH)

GRID_DTYPE = np.dtype([("indexmin", np.int32), ("indexmax", np.int32)])
LOOKUP_DTYPE = np.dtype([("idx", np.uint32), ("pad", np.uint32), ("gridLinearID", np.uint64)])
SCHED_DTYPE = np.dtype([("nbPoints", np.int32), ("cellId", np.int32)])

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "no device",
          4: "out of device memory", 5: "HIP error", 6: "result capacity exceeded",
          7: "I/O error", 8: "busy"}


class HegjError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"hegj status {status} ({STATUS.get(status, '?')}): {msg}")
        self.status = status


class _IndexOpts(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("indexed_dims", C.c_uint32), ("dtype", C.c_uint32),
                ("emulate_cell0", C.c_uint32), ("device", C.c_int32)]


class _IndexInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("dim", C.c_uint32), ("indexed_dims", C.c_uint32),
                ("dtype", C.c_uint32), ("nNonEmptyCells", C.c_uint32), ("n_rows", C.c_uint32),
                ("n_cell0_queries", C.c_uint32), ("emulate_cell0", C.c_uint32),
                ("totalCells", C.c_uint64), ("epsilon", C.c_double), ("threshold", C.c_double),
                ("minArr", C.c_double * MAX_INDEXED_DIMS), ("maxArr", C.c_double * MAX_INDEXED_DIMS),
                ("nCells", C.c_uint32 * MAX_INDEXED_DIMS), ("build_ms", C.c_double)]


class _JoinOpts(C.Structure):
    _fields_ = [("result_capacity", C.c_uint64), ("unit_order", C.c_uint32),
                ("queries_per_unit", C.c_uint32), ("refine", C.c_uint32),
                ("xcd_parts", C.c_uint32), ("units_per_grab", C.c_uint32),
                ("part_by_block", C.c_uint32), ("use_graph", C.c_uint32),
                ("no_prune", C.c_uint32), ("block_waves", C.c_uint32), ("batches", C.c_uint32)]


class _ResultInfo(C.Structure):
    _fields_ = [("total_neighbors", C.c_uint64), ("pair_hash", C.c_uint64),
                ("total_candidates", C.c_uint64), ("evals_executed", C.c_uint64),
                ("band_rechecks", C.c_uint64), ("n_queries", C.c_uint32),
                ("n_batches", C.c_uint32), ("n_units", C.c_uint32), ("n_deferred", C.c_uint32),
                ("estimate_ms", C.c_double), ("join_ms", C.c_double), ("workload_ms", C.c_double),
                ("total_ms", C.c_double), ("join_kernel_avg_ms", C.c_double),
                ("join_launches", C.c_uint32), ("refine_used", C.c_uint32),
                ("graph_replay", C.c_uint32), ("reserved", C.c_uint32),
                ("estimated_pairs", C.c_uint64)]


class _HybridOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("cpu_threads", C.c_uint32), ("cpu_batch", C.c_uint32),
                ("gpu_chunk", C.c_uint32), ("gpu_fraction", C.c_double)]


class _HybridInfo(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("cpu_threads", C.c_uint32), ("n_queries_gpu", C.c_uint32),
                ("n_queries_cpu", C.c_uint32), ("pairs_gpu", C.c_uint64), ("pairs_cpu", C.c_uint64),
                ("candidates_cpu", C.c_uint64), ("gpu_fraction", C.c_double), ("gpu_s", C.c_double),
                ("cpu_s", C.c_double), ("reorder_s", C.c_double), ("sort_s", C.c_double),
                ("gpu_end_s", C.c_double), ("cpu_end_s", C.c_double)]


MODES = {"gpu": 0, "hybrid": 1, "static": 2, "cpu": 3}


class _ServeOpts(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("indexed_dims", C.c_uint32), ("n_levels", C.c_uint32),
                ("slots", C.c_uint32), ("cell_sides", C.c_double * 32),
                ("result_bytes", C.c_uint64), ("max_queries", C.c_uint32), ("device", C.c_int32)]


class _RequestInfo(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("level", C.c_uint32),
                ("n_queries", C.c_uint32), ("refine_used", C.c_uint32), ("pairs", C.c_uint64),
                ("pair_hash", C.c_uint64), ("evals", C.c_uint64), ("result_bytes", C.c_uint64),
                ("band_rechecks", C.c_uint64),
                ("queue_ms", C.c_double), ("count_ms", C.c_double), ("admit_ms", C.c_double),
                ("fill_ms", C.c_double), ("latency_ms", C.c_double)]


class _ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("pairs", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_ms", C.c_double), ("pool_bytes", C.c_uint64),
                ("pool_in_use", C.c_uint64), ("pool_peak", C.c_uint64), ("wall_ms", C.c_double)]


def _struct_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if name == "reserved":
            continue
        if name == "refine_used":
            v = {1: "fp64", 2: "mfma", 3: "f32"}.get(v, "none")
        if isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


_vp = C.c_void_p
_lib.hegj_status_string.restype = C.c_char_p
_lib.hegj_last_error.restype = C.c_char_p
_lib.hegj_version.restype = C.c_char_p
_lib.hegj_dataset_generate_expo.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, _vp]
_lib.hegj_dataset_load.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.c_uint32)]
_lib.hegj_dataset_load_ex.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32,
                                      C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_uint32)]
_lib.hegj_dataset_free.argtypes = [_vp]
_lib.hegj_stream_copy_bandwidth.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double)]
_lib.hegj_release_cached_memory.argtypes = [C.c_int]
_lib.hegj_release_cached_memory.restype = None
_lib.hegj_index_opts_default.argtypes = [C.POINTER(_IndexOpts), C.c_uint32]
_lib.hegj_index_build.argtypes = [_vp, C.c_uint32, C.c_double, C.POINTER(_IndexOpts), C.POINTER(_vp)]
_lib.hegj_index_build_device.argtypes = _lib.hegj_index_build.argtypes
_lib.hegj_index_free.argtypes = [_vp]
_lib.hegj_index_get_info.argtypes = [_vp, C.POINTER(_IndexInfo)]
_lib.hegj_index_copy_grid.argtypes = [_vp, _vp, _vp, _vp]
_lib.hegj_index_copy_cell_ids.argtypes = [_vp, _vp]
_lib.hegj_sort_by_workload.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_uint64)]
_lib.hegj_join_opts_default.argtypes = [C.POINTER(_JoinOpts)]
_lib.hegj_selfjoin.argtypes = [_vp, C.POINTER(_JoinOpts), C.POINTER(_vp)]
_lib.hegj_result_free.argtypes = [_vp]
_lib.hegj_result_get_info.argtypes = [_vp, C.POINTER(_ResultInfo)]
_lib.hegj_result_copy_counts.argtypes = [_vp, _vp]
_lib.hegj_result_copy_csr.argtypes = [_vp, _vp, _vp, C.c_int]
_lib.hegj_f32_thresholds.argtypes = [C.c_double, C.c_uint32, C.c_double, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]
_lib.hegj_join_opts_tuned.argtypes = [C.POINTER(_JoinOpts), C.c_uint32]
_lib.hegj_tuning_set.argtypes = [C.c_uint32, C.POINTER(_JoinOpts)]
_lib.hegj_tuning_load.argtypes = [C.c_char_p, C.POINTER(C.c_uint32)]
_lib.hegj_selfjoin_noalloc.argtypes = [_vp, C.POINTER(_JoinOpts), _vp, C.c_uint64, _vp, _vp,
                                       C.POINTER(_ResultInfo)]
_lib.hegj_hybrid_opts_default.argtypes = [C.POINTER(_HybridOpts), C.c_uint32]
_lib.hegj_hybrid_opts_default.restype = None
_lib.hegj_selfjoin_hybrid.argtypes = [_vp, C.POINTER(_JoinOpts), C.POINTER(_HybridOpts),
                                      C.POINTER(_vp)]
_lib.hegj_result_get_hybrid_info.argtypes = [_vp, C.POINTER(_HybridInfo)]
_lib.hegj_superego_rows.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, _vp,
                                    C.c_uint32, _vp, C.POINTER(C.POINTER(C.c_uint32)),
                                    C.POINTER(C.c_uint64)]
_lib.hegj_gb_operands.argtypes = [_vp, C.c_uint32, C.c_uint32, _vp, C.c_double, _vp, _vp, _vp,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
_lib.hegj_free.argtypes = [_vp]
_lib.hegj_free.restype = None
_lib.hegj_static_partition_model.argtypes = [C.c_uint32, C.c_uint32, C.c_double]
_lib.hegj_static_partition_model.restype = C.c_double
_lib.hegj_static_partition_set_model.argtypes = [_vp, _vp]
_lib.hegj_static_partition_set_model.restype = None
_lib.hegj_serve_open.argtypes = [_vp, C.c_uint32, C.POINTER(_ServeOpts), C.POINTER(_vp)]
_lib.hegj_serve_submit.argtypes = [_vp, _vp, C.c_uint32, C.c_double, C.POINTER(C.c_uint64)]
_lib.hegj_serve_wait.argtypes = [_vp, C.c_uint64, C.POINTER(_RequestInfo)]
_lib.hegj_serve_poll.argtypes = [_vp, C.c_uint64, C.POINTER(_RequestInfo)]
_lib.hegj_serve_copy.argtypes = [_vp, C.c_uint64, _vp, _vp]
_lib.hegj_serve_release.argtypes = [_vp, C.c_uint64]
_lib.hegj_serve_get_stats.argtypes = [_vp, C.POINTER(_ServeStats), C.c_int]
_lib.hegj_serve_close.argtypes = [_vp]


def _check(status):
    if status != 0:
        raise HegjError(status, _lib.hegj_last_error().decode())


TUNING_PATH = os.path.join(_PKG, "tuning.tsv")


def load_tuning(path=TUNING_PATH):
    """Register a tuning table (dim qb order parts grab refine per line)."""
    n = C.c_uint32()
    _check(_lib.hegj_tuning_load(str(path).encode(), C.byref(n)))
    return n.value


if os.path.exists(TUNING_PATH):
    load_tuning()


def f32_thresholds(epsilon, dim, radius):
    """(tacc, trej) of the FP32 filter (see hegj_f32_thresholds)."""
    a, b = C.c_float(), C.c_float()
    _check(_lib.hegj_f32_thresholds(float(epsilon), int(dim), float(radius), C.byref(a), C.byref(b)))
    return a.value, b.value


def version() -> str:
    return _lib.hegj_version().decode()


def device_count() -> int:
    return _lib.hegj_device_count()


def stream_copy_bandwidth(nbytes=1 << 30, device=0):
    """Attainable HBM GB/s of a plain 16-byte-vector copy (read + write)."""
    g = C.c_double()
    _check(_lib.hegj_stream_copy_bandwidth(nbytes, device, C.byref(g)))
    return g.value


def release_cached_memory(device=-1):
    """Return the device buffers the self-join keeps for reuse (-1: all devices)."""
    _lib.hegj_release_cached_memory(device)


def library():
    """The loaded libhegjoin.so (ctypes.CDLL)."""
    return _lib


def generate_expo(n, dim, seed=0, lam=40.0):
    """The reference's dataset generator (datasets/dataset_gen_fixed_len_expo_dist.cpp):
    mt19937(seed) + exponential(lam), values redrawn outside [0, 1]."""
    out = np.empty((n, dim), np.float64)
    _check(_lib.hegj_dataset_generate_expo(seed, n, dim, lam, out.ctypes.data))
    return out


def load_dataset(path, dim, f32=False, bin_cache=False):
    """importNDDataset-compatible text loader (*.bin: raw row-major doubles).
    f32: parse as DTYPE=float; bin_cache: read / write the binary cache
    <path>.d<dim>[f].hegjbin beside a text file."""
    p = C.POINTER(C.c_double)()
    n = C.c_uint32()
    flags = (1 if bin_cache else 0) | (2 if f32 else 0)
    _check(_lib.hegj_dataset_load_ex(str(path).encode(), dim, flags, C.byref(p), C.byref(n)))
    try:
        return np.ctypeslib.as_array(p, shape=(n.value * dim,)).copy().reshape(n.value, dim) \
            if n.value else np.zeros((0, dim))
    finally:
        _lib.hegj_dataset_free(p)


def superego_rows(database, epsilon, qids=None, threads=0):
    """Super-EGO on the host (Util::egoJoinV2 per query, the CPU share of the
    hybrid modes): CSR (offsets, neighbours) over qids (default: every point),
    each row in EGO order."""
    arr = np.ascontiguousarray(database, dtype=np.float64)
    n, dim = arr.shape
    q = None if qids is None else np.ascontiguousarray(qids, np.uint32)
    nq = n if q is None else len(q)
    counts = np.zeros(nq, np.uint32)
    p = C.POINTER(C.c_uint32)()
    tot = C.c_uint64()
    _check(_lib.hegj_superego_rows(arr.ctypes.data, n, dim, float(epsilon), threads,
                                   None if q is None else q.ctypes.data, nq, counts.ctypes.data,
                                   C.byref(p), C.byref(tot)))
    try:
        nb = np.ctypeslib.as_array(p, shape=(tot.value,)).copy() if tot.value else \
            np.zeros(0, np.uint32)
    finally:
        _lib.hegj_free(p)
    offsets = np.zeros(nq + 1, np.uint64)
    np.cumsum(counts, out=offsets[1:])
    return offsets, nb


def gb_operands(points, centre, coef_scale=1.0):
    """The MFMA refine's operands for points (n, dim) centred at centre:
    (A', B') as uint16 BF16 bit patterns (n, kp), the guards w (n,) float32,
    kp and the guard coefficient used (host code shared with the device)."""
    pts = np.ascontiguousarray(points, np.float64)
    n, dim = pts.shape
    c = np.ascontiguousarray(centre, np.float64)
    kp = C.c_uint32()
    coef = C.c_double()
    _check(_lib.hegj_gb_operands(None, 0, dim, None, coef_scale, None, None, None, C.byref(kp),
                                 C.byref(coef)))
    A = np.zeros((n, kp.value), np.uint16)
    B = np.zeros((n, kp.value), np.uint16)
    w = np.zeros(n, np.float32)
    _check(_lib.hegj_gb_operands(pts.ctypes.data, n, dim, c.ctypes.data, coef_scale, A.ctypes.data,
                                 B.ctypes.data, w.ctypes.data, C.byref(kp), C.byref(coef)))
    return A, B, w, kp.value, coef.value


def static_partition_model(n, dim, epsilon):
    """The modelled GPU share of the queries (StaticPartition.cpp)."""
    return _lib.hegj_static_partition_model(n, dim, float(epsilon))


def set_static_partition_model(gpu, cpu):
    """Replace the (c0, c1, c2, c3) coefficients of the GPU and CPU time models."""
    g = np.ascontiguousarray(gpu, np.float64)
    c = np.ascontiguousarray(cpu, np.float64)
    assert g.shape == (4,) and c.shape == (4,)
    _lib.hegj_static_partition_set_model(g.ctypes.data, c.ctypes.data)


class GridIndex:
    """Device-resident grid index over a dataset (reference: gridIndexingGPU).

    database: (n, dim) float64 host array, or a device pointer (int) with
    n and dim given.  indexed_dims = NUMINDEXEDDIM (default min(dim, 6)).
    emulate_cell0 reproduces the reference's cell-0 wrap (default True).
    """

    def __init__(self, database, epsilon, indexed_dims=None, emulate_cell0=True, device=0,
                 n=None, dim=None, dtype="f64"):
        """dtype: "f64" (DTYPE=double) or "f32" (DTYPE=float: coordinates and
        epsilon rounded to float, grid and distances in FP32, params.h:24)."""
        self._h = _vp()
        if isinstance(database, int):
            if n is None or dim is None:
                raise ValueError("n and dim are required with a device pointer")
            ptr, build = database, _lib.hegj_index_build_device
            self._keep = None
        else:
            arr = np.ascontiguousarray(database, dtype=np.float64)
            if arr.ndim != 2:
                raise ValueError("database must be (n, dim)")
            n, dim = arr.shape
            ptr, build = arr.ctypes.data, _lib.hegj_index_build
            self._keep = arr
        o = _IndexOpts()
        _lib.hegj_index_opts_default(C.byref(o), dim)
        if indexed_dims is not None:
            o.indexed_dims = indexed_dims
        o.emulate_cell0 = 1 if emulate_cell0 else 0
        o.device = device
        o.dtype = {"f64": 0, "f32": 1}[dtype]
        _check(build(ptr, n, float(epsilon), C.byref(o), C.byref(self._h)))
        self._keep = None
        self.n, self.dim = int(n), int(dim)
        self.dtype = dtype

    def close(self):
        if self._h:
            _lib.hegj_index_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def info(self) -> dict:
        ii = _IndexInfo()
        _check(_lib.hegj_index_get_info(self._h, C.byref(ii)))
        d = _struct_dict(ii)
        k = d["indexed_dims"]
        for key in ("minArr", "maxArr", "nCells"):
            d[key] = d[key][:k]
        return d

    def copy_grid(self):
        """(index[C] grid, gridCellLookupArr[C], indexLookupArr[n]) in the
        reference's struct layouts (structs.h:83-98)."""
        c = self.info["nNonEmptyCells"]
        grid = np.zeros(c, GRID_DTYPE)
        lookup = np.zeros(c, LOOKUP_DTYPE)
        ila = np.zeros(self.n, np.uint32)
        _check(_lib.hegj_index_copy_grid(self._h, grid.ctypes.data, lookup.ctypes.data,
                                         ila.ctypes.data))
        return grid, lookup, ila

    def cell_ids(self):
        ids = np.zeros(self.n, np.uint64)
        _check(_lib.hegj_index_copy_cell_ids(self._h, ids.ctypes.data))
        return ids

    def sort_by_workload(self):
        """(sortedCells[C] {nbPoints, cellId}, originPointIndex[n], totalCandidates)."""
        c = self.info["nNonEmptyCells"]
        cells = np.zeros(c, SCHED_DTYPE)
        opi = np.zeros(self.n, np.uint32)
        tot = C.c_uint64()
        _check(_lib.hegj_sort_by_workload(self._h, cells.ctypes.data, opi.ctypes.data,
                                          C.byref(tot)))
        return cells, opi, tot.value

    def selfjoin(self, result_capacity=0, unit_order=None, queries_per_unit=None, xcd_parts=None,
                 refine=None, units_per_grab=None, part_by_block=0, use_graph=False, prune=True,
                 block_waves=None, batches=None):
        """Options left at None take the tuned defaults for this
        dimensionality (tuning.tsv, written by hegjoin_amd.autotune).
        unit_order: "morton", "spatial" (linear cell id) or "workload"
        (descending candidates, the reference's order); refine: "auto",
        "fp64" or "mfma" (split-BF16 MFMA + guard band + FP64 band
        recheck) or "f32" (FP32 distance filter + exact FP64 recheck of its
        band); xcd_parts: unit ranges, one per XCD (0 = auto);
        use_graph: keep the FP64 plan in the index and replay a captured
        hipGraph of the batch (repeated joins with the same options);
        prune=False evaluates every point of every stencil cell (the pair set
        is the same either way); block_waves: waves per workgroup of the
        grid join (1, 2 or 4); batches: device result batches the estimated
        result is split into (with result_capacity 0)."""
        o = self._join_opts(result_capacity, unit_order, queries_per_unit, xcd_parts, refine,
                            units_per_grab, part_by_block, use_graph, prune,
                            block_waves=block_waves, batches=batches)
        r = _vp()
        _check(_lib.hegj_selfjoin(self._h, C.byref(o), C.byref(r)))
        return NeighborTable(r, self)

    def selfjoin_hybrid(self, mode="hybrid", gpu_fraction=0.0, cpu_threads=0, cpu_batch=0,
                        gpu_chunk=0, **join_kw):
        """Algorithms 1-3 (hegj_selfjoin_hybrid): mode "hybrid" (the GPU
        takes units from the head of the workload-ordered queue, Super-EGO CPU
        threads batches from its tail), "static" (the GPU takes the first
        gpu_fraction of the queries; <= 0: the model) or "cpu" (Super-EGO
        alone).  join_kw as for selfjoin.  The result's hybrid_info has the
        GPU/CPU split."""
        o = self._join_opts(grid_only=True, **join_kw)
        h = _HybridOpts()
        _lib.hegj_hybrid_opts_default(C.byref(h), MODES[mode])
        h.gpu_fraction = gpu_fraction
        h.cpu_threads = cpu_threads
        h.cpu_batch = cpu_batch
        h.gpu_chunk = gpu_chunk
        r = _vp()
        _check(_lib.hegj_selfjoin_hybrid(self._h, C.byref(o), C.byref(h), C.byref(r)))
        return NeighborTable(r, self)

    def selfjoin_noalloc(self, neighbors, capacity, row_off, row_cnt, **join_kw):
        """hegj_selfjoin_noalloc: the self-join into caller-owned device
        buffers (integer device addresses) neighbors[capacity] (uint32),
        row_off[n] (uint64), row_cnt[n] (uint32), with no device allocation.
        Returns the result info; HegjError(6) when the rows do not fit."""
        o = self._join_opts(grid_only=True, **join_kw)
        ri = _ResultInfo()
        st = _lib.hegj_selfjoin_noalloc(self._h, C.byref(o), neighbors, capacity, row_off, row_cnt,
                                        C.byref(ri))
        if st != 0:
            err = HegjError(st, _lib.hegj_last_error().decode())
            err.info = _struct_dict(ri)
            raise err
        return _struct_dict(ri)

    def _join_opts(self, result_capacity=0, unit_order=None, queries_per_unit=None,
                   xcd_parts=None, refine=None, units_per_grab=None, part_by_block=0,
                   use_graph=False, prune=True, grid_only=False, block_waves=None, batches=None):
        o = _JoinOpts()
        _lib.hegj_join_opts_tuned(C.byref(o), self.dim)
        o.result_capacity = result_capacity
        if unit_order is not None:
            o.unit_order = {"spatial": 0, "workload": 1, "morton": 2}[unit_order]
        if queries_per_unit is not None:
            o.queries_per_unit = queries_per_unit
        if xcd_parts is not None:
            o.xcd_parts = xcd_parts
        if refine is not None:
            o.refine = {"auto": 0, "fp64": 1, "mfma": 2, "f32": 3}[refine]
        elif o.refine == 2 and (self.dtype == "f32" or grid_only):
            o.refine = 0  # the tuned MFMA refine is for DTYPE=double self-joins
        if units_per_grab is not None:
            o.units_per_grab = units_per_grab
        o.part_by_block = part_by_block
        o.use_graph = 1 if use_graph else 0
        o.no_prune = 0 if prune else 1
        if block_waves is not None:
            o.block_waves = block_waves
        if batches is not None:
            o.batches = batches
        return o


class NeighborTable:
    """Result of a self-join: per-query neighbour rows on the device."""

    def __init__(self, handle, index):
        self._h = handle
        self._index = index  # the result refers to the index's device arrays
        self.n = index.n

    def close(self):
        if self._h:
            _lib.hegj_result_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self) -> dict:
        ri = _ResultInfo()
        _check(_lib.hegj_result_get_info(self._h, C.byref(ri)))
        return _struct_dict(ri)

    @property
    def hybrid_info(self) -> dict:
        """GPU / CPU split of a hegj_selfjoin_hybrid result."""
        hi = _HybridInfo()
        _check(_lib.hegj_result_get_hybrid_info(self._h, C.byref(hi)))
        return _struct_dict(hi)

    def counts(self):
        c = np.zeros(self.n, np.uint32)
        _check(_lib.hegj_result_copy_counts(self._h, c.ctypes.data))
        return c

    def csr(self, sort_rows=True):
        """(offsets[n+1] uint64, neighbors[total] uint32), rows by point id."""
        total = self.info["total_neighbors"]
        off = np.zeros(self.n + 1, np.uint64)
        nb = np.zeros(max(total, 1), np.uint32)
        _check(_lib.hegj_result_copy_csr(self._h, off.ctypes.data, nb.ctypes.data,
                                         1 if sort_rows else 0))
        return off, nb[:total]


class Server:
    """Device-resident index answering range-join requests (continuous
    batching over `slots` streams, admission against a result-buffer pool of
    `result_bytes`).  A request (queries (nq, dim), eps) returns, per query,
    the ids of every point p with sqrt(fma-chain(p - q)) <= eps.

    cell_sides: ascending grid levels; a request runs on the smallest cell
    side >= eps (times 1 - 2^-30)."""

    def __init__(self, database, cell_sides, indexed_dims=None, slots=4,
                 result_bytes=1 << 30, max_queries=1 << 20, device=0):
        arr = np.ascontiguousarray(database, dtype=np.float64)
        if arr.ndim != 2:
            raise ValueError("database must be (n, dim)")
        n, dim = arr.shape
        o = _ServeOpts()
        o.dim = dim
        o.indexed_dims = indexed_dims if indexed_dims is not None else min(dim, 6)
        sides = list(cell_sides)
        if not 1 <= len(sides) <= len(o.cell_sides):
            raise ValueError(f"1 to {len(o.cell_sides)} cell sides")
        o.n_levels = len(sides)
        for i, c in enumerate(sides):
            o.cell_sides[i] = float(c)
        o.slots = slots
        o.result_bytes = int(result_bytes)
        o.max_queries = int(max_queries)
        o.device = device
        self._h = _vp()
        self.dim = dim
        _check(_lib.hegj_serve_open(arr.ctypes.data, n, C.byref(o), C.byref(self._h)))

    def submit(self, queries, eps):
        """Queue a request; returns its ticket (the queries are copied)."""
        q = np.ascontiguousarray(queries, dtype=np.float64)
        if q.ndim != 2 or q.shape[1] != self.dim:
            raise ValueError("queries must be (nq, dim)")
        t = C.c_uint64()
        _check(_lib.hegj_serve_submit(self._h, q.ctypes.data, q.shape[0], float(eps), C.byref(t)))
        return t.value

    def wait(self, ticket):
        """Block until the request is done; returns its info dict (raises
        HegjError when the request failed)."""
        info = _RequestInfo()
        _check(_lib.hegj_serve_wait(self._h, ticket, C.byref(info)))
        return _struct_dict(info)

    def poll(self, ticket):
        """The request's info dict when it has finished, None while it runs
        (raises HegjError when it failed)."""
        info = _RequestInfo()
        st = _lib.hegj_serve_poll(self._h, ticket, C.byref(info))
        if st == 8:  # HEGJ_ERR_BUSY
            return None
        _check(st)
        return _struct_dict(info)

    def result(self, ticket, release=True):
        """(offsets[nq+1], neighbors[pairs]) of a finished request."""
        info = self.wait(ticket)
        off = np.zeros(info["n_queries"] + 1, np.uint64)
        nb = np.zeros(max(info["pairs"], 1), np.uint32)
        _check(_lib.hegj_serve_copy(self._h, ticket, off.ctypes.data, nb.ctypes.data))
        if release:
            self.release(ticket)
        return off, nb[:info["pairs"]]

    def release(self, ticket):
        _check(_lib.hegj_serve_release(self._h, ticket))

    def stats(self, reset=False):
        st = _ServeStats()
        _check(_lib.hegj_serve_get_stats(self._h, C.byref(st), 1 if reset else 0))
        return _struct_dict(st)

    def close(self):
        if self._h:
            _lib.hegj_serve_close(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def selfjoin(database, epsilon, indexed_dims=None, emulate_cell0=True, device=0, **kw):
    """One-shot: build the index, join, return (offsets, neighbors, info)."""
    with GridIndex(database, epsilon, indexed_dims, emulate_cell0, device) as ix:
        t = ix.selfjoin(**kw)
        off, nb = t.csr()
        info = t.info
        t.close()
        return off, nb, info
