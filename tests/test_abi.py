"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/hegjoin.h declares, mirrors the reference struct layouts, and
rejects bad arguments with a status instead of crashing; the CLI keeps the
reference's argument contract (main.cu:47-71)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hegjoin.h")
CLI = os.path.join(ROOT, "hegjoin_amd", "main")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)  # drop comments
    decl = r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\*?\s*(hegj_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    import hegjoin_amd as H
    lib = H.library()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (hegj_\w+)", out))
    assert set(syms) <= exported


def test_struct_layouts_match_reference():
    import hegjoin_amd as H
    assert H.GRID_DTYPE.itemsize == 8          # struct grid
    assert H.LOOKUP_DTYPE.itemsize == 16       # struct gridCellLookup (padded)
    assert H.LOOKUP_DTYPE.fields["gridLinearID"][1] == 8
    assert H.SCHED_DTYPE.itemsize == 8         # struct schedulingCell


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "hegjoin.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
    S(hegj_index_opts) S(hegj_index_info) S(hegj_join_opts) S(hegj_result_info)
    S(hegj_serve_opts) S(hegj_request_info) S(hegj_serve_stats)
    S(hegj_hybrid_opts) S(hegj_hybrid_info)
    O(hegj_hybrid_opts, gpu_fraction) O(hegj_hybrid_info, candidates_cpu) O(hegj_hybrid_info, cpu_end_s)
    O(hegj_result_info, estimated_pairs)
    O(hegj_index_info, build_ms) O(hegj_join_opts, part_by_block) O(hegj_join_opts, use_graph)
    O(hegj_join_opts, no_prune) O(hegj_join_opts, block_waves) O(hegj_join_opts, batches)
    O(hegj_result_info, refine_used) O(hegj_result_info, graph_replay) O(hegj_serve_opts, result_bytes)
    O(hegj_serve_opts, device) O(hegj_request_info, latency_ms) O(hegj_request_info, band_rechecks)
    O(hegj_request_info, refine_used) O(hegj_serve_stats, wall_ms)
    return 0;
}
"""


def test_ctypes_mirrors_match_c_layouts(tmp_path):
    """Every ctypes structure of the Python mirror has the C header's size and
    field offsets (compiled probe)."""
    import hegjoin_amd as H
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = str(tmp_path / "probe")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run(
        [exe], capture_output=True, text=True, check=True).stdout.splitlines())
    mirror = {"hegj_index_opts": H._IndexOpts, "hegj_index_info": H._IndexInfo,
              "hegj_join_opts": H._JoinOpts, "hegj_result_info": H._ResultInfo,
              "hegj_serve_opts": H._ServeOpts, "hegj_request_info": H._RequestInfo,
              "hegj_serve_stats": H._ServeStats, "hegj_hybrid_opts": H._HybridOpts,
              "hegj_hybrid_info": H._HybridInfo}
    for name, cls in mirror.items():
        assert int(got[name]) == C.sizeof(cls), name
    for key, val in got.items():
        if "." in key:
            name, field = key.split(".")
            assert getattr(mirror[name], field).offset == int(val), key


def test_info_structs_convert_to_dicts():
    """Python mirror: scalar fields stay scalars, arrays become lists, the
    refine code becomes its name."""
    import hegjoin_amd as H
    r = H._ResultInfo()
    r.refine_used, r.total_neighbors = 2, 7
    d = H._struct_dict(r)
    assert d["refine_used"] == "mfma" and d["total_neighbors"] == 7 and "reserved" not in d
    ii = H._IndexInfo()
    ii.nCells[0] = 5
    d = H._struct_dict(ii)
    assert isinstance(d["nCells"], list) and d["nCells"][0] == 5


def test_serve_argument_validation():
    import hegjoin_amd as H
    lib = H.library()
    pts = np.random.rand(100, 3)
    o = H._ServeOpts()
    h = C.c_void_p()

    def opts(**kw):
        o.dim, o.indexed_dims, o.n_levels, o.slots = 3, 3, 2, 2
        o.cell_sides[0], o.cell_sides[1] = 0.05, 0.1
        o.result_bytes, o.max_queries, o.device = 1 << 20, 1000, 0
        for k, v in kw.items():
            setattr(o, k, v)
        return C.byref(o)
    bad = [dict(indexed_dims=4), dict(n_levels=0), dict(n_levels=9), dict(slots=0),
           dict(slots=17), dict(max_queries=0), dict(result_bytes=100)]
    for kw in bad:
        assert lib.hegj_serve_open(pts.ctypes.data, 100, opts(**kw), C.byref(h)) == 1, kw
    o2 = opts()
    o.cell_sides[1] = 0.05  # not ascending
    assert lib.hegj_serve_open(pts.ctypes.data, 100, o2, C.byref(h)) == 1
    assert lib.hegj_serve_open(None, 100, opts(), C.byref(h)) == 1
    assert lib.hegj_serve_close(None) == 0
    assert lib.hegj_serve_submit(None, None, 0, 0.1, None) == 1


def test_version_and_no_device_status():
    import hegjoin_amd as H
    assert "gfx950" in H.version()
    if H.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(H.HegjError) as e:
        H.GridIndex(np.random.rand(10, 2), 0.1)
    assert e.value.status == 3


def test_argument_validation_statuses():
    import hegjoin_amd as H
    lib = H.library()
    o = H._IndexOpts()
    lib.hegj_index_opts_default(C.byref(o), 2)
    assert (o.dim, o.indexed_dims, o.dtype, o.emulate_cell0) == (2, 2, 0, 1)
    lib.hegj_index_opts_default(C.byref(o), 90)
    assert o.indexed_dims == 6
    h = C.c_void_p()
    pts = np.random.rand(10, 2)
    bad = [(1.5, 2, 2), (0.0, 2, 2), (0.1, 2, 3), (0.1, 2, 0), (0.1, 8, 7)]
    for eps, dim, k in bad:
        lib.hegj_index_opts_default(C.byref(o), dim)
        o.indexed_dims = k
        st = lib.hegj_index_build(pts.ctypes.data, 10, eps, C.byref(o), C.byref(h))
        assert st == 1, (eps, dim, k)
    lib.hegj_index_opts_default(C.byref(o), 2)
    o.dtype = 2  # neither HEGJ_DTYPE_F64 nor HEGJ_DTYPE_F32
    assert lib.hegj_index_build(pts.ctypes.data, 10, 0.1, C.byref(o), C.byref(h)) == 1
    assert lib.hegj_index_build(None, 10, 0.1, C.byref(o), C.byref(h)) == 1
    assert lib.hegj_selfjoin(None, None, C.byref(h)) == 1
    assert lib.hegj_index_free(None) == 0


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("args,msg", [
    (["x.txt", "0.1", "2"], "Expected 6 or 5 args"),
    (["x.txt", "1.5", "2", "0"], "epsilon should be between"),
    (["x.txt", "0", "2", "0"], "epsilon should be between"),
    (["x.txt", "0.1", "2", "4"], "algorithm must be"),
    (["x.txt", "0.1", "2", "2", "1.5"], "partitioning should be between"),
    (["x.txt", "0.1", "2", "2", "0"], "partitioning should be between"),
])
def test_cli_argument_contract(args, msg):
    r = subprocess.run([CLI, *args], capture_output=True, text=True)
    assert r.returncode == 1
    assert msg in r.stderr


def test_dataset_generator_matches_oracle(oracle):
    """The product generator reproduces the reference generator's values
    (the oracle's copy is pinned to the reference's own text output)."""
    import hegjoin_amd as H
    for dim, seed in [(2, 0), (6, 7), (90, 2137834274)]:
        a = H.generate_expo(3000, dim, seed=seed)
        b = oracle.generate_expo(3000, dim, seed=seed)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_dataset_loader_matches_oracle(oracle, tmp_path):
    import hegjoin_amd as H
    pts = oracle.generate_expo(20000, 3, seed=3)
    for prec in (6, 17):
        p = str(tmp_path / f"d{prec}.txt")
        oracle.write_text(p, pts, prec)
        a = H.load_dataset(p, 3)
        b = oracle.load_text(p, 3)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    pb = str(tmp_path / "d.bin")
    pts.tofile(pb)
    assert np.array_equal(H.load_dataset(pb, 3), pts)
    p = str(tmp_path / "r.txt")
    open(p, "w").write("0.1, 0.2\n0.3,0.4\n 0.5 ,\t0.6\n0.7\n")
    assert H.load_dataset(p, 2).tolist() == [[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]]


def test_tuning_table_roundtrip(tmp_path):
    """autotune.write_tables -> hegj_tuning_load -> hegj_join_opts_tuned."""
    import hegjoin_amd as H
    from hegjoin_amd import autotune as T
    best = dict(refine="fp64", unit_order="workload", queries_per_unit=8, xcd_parts=1,
                units_per_grab=4, block_waves=2, batches=3, ms=1.0)
    tsv, js = T.write_tables(7, dict(dim=7), best, [best], str(tmp_path / "t.tsv"),
                             str(tmp_path / "t.json"))
    assert H.load_tuning(tsv) == 1
    o = H._JoinOpts()
    H.library().hegj_join_opts_tuned(C.byref(o), 7)
    assert (o.queries_per_unit, o.unit_order, o.xcd_parts, o.units_per_grab, o.refine) == (8, 1, 1, 4, 1)
    assert (o.block_waves, o.batches) == (2, 3)
    H.library().hegj_join_opts_tuned(C.byref(o), 5)  # no entry: plain defaults
    assert (o.queries_per_unit, o.unit_order) == (0, 2)
    bad = tmp_path / "bad.tsv"
    bad.write_text("# comment\n3 4 2 1\n")
    with pytest.raises(H.HegjError):
        H.load_tuning(str(bad))
    bad.write_text("3 5 2 1 0 1\n")  # queries_per_unit 5 is not a kernel
    with pytest.raises(H.HegjError):
        H.load_tuning(str(bad))
    bad.write_text("3 4 2 1 0 1 3 1\n")  # 3 waves per workgroup: not offered
    with pytest.raises(H.HegjError):
        H.load_tuning(str(bad))
    six = tmp_path / "six.tsv"  # tables without the last two columns still load
    six.write_text("3 4 2 1 0 1\n")
    assert H.load_tuning(str(six)) == 1
    H.library().hegj_join_opts_tuned(C.byref(o), 3)
    assert (o.queries_per_unit, o.block_waves, o.batches) == (4, 0, 0)
    with pytest.raises(H.HegjError):
        H.load_tuning(str(tmp_path / "missing.tsv"))


def test_bench_rooflines_pick_the_binding_roof():
    """2-D (few flops per pair, many pairs) binds on HBM; 6-D FP64 on compute;
    the MFMA refine uses 2D flops per pair against bf16 peak / 3."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    B = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(B)
    roof, other = B.rooflines(1_000_000, 2, 20_000, 1.06e9, 3.0e9, 10.0, "fp64")
    assert roof["bound"] == "hbm" and other["bound"] == "valu-fp64"
    assert abs(roof["work"] - (8 * 2e6 + 4e6 + 24 * 2e4 + 8 * 1.06e9)) < 1
    roof, other = B.rooflines(5_000_000, 6, 1_860_000, 3.0e8, 2.4e10, 100.0, "fp64")
    assert roof["bound"] == "valu-fp64" and abs(roof["peak"] - 78.6) < 1e-9
    # the FP32 filter: unpacked FP32 VALU (the build emits no packed FP32)
    roof, other = B.rooflines(5_000_000, 6, 1_860_000, 3.0e8, 2.4e10, 100.0, "f32")
    assert roof["bound"] == "valu-fp32" and abs(roof["peak"] - 157.3 / 2) < 1e-9
    assert abs(roof["achieved"] - 3 * 6 * 2.4e10 / 0.1 / 1e12) < 1e-9
    roof, other = B.rooflines(2_000_000, 18, 2900, 7.4e7, 2.8e12, 500.0, "mfma")
    assert roof["bound"] == "mfma" and abs(roof["peak"] - 2500 / 3) < 1e-9
    assert abs(roof["frac"] - 2 * 18 * 2.8e12 / 0.5 / 1e12 / (2500 / 3)) < 1e-12


def _fmaf(a, b, c):
    """float32 fma: the exact product of two float32 values fits a double;
    one rounding of product + c to float32 (c and the product are exact in
    float64 up to a final rounding far below float32 precision)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def test_f32_filter_thresholds_are_sound(oracle):
    """The FP32 filter's decisions agree with the reference's FP64 chain for
    every pair outside the band: pairs planted at eps(1 +- tiny) with
    coordinates up to `radius` from the centre, in 2, 6 and 8 dims."""
    import hegjoin_amd as H
    rng = np.random.default_rng(5)
    checked = 0
    for dim, eps, R in [(2, 0.001, 0.5), (6, 0.0063, 0.5), (6, 0.002, 8.0), (8, 0.05, 1.0)]:
        tacc, trej = H.f32_thresholds(eps, dim, R)
        T = oracle.threshold(eps)
        assert tacc < T < trej and (trej - tacc) < 0.02 * T
        centre = rng.uniform(-3, 3, dim)
        q = centre + rng.uniform(-R, R, (4000, dim)) * 0.999
        d = rng.standard_normal((4000, dim))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rel = rng.choice([-1e-3, -1e-5, -1e-7, 0.0, 1e-7, 1e-5, 1e-3], 4000)
        c = q + d * (eps * (1 + rel))[:, None]
        c = centre + np.clip(c - centre, -R, R)
        qf = (q - centre).astype(np.float32)   # fl32(fl64(x - c)): numpy rounds to nearest
        cf = (c - centre).astype(np.float32)
        for i in range(len(q)):
            acc32 = np.float32(0)
            acc = 0.0
            for l in range(dim):
                df = np.float32(cf[i, l] - qf[i, l])
                acc32 = _fmaf(df, df, acc32)
                dd = c[i, l] - q[i, l]
                # FP64 multiply-add: differs from the reference's fma chain by
                # ~1e-15 relative, far inside the band (>= 1e-6 relative)
                acc = dd * dd + acc
            # outside the band the filter's decision must be the FP64 one
            if acc32 <= tacc:
                assert acc <= T, (dim, i)
            elif acc32 > trej:
                assert acc > T, (dim, i)
            checked += 1
    assert checked == 16000


def test_f32_thresholds_reject_wide_bands():
    import hegjoin_amd as H
    with pytest.raises(H.HegjError) as e:
        H.f32_thresholds(1e-6, 6, 1e3)  # tiny eps, huge coordinates
    assert e.value.status == 2


def test_release_cached_memory_is_safe_without_a_device():
    import hegjoin_amd as H
    H.release_cached_memory()
    H.release_cached_memory(0)


def test_dataset_binary_cache(oracle, tmp_path):
    """HEGJ_LOAD_BIN_CACHE: the first load parses the text and writes
    <path>.d<dim>.hegjbin; later loads read that file while the text is
    unchanged (shown by planting a value in the cache), and re-parse once the
    text changes."""
    import hegjoin_amd as H
    pts = oracle.generate_expo(5000, 3, seed=5)
    p = str(tmp_path / "d.txt")
    oracle.write_text(p, pts, 17)
    a = H.load_dataset(p, 3, bin_cache=True)
    assert np.array_equal(a, pts)
    cache = p + ".d3.hegjbin"
    raw = bytearray(open(cache, "rb").read())
    assert raw[:8] == b"HEGJBIN1" and len(raw) == 56 + pts.nbytes
    planted = np.frombuffer(bytes(raw[56:]), np.float64).copy()
    planted[7] = 123.0
    raw[56:] = planted.tobytes()
    open(cache, "wb").write(bytes(raw))
    b = H.load_dataset(p, 3, bin_cache=True)
    assert b.reshape(-1)[7] == 123.0  # read from the cache, not re-parsed
    assert np.array_equal(H.load_dataset(p, 3), pts)  # no flag: the text
    with open(p, "a") as f:  # the text changes: the cache is stale
        f.write("0.5,0.25,0.125\n")
    c = H.load_dataset(p, 3, bin_cache=True)
    assert c.shape == (5001, 3) and np.array_equal(c[:5000], pts)
    assert np.array_equal(H.load_dataset(p, 3, bin_cache=True), c)
    # float parsing keeps its own cache
    f = H.load_dataset(p, 3, f32=True, bin_cache=True)
    assert os.path.exists(p + ".d3f.hegjbin")
    assert f.shape == c.shape


def _decimal(fr, digits=60):
    """Exact decimal text of a dyadic fraction in [0, 1) (60 digits suffice
    for the values below)."""
    q = fr * 10**digits
    assert q.denominator == 1
    t = str(q.numerator).rjust(digits + 1, "0")
    return t[:-digits] + "." + t[-digits:]


def test_dataset_float_parse_is_correctly_rounded(tmp_path):
    """HEGJ_LOAD_F32 (`ss >> float` with DTYPE=float): every value is the
    float nearest to the decimal text, ties to even, checked with exact
    rational arithmetic.  The cases include exact midpoints between adjacent
    floats and midpoints nudged by 1e-45, which a parse through an
    intermediate double rounds the wrong way."""
    from fractions import Fraction
    import hegjoin_amd as H
    rng = np.random.default_rng(9)
    vals = [repr(float(v)) for v in rng.random(2000)]
    ties = set()
    for v in rng.random(300).astype(np.float32):
        lo = Fraction(float(v))
        mid = (lo + Fraction(float(np.nextafter(v, np.float32(2))))) / 2
        ties.add(len(vals))
        vals.append(_decimal(mid))
        vals.append(_decimal(mid + Fraction(1, 10**45)))
        vals.append(_decimal(mid - Fraction(1, 10**45)))
    p = str(tmp_path / "f.txt")
    open(p, "w").write("\n".join(vals) + "\n")
    got = H.load_dataset(p, 1, f32=True).reshape(-1)
    assert len(got) == len(vals)
    for i, (s, g) in enumerate(zip(vals, got)):
        x = Fraction(s)
        g32 = np.float32(g)
        assert float(g32) == g
        err = abs(Fraction(float(g32)) - x)
        for nb in (np.nextafter(g32, np.float32(-2)), np.nextafter(g32, np.float32(2))):
            assert err <= abs(Fraction(float(nb)) - x), s
        if i in ties:  # exactly halfway: the even neighbour
            assert int(g32.view(np.uint32)) % 2 == 0, s
