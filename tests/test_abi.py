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
    o.dtype = 1
    assert lib.hegj_index_build(pts.ctypes.data, 10, 0.1, C.byref(o), C.byref(h)) == 2
    assert lib.hegj_index_build(None, 10, 0.1, C.byref(o), C.byref(h)) == 1
    assert lib.hegj_selfjoin(None, None, C.byref(h)) == 1
    assert lib.hegj_index_free(None) == 0


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("args,msg", [
    (["x.txt", "0.1", "2"], "Expected 6 or 5 args"),
    (["x.txt", "1.5", "2", "0"], "epsilon should be between"),
    (["x.txt", "0", "2", "0"], "epsilon should be between"),
    (["x.txt", "0.1", "2", "3"], "algorithm 0"),
    (["x.txt", "0.1", "2", "1", "0.5"], "algorithm 0"),
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
