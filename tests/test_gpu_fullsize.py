"""The self-join workloads at their full sizes against the oracle's pair count
and pair hash (tests/golden/fullsize.json, written by make_fullsize.py):
2-D 1M (eps 0.001), 6-D 5M (64 neighbours), 18-D 2M (32 neighbours) and 90-D
1M (16 neighbours), points from the reference generator with seed 0.

Every refine the library offers for a workload must give the oracle's pair
set: the FP64 chain, the FP32 filter with its FP64 band recheck (2-D, 6-D)
and the split-BF16 MFMA guard band with its FP64 band recheck (18-D, 90-D).
The hash is sum of mix64(query << 32 | neighbour) over all pairs, so equal
hashes and counts mean equal pair multisets up to a 2^-64 collision."""
import json
import os

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "fullsize.json")) as _f:
    FULL = json.load(_f)

# (workload, refine): the plain FP64 kernel at 90-D takes minutes (every
# point is a candidate of every query); it is timed by bench.py instead
CASES = [("2d", "auto"), ("2d", "fp64"), ("2d", "f32"), ("6d", "auto"), ("6d", "fp64"),
         ("6d", "f32"), ("18d", "mfma"), ("18d", "fp64"), ("90d", "mfma")]
CASES = [c for c in CASES if c[0] in FULL]


@pytest.fixture(scope="module")
def H():
    import hegjoin_amd as H
    if H.device_count() == 0:
        pytest.fail("no HIP device visible for a GPU test")
    return H


@pytest.fixture(scope="module")
def indexes(H):
    """One resident workload at a time (the 6-D index and a 1.2e9-pair 2-D
    result would otherwise share the device), closed at the end."""
    held = {}

    def get(key):
        if key not in held:
            for k in list(held):
                held.pop(k).close()
            f = FULL[key]
            pts = H.generate_expo(f["n"], f["dim"], seed=f["seed"])
            held[key] = H.GridIndex(pts, f["eps"], indexed_dims=f["indexed_dims"])
        return held[key]
    yield get
    for ix in held.values():
        ix.close()


@pytest.mark.parametrize("key,refine", CASES, ids=[f"{k}-{r}" for k, r in CASES])
def test_fullsize_pair_set_equals_oracle(H, indexes, key, refine):
    f = FULL[key]
    print("case", key, refine, f["n"], "x", f["dim"], flush=True)
    ix = indexes(key)
    t = ix.selfjoin(refine=refine) if refine != "auto" else ix.selfjoin()
    info = t.info
    t.close()
    print(f"  {info['refine_used']}: {info['total_neighbors']} pairs, join kernel "
          f"{info['join_kernel_avg_ms']:.1f} ms, band rechecks {info['band_rechecks']}", flush=True)
    if refine != "auto":
        assert info["refine_used"] == refine
    assert info["total_neighbors"] == f["pairs"]
    assert "%016x" % info["pair_hash"] == f["pair_hash"]


# the MFMA workloads at full size (bench.py CONFIGS): no CPU oracle finishes
# them in test time, so the plain FP64 kernel (the reference's chain, pinned
# against the oracle on every fixture) is the yardstick
@pytest.mark.parametrize("key", ["18d", "90d"])
def test_fullsize_mfma_equals_fp64_kernel(H, key):
    """18-D 2M and 90-D 1M: the split-BF16 MFMA guard band with its FP64 band
    recheck gives the plain FP64 kernel's pair set (count and pair hash),
    with rechecks in the band.  Prints both kernels' times."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import CONFIGS
    c = CONFIGS[key]
    print("case", key, c["n"], "x", c["dim"], flush=True)
    pts = H.generate_expo(c["n"], c["dim"], seed=0)
    with H.GridIndex(pts, c["eps"], indexed_dims=c["k"]) as ix:
        got = {}
        for refine in ("mfma", "fp64"):
            t = ix.selfjoin(refine=refine)
            got[refine] = t.info
            t.close()
            print(f"  {refine}: {got[refine]['total_neighbors']} pairs, hash "
                  f"{got[refine]['pair_hash']:016x}, join kernel {got[refine]['join_kernel_avg_ms']:.1f} ms, "
                  f"band rechecks {got[refine]['band_rechecks']}", flush=True)
    m, f = got["mfma"], got["fp64"]
    assert m["refine_used"] == "mfma" and f["refine_used"] == "fp64"
    assert m["total_neighbors"] == f["total_neighbors"]
    assert m["pair_hash"] == f["pair_hash"]
    assert m["band_rechecks"] > 0
