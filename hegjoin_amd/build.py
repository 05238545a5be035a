"""Builds libhegjoin.so and the `main` CLI for gfx950 with hipcc, in-tree.

    python hegjoin_amd/build.py            # library + CLI
    python hegjoin_amd/build.py --params "-DGPUNUMDIM=6 -DNUMINDEXEDDIM=6"

Objects go to hegjoin_amd/_build/, the library to hegjoin_amd/libhegjoin.so,
the CLI to hegjoin_amd/main (rpath $ORIGIN).  PARAMS mirrors the reference
makefile's $(PARAMS) hook (makefile:17): compile-time GPUNUMDIM /
NUMINDEXEDDIM / DTYPE for the CLI.
"""
from __future__ import annotations

import argparse
import os
import shlex
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhegjoin.so")
CLI = os.path.join(PKG, "main")
ARCH = os.environ.get("HEGJ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hegj_prims.hip", "hegj_index.hip", "hegj_join.hip", "hegj_util.hip",
           "hegj_mfma.hip", "hegj_query.hip", "hegj_serve.cpp", "hegj_dataset.cpp", "hegj_ego.cpp",
           "hegj_scratch.hip", "hegj_api.cpp"]

# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Per-file additions.  The MFMA join keeps its accumulator tiles in ordinary
# VGPRs (gfx950 MFMAs read and write either file): in the AGPR form every
# reject test first copied each accumulator out with v_accvgpr_read (128 of
# them per two chunks at 18-D; 501 -> 486 ms per launch without them).
FILE_FLAGS = {"hegj_mfma.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(params: str = "", jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s + ".o") for s in srcs]

    def compile_one(i):
        cmd = ["hipcc", *HIPFLAGS, *FILE_FLAGS.get(srcs[i], []), "-c", os.path.join(CSRC, srcs[i]),
               "-o", objs[i]]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli_cmd = ["g++", "-O2", "-std=c++17", "-Wall", *shlex.split(params),
               "-I" + os.path.join(PKG, "..", "include"), os.path.join(CSRC, "main.cpp"),
               "-o", CLI + ".tmp", "-L" + PKG, "-lhegjoin", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    if os.path.exists(os.path.join(CSRC, "main.cpp")):
        _run(cli_cmd)
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--params", default="")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.params, a.jobs, a.verbose))
