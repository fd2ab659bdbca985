"""Builds libtbsla_hip.so in-tree with hipcc for gfx950.

    python -m tbsla_amd.build          # or tbsla_amd.build.build()

Every HIP source under tbsla_amd/csrc is compiled to an object, in parallel,
and linked into tbsla_amd/libtbsla_hip.so.  Packed FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtbsla_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


CPP_PROGRAMS = [
    # (source, output, rpath relative to the output)
    (os.path.join(PKG, "apps", "tbsla_perf_hip.cpp"), os.path.join(PKG, "bin", "tbsla_perf_hip"),
     "$ORIGIN/.."),
    (os.path.join(PKG, "apps", "tbsla_attention_hip.cpp"),
     os.path.join(PKG, "bin", "tbsla_attention_hip"), "$ORIGIN/.."),
    (os.path.join(PKG, "apps", "tbsla_perf_page_rank_hip.cpp"),
     os.path.join(PKG, "bin", "tbsla_perf_page_rank_hip"), "$ORIGIN/.."),
    (os.path.join(PKG, "apps", "tbsla_perf_cg_hip.cpp"), os.path.join(PKG, "bin", "tbsla_perf_cg_hip"),
     "$ORIGIN/.."),
    (os.path.join(PKG, "apps", "tbsla_mm_to_bin_hip.cpp"),
     os.path.join(PKG, "bin", "tbsla_mm_to_bin_hip"), "$ORIGIN/.."),
    (os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp"),
     os.path.join(ROOT, "tests", "cpp", "test_dropin"), "$ORIGIN/../../tbsla_amd"),
]


def build_programs() -> None:
    """Host programs over the C ABI (plain g++, no HIP headers needed)."""
    for src, out, rpath in CPP_PROGRAMS:
        if not os.path.exists(src):
            continue
        os.makedirs(os.path.dirname(out), exist_ok=True)
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", out,
               "-L" + PKG, "-ltbsla_hip", "-Wl,-rpath," + rpath]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on {src}:\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_programs()
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
