"""Compile libgmxcorr_hip.so for gfx950 in-tree with hipcc.

    python -m gmx_correlation_amd.build

Cross-compiles without a GPU.  Packed FP32 instructions are disabled for all
device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libgmxcorr_hip.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

SOURCES = [
    "ksg_device.hip", "ksg_host.hip", "dense_device.hip", "dense_host.hip",
    "surrogate_device.hip", "surrogate_host.hip", "batch_host.hip",
    "te_device.hip", "te_host.hip", "xtc_host.hip",
]

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + INCLUDE,
    "-I" + CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every source to an object in parallel (each file launches only
    its own kernels, so no relocatable device code is needed), then link the
    shared library.  Objects live in a temporary directory."""
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1), 16)
    with tempfile.TemporaryDirectory(prefix="gmxcorr_build_") as tmpdir:
        def compile_one(src: str) -> str:
            obj = os.path.join(tmpdir, os.path.splitext(src)[0] + ".o")
            cmd = [hipcc(), *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            return obj

        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = list(ex.map(compile_one, SOURCES))
        tmp = LIB_PATH + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
