"""Builds librick_hip.so (gfx950) in-tree.

    python -m rick_amd.build          # or __graft_entry__.build()

Each source compiles to an object under rick_amd/_build/ in parallel, then
one shared library is linked next to this file.  All HIP code is built with
packed-FP32 instructions disabled (-Xclang -target-feature -Xclang
-packed-fp32-ops); the gridding and phase sources additionally pin
-ffp-contract=off in the source.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librick_hip.so"
ARCH = os.environ.get("RICK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
             "-Wno-unused-result"]
HIP_SOURCES = ["grid.hip", "fft.hip", "fft_gen.hip", "phase.hip", "prep.hip", "api.cpp",
               "io.cpp"]
CXX_SOURCES = ["tables.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src.rsplit(".", 1)[0] + ".o")
        jobs.append([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src.rsplit(".", 1)[0] + ".o")
        jobs.append(["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-c",
                     str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [str(BUILD / (s.rsplit(".", 1)[0] + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
