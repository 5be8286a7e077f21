"""Builds the in-tree HIP library ``mprgp_amd/libmprgp_amd.so`` for gfx950.

hipcc compiles every HIP source for gfx950 only.  Packed FP32 instructions
are disabled (``-packed-fp32-ops``) on the device side.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmprgp_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


ASAN_LIB = os.path.join(HERE, "libmprgp_amd_hostasan.so")


def build_host_asan(verbose: bool = False) -> str:
    """The same library with the host code under AddressSanitizer (device code
    unchanged: -fsanitize applies to the host pass only), for the graph
    life-cycle check tests/cpp/graph_churn.cpp.  Not loaded by the package."""
    tmp = ASAN_LIB + ".tmp"
    flags = [f for f in FLAGS if f != "-O3"] + ["-O1", "-g", "-Xarch_host", "-fsanitize=address",
                                                "-Xarch_host", "-fno-omit-frame-pointer"]
    cmd = [HIPCC, *flags, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmprgp_amd_hostasan.so")
    os.replace(tmp, ASAN_LIB)
    return ASAN_LIB


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmprgp_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
