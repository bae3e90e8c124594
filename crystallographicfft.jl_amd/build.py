"""Build libcfft_hip.so (gfx950) in-tree with hipcc.

Usage: python crystallographicfft.jl_amd/build.py
The library is written next to this file as ``libcfft_hip.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfft_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no packed-FP32 instructions are generated (hardware note for gfx950)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = [os.path.join(CSRC, "cfft_hip.hip"), os.path.join(CSRC, "plan.cpp")]


def build(verbose: bool = True) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libcfft_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
