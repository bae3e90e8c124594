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

SOURCES = [os.path.join(CSRC, f) for f in ("cfft_hip.hip", "blocks.hip", "scft.hip", "mlp.hip", "dense.hip",
                                            "ensemble.hip", "p3c.hip", "plan.cpp")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose: bool = True) -> str:
    """Compile every source to an object in parallel, then link."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    cflags = [f for f in FLAGS if f != "-shared"]
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, _run([HIPCC, *cflags, "-c", "-o", obj, src], verbose)))
    failed = False
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building libcfft_hip.so")
    tmp = LIB + ".tmp"
    pr = _run([HIPCC, *FLAGS, "-o", tmp, *objs], verbose)
    out, _ = pr.communicate()
    if pr.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError("hipcc failed linking libcfft_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
