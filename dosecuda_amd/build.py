"""Builds libdosehip.so (the engine) in-tree with hipcc for gfx950.

    python -m dosecuda_amd.build

The library is written next to this file so that it travels with the source
tree; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdosehip.so")

SOURCES = ["impt.hip", "imrt.hip", "hu.hip", "influence.hip", "surrogate.hip", "scheduler.hip", "tuner.hip", "noalloc.hip", "host_tables.cpp", "runtime.cpp", "api.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # decision arithmetic (ray exits, rounded neighbour indices) must not be
    # fused behind our back; explicit fmaf() stays fused
    "-ffp-contract=off",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the engine needs ROCm's hipcc to build")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdosehip.so")
    os.replace(tmp, LIB)
    return LIB


def build_sanitize_host(verbose: bool = False) -> str:
    """CPU-only executable of the GPU-free host set-up (host_tables.cpp) and
    its driver (sanitize_host.cpp), built with AddressSanitizer and
    UndefinedBehaviorSanitizer on the host side only; written to
    dosecuda_amd/_build/sanitize_host.  Returns its path."""
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "sanitize_host")
    cmd = [hipcc(), "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fno-omit-frame-pointer",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all",
           "-o", exe, os.path.join(CSRC, "host_tables.cpp"), os.path.join(CSRC, "sanitize_host.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building the host sanitizer check")
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
