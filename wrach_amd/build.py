"""Builds libwrach_hip.so in-tree with hipcc for gfx950.

    python -m wrach_amd.build        # or __graft_entry__.build()

Flags that matter for results:
  -ffp-contract=off       no FMA contraction: a*b+c rounds twice, like the
                          reference's Rust; f32 division and sqrtf stay the
                          correctly rounded sequences hipcc emits by default.
  -target-feature -packed-fp32-ops
                          no packed FP32 instructions in the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libwrach_hip.so")
SOURCES = ["kernels.hip", "learn.hip", "abi.hip"]
ARCH = os.environ.get("WRACH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libwrach_hip.so)")


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-I", INCLUDE,
        "-I", CSRC,
    ]


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    cc = hipcc()
    objs = []
    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(PKG, "_build", src.replace(".hip", ".o"))
        cmd = [cc, *compile_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
