"""Builds the engine's shared library (HIP for gfx950) in-tree.

``python -m opensfm_amd._build`` or ``__graft_entry__.build()``.  The library
lands in ``opensfm_amd/_lib/`` so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libopensfm_amd.so")
SOURCES = ["pm_kernels.hip", "pm_capi.cpp", "pm_batch.cpp", "nn_kernels.hip", "nn_capi.cpp",
           "fuse_kernels.hip", "fuse_capi.cpp", "covis_kernels.hip", "covis_capi.cpp",
           "lz4_codec.cpp"]
ARCH = os.environ.get("OPENSFM_AMD_ARCH", "gfx950")

# -ffp-contract=off: the exact scorer and the host-side homography/weight
# tables follow the reference's FP32/FP64 operation order; FMAs are asked for
# explicitly where wanted.  -packed-fp32-ops: no packed FP32 VALU ops.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", CSRC, "-o", tmp] + [
        os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
