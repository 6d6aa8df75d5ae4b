"""Builds the engine's shared library (HIP for gfx950) in-tree.

``python -m opensfm_amd._build`` or ``__graft_entry__.build()``.  The library
lands in ``opensfm_amd/_lib/`` so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libopensfm_amd.so")
OBJDIR = os.path.join(LIBDIR, "obj")
SOURCES = ["pm_kernels.hip", "pm_capi.cpp", "pm_batch.cpp", "nn_kernels.hip", "nn_capi.cpp",
           "fuse_kernels.hip", "fuse_capi.cpp", "covis_kernels.hip", "covis_capi.cpp",
           "lz4_codec.cpp", "pm_noalloc.hip"]
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


def _compile(src: str, obj: str, verbose: bool) -> None:
    cmd = [hipcc()] + [f for f in FLAGS if f != "-shared"] + ["-I", CSRC, "-c", "-o", obj + ".tmp",
                                                           os.path.join(CSRC, src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n" + res.stdout + res.stderr)
    os.replace(obj + ".tmp", obj)


def _stale(src: str, obj: str) -> bool:
    """True when obj is older than its source or any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [
        os.path.join(d, f) for d in (CSRC, os.path.join(os.path.dirname(HERE), "include"))
        for f in os.listdir(d) if f.endswith(".h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source (concurrently, one hipcc per file) and links the
    library.  incremental=True recompiles only objects older than their
    source or a header."""
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    todo = [(s, o) for s, o in zip(SOURCES, objs) if not incremental or _stale(s, o)]
    workers = max(1, min(len(todo), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_compile, s, o, verbose) for s, o in todo]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
