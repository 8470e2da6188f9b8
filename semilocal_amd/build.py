"""Builds the engine's shared library (HIP for gfx950) in-tree.

``python -m semilocal_amd.build`` compiles every ``csrc/*.hip`` file into
``semilocal_amd/libsemilocal_hip.so``.  The library is the product: host C++
and the hand-written HIP kernels behind the C ABI in include/semilocal_hip.h.
``build_cli()`` then links the command-line drivers (``cli/*.cpp``, host C++
over the C ABI) into ``semilocal_amd/bin/``.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libsemilocal_hip.so")
ARCH = os.environ.get("SEMILOCAL_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, *(extra or []), "-I", os.path.join(ROOT, "include"), "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose: bool = False) -> list[str]:
    """Links every ``cli/*.cpp`` driver against the library into ``bin/``
    (rpath ``$ORIGIN/..``, so the drivers find the in-tree library)."""
    out_dir = os.path.join(PKG, "bin")
    os.makedirs(out_dir, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    built = []
    for src in sorted(glob.glob(os.path.join(PKG, "cli", "*.cpp"))):
        exe = os.path.join(out_dir, os.path.splitext(os.path.basename(src))[0])
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-march=x86-64-v3", "-I", os.path.join(ROOT, "include"),
               "-o", exe + ".tmp", src, "-L", PKG, "-lsemilocal_hip", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        built.append(exe)
    return built


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
    build_cli(verbose=True)
