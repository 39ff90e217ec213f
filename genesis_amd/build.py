"""Build recipe for the gfx950 shared library (``libgenesis_gpu.so``).

hipcc cross-compiles for gfx950 without a GPU present.  The library is
built in-tree so that it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgenesis_gpu.so")
SOURCES = [os.path.join(CSRC, "genesis_gpu.hip")]
HEADERS = [os.path.join(CSRC, "gs_device.hpp"), os.path.join(ROOT, "include", "genesis_gpu.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = True) -> str:
    """Compile every HIP source into libgenesis_gpu.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), *SOURCES, "-o", tmp]
    if verbose:
        print("[genesis_amd] " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
