"""Build libmercury_hip.so in-tree with hipcc for gfx950.

The library is compiled with -ffp-contract=off so that no multiply-add is
fused: the CPU path it reproduces (mercury6_2.for built with
-ffp-contract=off, reference Makefile:5) rounds every product and sum.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmercury_hip.so"
SOURCES = [CSRC / "mercury_hip.hip"]
DEPENDS = SOURCES + sorted(CSRC.glob("*.hpp")) + [ROOT / "include" / "mercury_hip.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MERCURY_HIP_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(force: bool = True, verbose: bool = False) -> Path:
    """Compile the library (always when force, else only if sources are newer)."""
    if not force and LIB.exists():
        newest = max(p.stat().st_mtime for p in DEPENDS)
        if LIB.stat().st_mtime >= newest:
            return LIB
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
