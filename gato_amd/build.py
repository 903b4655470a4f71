"""Builds the gato_amd HIP library (gfx950) in-tree: gato_amd/libgato_amd.so.

One hipcc invocation compiles every kernel source of gato_amd/csrc together
with the C-ABI layer.  Packed-FP32 instructions are switched off
(-packed-fp32-ops) for every device compilation.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libgato_amd.so"
SOURCES = ["stencil.hip", "reduce.hip", "poisson.hip", "potentials.hip", "lanczos.hip",
           "block.hip"]
ARCH = os.environ.get("GATO_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.exists(cand) or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result", "-I", str(HERE.parent / "include"),
        *[str(CSRC / s) for s in SOURCES], "-o", str(tmp),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
