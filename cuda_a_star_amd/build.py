"""Builds libgastar.so (HIP for gfx950 + host C++) in-tree with hipcc.

Usage: python -m cuda_a_star_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgastar.so")
CLI = os.path.join(HERE, "gastar_cli")

SOURCES = [os.path.join(CSRC, "engine.hip"), os.path.join(CSRC, "checks.hip"), os.path.join(CSRC, "instances.cpp")]
CLI_SOURCE = os.path.join(CSRC, "cli.hip")
HEADERS = [os.path.join(INCLUDE, "gastar.h"), os.path.join(INCLUDE, "gastar", "core.hpp")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GASTAR_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    # keep the compiler from emitting packed fp32 instructions (none are needed here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]
FLAGS = COMMON + ["-fPIC", "-shared"]


def _hipcc(args, out, verbose):
    tmp = out + ".tmp"
    cmd = [HIPCC, *args, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    """Compile libgastar.so and the gastar_cli executable (always recompiles)."""
    _hipcc([*FLAGS, *SOURCES], LIB, verbose)
    _hipcc([*COMMON, CLI_SOURCE], CLI, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
