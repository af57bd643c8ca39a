"""Build recipe for the gfx950 library ``dclaw_amd/_build/libdclaw_hip.so``.

hipcc cross-compiles for gfx950 without a GPU.  Flags that matter for the
numerics: ``-ffp-contract=off`` (no FMA contraction: the reference's
statement-by-statement rounding is kept), and packed FP32 disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libdclaw_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-I" + INCLUDE, "-I" + CSRC,
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(extra=()) -> str:
    """Compile the kernels and C-ABI into the shared library (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(CSRC, "dclaw_kernels.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, *DEVICE, *extra, "-shared", "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_hip(sys.argv[1:])
