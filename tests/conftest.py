"""Shared test configuration: the `gpu` marker, import paths and builders of
the CPU-side test tools (the oracle library and the host lane emulator)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EMU_SRC = os.path.join(ROOT, "tests", "emu", "strip_emu.cpp")
EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libstrip_emu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def build_emulator() -> str:
    """Compile the host lane emulator (product kernel body on the CPU).  It
    is host code only and links no HIP runtime (--as-needed drops it), so a
    GPU test process can load it beside torch's runtime."""
    srcs = [EMU_SRC] + [os.path.join(ROOT, "dclaw_amd", "csrc", f)
                        for f in ("dclaw_strip.h", "dclaw_physics.h", "dclaw_amr.h",
                                  "dclaw_tangent.h", "dclaw_tangent_k.h", "dclaw_adjoint.h",
                                  "dclaw_adjoint_src2.h",
                                  "dclaw_strip_body.inc", "dclaw_prepare_body.inc",
                                  "dclaw_physics_body.inc", "dcmath.h")]
    if os.path.exists(EMU_LIB) and all(os.path.getmtime(EMU_LIB) >= os.path.getmtime(s)
                                       for s in srcs):
        return EMU_LIB
    os.makedirs(os.path.dirname(EMU_LIB), exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
           "-fPIC", "-ffp-contract=off", "-fno-fast-math",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "dclaw_amd", "csrc"),
           "-Wl,--as-needed", "-shared", "-o", EMU_LIB, EMU_SRC]
    subprocess.run(cmd, check=True)
    return EMU_LIB


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def emu_lib():
    import ctypes
    return ctypes.CDLL(build_emulator())
