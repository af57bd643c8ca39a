"""Shared test configuration: the `gpu` marker, import paths and builders of
the CPU-side test tools (the oracle library and the host lane emulator)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EMU_SRCS = [os.path.join(ROOT, "tests", "emu", f) for f in ("strip_emu.cpp", "adjoint_emu.cpp")]
EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libstrip_emu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def build_emulator() -> str:
    """Compile the host lane emulator (product kernel body on the CPU).  It
    is host code only and links no HIP runtime (--as-needed drops it), so a
    GPU test process can load it beside torch's runtime.  Its two units
    (the strip body; the adjoint stages) compile in parallel."""
    hdrs = [os.path.join(ROOT, "dclaw_amd", "csrc", f)
            for f in ("dclaw_strip.h", "dclaw_physics.h", "dclaw_amr.h",
                      "dclaw_tangent.h", "dclaw_tangent_k.h", "dclaw_adjoint.h",
                      "dclaw_adjoint_src2.h",
                      "dclaw_strip_body.inc", "dclaw_prepare_body.inc",
                      "dclaw_physics_body.inc", "dcmath.h")]
    if os.path.exists(EMU_LIB) and all(os.path.getmtime(EMU_LIB) >= os.path.getmtime(s)
                                       for s in EMU_SRCS + hdrs):
        return EMU_LIB
    bdir = os.path.dirname(EMU_LIB)
    os.makedirs(bdir, exist_ok=True)
    flags = ["-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-fPIC",
             "-ffp-contract=off", "-fno-fast-math",
             "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "dclaw_amd", "csrc")]
    objs = [os.path.join(bdir, os.path.basename(src)[:-4] + ".o") for src in EMU_SRCS]
    procs = [subprocess.Popen(["/opt/rocm/bin/hipcc", *flags, "-c", "-o", o, src])
             for src, o in zip(EMU_SRCS, objs)]
    rcs = [pr.wait() for pr in procs]
    if any(rcs):
        raise RuntimeError(f"emulator build failed: {rcs}")
    tmp = EMU_LIB + ".tmp"
    subprocess.run(["/opt/rocm/bin/hipcc", "--cuda-host-only", "-Wl,--as-needed", "-shared",
                    "-o", tmp, *objs], check=True)
    os.replace(tmp, EMU_LIB)
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
