"""Accuracy of the portable elementary functions (dclaw_amd/csrc/dcmath.h)
against x86 80-bit long-double libm (64-bit mantissa) as the reference, and
the effect of the choice of libm on a step of the scheme."""
import ctypes
import math

import numpy as np
import pytest

from dclaw_amd import workloads as W
from helpers import TOL_STEP, advance_oracle, rel_linf_wet

_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def dcm(emu_lib):
    emu_lib.emu_dcm.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]

    def f(fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros_like(x)
        emu_lib.emu_dcm(fn, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                        out.ctypes.data_as(_dp), x.size)
        return out
    return f


def ulp_err(got, ref_ld):
    got_ld = got.astype(np.longdouble)
    ref64 = ref_ld.astype(np.float64)
    ulp = np.spacing(np.abs(ref64)).astype(np.longdouble)
    return np.max(np.abs(got_ld - ref_ld) / ulp)


RNG = np.random.default_rng(0)


@pytest.mark.parametrize("fn,npf,lo,hi,max_ulp", [
    (0, np.exp, -30.0, 30.0, 1.25),
    (0, np.exp, -740.0, 709.0, 1.25),
    (1, np.arctan, -1.0, 1.0, 1.0),
    (1, np.arctan, -50.0, 50.0, 1.0),
    (2, np.tan, -1.5, 1.5, 3.0),
    (3, np.tanh, -0.6, 0.6, 2.5),
    (3, np.tanh, -25.0, 25.0, 2.5),
    (4, np.log, 1e-6, 1e6, 1.0),
])
def test_accuracy(dcm, fn, npf, lo, hi, max_ulp):
    x = RNG.uniform(lo, hi, 20000)
    got = dcm(fn, x)
    ref = npf(x.astype(np.longdouble))
    assert ulp_err(got, ref) <= max_ulp


def test_special_values(dcm):
    assert dcm(0, [0.0, -800.0, 800.0, -np.inf]).tolist() == [1.0, 0.0, np.inf, 0.0]
    assert np.isnan(dcm(0, [np.nan])[0])
    assert dcm(1, [0.0, np.inf, -np.inf]).tolist() == [0.0, math.pi / 2, -math.pi / 2]
    assert dcm(3, [0.0, 30.0, -30.0]).tolist() == [0.0, 1.0, -1.0]
    assert np.isnan(dcm(2, [np.inf])[0])
    assert dcm(4, [1.0])[0] == 0.0 and dcm(4, [0.0])[0] == -np.inf
    assert dcm(5, [10.0, 0.0, 2.0], [0.0, 2.0, 10.0]).tolist() == [1.0, 0.0, 1024.0]


def test_subnormal_exp(dcm):
    x = np.linspace(-745.0, -708.0, 500)
    got = dcm(0, x)
    ref = np.exp(x.astype(np.longdouble))
    # subnormal results: absolute error within one subnormal step
    assert np.max(np.abs(got.astype(np.longdouble) - ref)) <= 2 * 4.9406564584124654e-324


def test_libm_choice_per_step(oracle_mod):
    """One step with dcmath vs glibc elementary functions: within the
    per-step bar (the two libms differ by about an ulp)."""
    O = oracle_mod
    pt = W.inclined_plane_mound()
    q, aux, dt = advance_oracle(O, pt, 30)
    a, _, ca = O.stepgrid(O.params_from(pt.params), q, aux, 1.0, 1.0, dt)
    with O.libm("glibc"):
        b, _, cb = O.stepgrid(O.params_from(pt.params), q, aux, 1.0, 1.0, dt)
    assert rel_linf_wet(a, b, 1e-3) <= TOL_STEP
    assert abs(ca - cb) <= TOL_STEP * cb


def test_scheme_amplifies_ulp_perturbations(oracle_mod):
    """Why the GPU must reproduce the CPU's elementary functions bit for bit:
    two oracle runs of the 200x200 workload that differ by 2 ulp in two
    momentum values at step 5 separate by orders of magnitude within 150
    steps, so no libm-independent 1e-9 bar over 1000 steps exists."""
    O = oracle_mod
    pt = W.inclined_plane_mound()
    P = O.params_from(pt.params)
    qa, aux, dt, _, _ = O.run(P, pt.q, pt.aux, 1.0, 1.0, 5, 1e-3, 0.9, 1e9)
    qb = qa.copy(order="F")
    for i, j in [(42, 103), (42, 115)]:
        qb[1, i, j] = np.nextafter(np.nextafter(qb[1, i, j], 2.0), 2.0)
    ra = O.run(P, qa, aux, 1.0, 1.0, 150, dt, 0.9, 1e9)[0]
    rb = O.run(P, qb, aux, 1.0, 1.0, 150, dt, 0.9, 1e9)[0]
    assert rel_linf_wet(rb[:, 2:-2, 2:-2], ra[:, 2:-2, 2:-2], 1e-3) > 1e-7
