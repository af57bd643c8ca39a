"""ctypes binding of the CPU oracle (oracle/dclaw_oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle the HIP path is compared against.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module; the product package ``dclaw_amd`` never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdclaw_oracle.so")
LIB_PATH_GLIBC = os.path.join(HERE, "_build", "libdclaw_oracle_glibc.so")
LIB_PATH_OMP = os.path.join(HERE, "_build", "libdclaw_oracle_omp.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OracleParams(ctypes.Structure):
    """Mirror of ``oracle_params`` in dclaw_oracle.h (same field order)."""

    _fields_ = [
        ("grav", ctypes.c_double),
        ("dry_tolerance", ctypes.c_double),
        ("speed_limit", ctypes.c_double),
        ("rho_s", ctypes.c_double),
        ("rho_f", ctypes.c_double),
        ("m_crit", ctypes.c_double),
        ("mref", ctypes.c_double),
        ("kref", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("alpha_c", ctypes.c_double),
        ("c1", ctypes.c_double),
        ("sigma_0", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("beta_seg", ctypes.c_double),
        ("delta_kr_order", ctypes.c_double),
        ("manning_coefficient", ctypes.c_double),
        ("friction_depth", ctypes.c_double),
        ("manning_max", ctypes.c_double),
        ("src2method", ctypes.c_int),
        ("alphamethod", ctypes.c_int),
        ("rarecorrectortest", ctypes.c_int),
        ("segregation", ctypes.c_int),
        ("friction_forcing", ctypes.c_int),
        ("dd_manning", ctypes.c_int),
        ("order", ctypes.c_int),
        ("transverse", ctypes.c_int),
        ("mthlim", ctypes.c_int * 5),
        ("mwaves", ctypes.c_int),
        ("entrainment_rate", ctypes.c_double),
        ("me", ctypes.c_double),
        ("chie", ctypes.c_double),
        ("entrainment", ctypes.c_int),
        ("entrainment_method", ctypes.c_int),
        ("curvature", ctypes.c_int),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_libs = {}
_variant = "dcmath"


class libm:
    """Context manager selecting the oracle build: "dcmath" (default; the
    portable elementary functions shared with the GPU) or "glibc"."""

    def __init__(self, variant: str):
        assert variant in ("dcmath", "glibc", "omp") or os.path.isabs(variant)
        self.variant = variant

    def __enter__(self):
        global _variant
        self.prev = _variant
        _variant = self.variant
        return self

    def __exit__(self, *exc):
        global _variant
        _variant = self.prev


def library(path: str) -> libm:
    """Context manager running the oracle functions of this module against
    another build of the oracle source at `path` (an absolute .so path): the
    mutation tests use it to check that the pins reject altered copies."""
    return libm(os.path.abspath(path))


def lib():
    if _variant not in _libs:
        path = {"dcmath": LIB_PATH, "glibc": LIB_PATH_GLIBC, "omp": LIB_PATH_OMP}.get(_variant,
                                                                              _variant)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(OracleParams)
        L.oracle_qfix.argtypes = [P, _dp, ctypes.c_double, _dp]
        L.oracle_setvars.argtypes = [P, _dp, _dp]
        L.oracle_riemanntype.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double,
                                                                   ctypes.c_double, _dp, _ip]
        L.oracle_riemann_solver.argtypes = [P, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_rpn2.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_rpt2.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_limiter.argtypes = [P, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_calc_taudir.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_b4step2.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_step2.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   _dp, _dp, _dp, _dp, _dp]
        L.oracle_src2.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_double]
        L.oracle_stepgrid.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                      _dp, _dp]
        L.oracle_stepgrid.restype = ctypes.c_double
        L.oracle_bc_fill.argtypes = [ctypes.c_int] * 4 + [_dp, _ip]
        L.oracle_run.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, _ip, _dp, _dp, ctypes.c_int,
                                 _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.oracle_run.restype = ctypes.c_int
        if hasattr(L, "oracle_run_patches"):
            _pp = ctypes.POINTER(_dp)
            L.oracle_run_patches.argtypes = [P, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _dp, _dp,
                                             _ip, _pp, _pp, ctypes.c_int, _dp, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int]
            L.oracle_run_patches.restype = ctypes.c_int
            L.oracle_openmp_threads.restype = ctypes.c_int
        L.oracle_dt_rule.argtypes = [ctypes.c_double] * 6 + [_dp, _dp, _ip]
        L.oracle_gauges.argtypes = [ctypes.c_int] * 4 + [_dp, _dp] + [ctypes.c_double] * 5 + [
            _dp, _dp, ctypes.c_int, _dp]
        _ci = ctypes.c_int
        L.oracle_amr_average.argtypes = [_dp, _dp, _ci, _ci, _dp, _dp, _ci, _ci] + [_ci] * 8 + [
            ctypes.c_double]
        L.oracle_amr_interp.argtypes = [_dp, _dp, _ci, _ci, _ci, _dp, _dp, _ci, _ci,
                                        ctypes.c_void_p, _ci, _ci] + [ctypes.c_double] * 4
        _cd = ctypes.c_double
        _vp = ctypes.c_void_p
        L.oracle_amr_coarse_box.argtypes = [_ci] * 6 + [_vp]
        L.oracle_amr_filpatch.argtypes = [_dp, _dp, _ci, _ci, _ci, _ci, _ci, _dp, _dp, _ci, _ci,
                                          _ci, _ci, _ci, _ci, _ci, _ci, _vp, _ci, _ci] + [_cd] * 10
        L.oracle_amr_filpatch.restype = _ci
        L.oracle_amr_intfil.argtypes = [_dp, _dp, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _vp,
                                        _dp, _vp, _cd, _dp, _ci, _ci, _ci, _ci, _ci, _ci]
        L.oracle_amr_trimbd.argtypes = [_vp, _ci, _ci, _ci, _ci, _vp]
        L.oracle_amr_trimbd.restype = _ci
        L.oracle_amr_bc2amr.argtypes = [_dp, _dp, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci, _ci,
                                        _ci, _vp]
        L.oracle_amr_setaux.argtypes = [_dp, _ci, _ci, _ci, _ci, _ci, _vp, _dp] + [_cd] * 8
        _libs[_variant] = L
    return _libs[_variant]


def params_from(dp) -> OracleParams:
    """Build the oracle's parameter struct from a ``DigParams``-like object
    (any object with the same attribute names)."""
    o = OracleParams()
    for name, ctype in OracleParams._fields_:
        if name == "mthlim":
            arr = (ctypes.c_int * 5)(*[int(x) for x in dp.mthlim])
            o.mthlim = arr
        elif ctype is ctypes.c_int:
            setattr(o, name, int(getattr(dp, name)))
        else:
            setattr(o, name, float(getattr(dp, name)))
    return o


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- point API
def qfix(params, q6, gz=None):
    q = np.ascontiguousarray(q6, dtype=np.float64).copy()
    prim = np.zeros(5)
    lib().oracle_qfix(ctypes.byref(params), _ptr(q), params.grav if gz is None else gz, _ptr(prim))
    return q, prim


def setvars(params, h, u, v, m, p, chi, gz):
    inp = np.array([h, u, v, m, p, chi, gz], dtype=np.float64)
    out = np.zeros(6)
    lib().oracle_setvars(ctypes.byref(params), _ptr(inp), _ptr(out))
    return dict(zip(["rho", "kperm", "alphainv", "m_eq", "tanpsi", "tau"], out))


def riemanntype(hL, hR, uL, uR, maxiter=1, drytol=1e-3, g=9.81):
    out = np.zeros(3)
    rare = (ctypes.c_int * 2)()
    lib().oracle_riemanntype(hL, hR, uL, uR, maxiter, drytol, g, _ptr(out),
                             ctypes.cast(rare, _ip))
    return out[0], out[1], out[2], bool(rare[0]), bool(rare[1])


SOLVER_INPUTS = ["hL", "hR", "huL", "huR", "hvL", "hvR", "hmL", "hmR", "pL", "pR", "bL", "bR",
                 "uL", "uR", "vL", "vR", "mL", "mR", "thetaL", "thetaR", "taudirL", "taudirR",
                 "chiL", "chiR", "fsL", "fsR"]


def riemann_solver(params, ixy=1, wallprob=False, **kw):
    x = np.zeros(30)
    for k, name in enumerate(SOLVER_INPUTS):
        x[k] = kw.get(name, 0.0)
    sw = np.zeros(3)
    fw = np.zeros(18)
    lib().oracle_riemann_solver(ctypes.byref(params), ixy, _ptr(x), int(wallprob), _ptr(sw),
                                _ptr(fw))
    return sw, fw.reshape(3, 6).T.copy()   # fw[m, mw] with m 0..5, mw 0..2


# ---------------------------------------------------------------- slice API
def rpn2(params, ixy, q1d, aux2, mbc=2):
    """q1d: (7, mx+2mbc) Fortran-ordered slice; aux2: (maux, mx+2mbc).
    Returns (q1d_after, fwave (7,mwaves,n), s (mwaves,n), amdq, apdq)."""
    q = np.array(q1d, dtype=np.float64, order="F", copy=True)
    aux = np.asfortranarray(aux2, dtype=np.float64)
    n = q.shape[1]
    mx = n - 2 * mbc
    mw = params.mwaves
    fwave = np.zeros((7, mw, n), order="F")
    s = np.zeros((mw, n), order="F")
    amdq = np.zeros((7, n), order="F")
    apdq = np.zeros((7, n), order="F")
    lib().oracle_rpn2(ctypes.byref(params), ixy, mx, mbc, aux.shape[0], _ptr(q), _ptr(aux),
                      _ptr(fwave), _ptr(s), _ptr(amdq), _ptr(apdq))
    return q, fwave, s, amdq, apdq


def rpt2(params, ixy, imp, q1d, aux1, aux2, aux3, asdq, mbc=2):
    q = np.asfortranarray(q1d, dtype=np.float64)
    n = q.shape[1]
    mx = n - 2 * mbc
    a1, a2, a3 = (np.asfortranarray(a, dtype=np.float64) for a in (aux1, aux2, aux3))
    asdq = np.asfortranarray(asdq, dtype=np.float64)
    bm = np.zeros((7, n), order="F")
    bp = np.zeros((7, n), order="F")
    lib().oracle_rpt2(ctypes.byref(params), ixy, imp, mx, mbc, a2.shape[0], _ptr(q), _ptr(a1),
                      _ptr(a2), _ptr(a3), _ptr(asdq), _ptr(bm), _ptr(bp))
    return bm, bp


def limiter(params, fwave, s, mbc=2):
    fw = np.array(fwave, dtype=np.float64, order="F", copy=True)
    s = np.asfortranarray(s, dtype=np.float64)
    n = fw.shape[2]
    lib().oracle_limiter(ctypes.byref(params), n - 2 * mbc, mbc, _ptr(fw), _ptr(s))
    return fw


# ---------------------------------------------------------------- patch API
def _dims(q, mbc):
    return q.shape[1] - 2 * mbc, q.shape[2] - 2 * mbc


def calc_taudir(params, q, aux, dx, dy, mbc=2):
    aux = np.array(aux, dtype=np.float64, order="F", copy=True)
    q = np.asfortranarray(q, dtype=np.float64)
    mx, my = _dims(q, mbc)
    lib().oracle_calc_taudir(ctypes.byref(params), mx, my, mbc, aux.shape[0], dx, dy, _ptr(q),
                             _ptr(aux))
    return aux


def b4step2(params, q, aux, dx, dy, mbc=2):
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    aux = np.array(aux, dtype=np.float64, order="F", copy=True)
    mx, my = _dims(q, mbc)
    lib().oracle_b4step2(ctypes.byref(params), mx, my, mbc, aux.shape[0], dx, dy, _ptr(q),
                         _ptr(aux))
    return q, aux


def step2(params, q, aux, dx, dy, dt, mbc=2):
    q = np.asfortranarray(q, dtype=np.float64)
    aux = np.asfortranarray(aux, dtype=np.float64)
    mx, my = _dims(q, mbc)
    fm, fp, gm, gp = (np.zeros_like(q, order="F") for _ in range(4))
    cfl = ctypes.c_double(0.0)
    lib().oracle_step2(ctypes.byref(params), mx, my, mbc, aux.shape[0], _ptr(q), _ptr(aux), dx,
                       dy, dt, ctypes.byref(cfl), _ptr(fm), _ptr(fp), _ptr(gm), _ptr(gp))
    return cfl.value, fm, fp, gm, gp


def src2(params, q, aux, dt, mbc=2, dx=1.0, dy=1.0):
    """src2 on the interior; dx, dy are read only with curvature on."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    aux = np.asfortranarray(aux, dtype=np.float64)
    mx, my = _dims(q, mbc)
    lib().oracle_src2(ctypes.byref(params), mx, my, mbc, aux.shape[0], dx, dy, _ptr(q),
                      _ptr(aux), dt)
    return q


def stepgrid(params, q, aux, dx, dy, dt, mbc=2):
    """One step in place semantics on copies: returns (q, aux, cflgrid)."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    aux = np.array(aux, dtype=np.float64, order="F", copy=True)
    mx, my = _dims(q, mbc)
    work = np.zeros(4 * q.size)
    cfl = lib().oracle_stepgrid(ctypes.byref(params), mx, my, mbc, aux.shape[0], dx, dy, dt,
                                _ptr(q), _ptr(aux), _ptr(work))
    return q, aux, cfl


def bc_fill(q, bc, mbc=2):
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    mx, my = _dims(q, mbc)
    bcs = (ctypes.c_int * 4)(*bc)
    lib().oracle_bc_fill(mx, my, mbc, q.shape[0], _ptr(q), ctypes.cast(bcs, _ip))
    return q


def run(params, q, aux, dx, dy, nsteps, dt0, cfl_desired, dt_max, bc=(1, 1, 1, 1), mbc=2):
    """Fixed-patch time loop.  Returns (q, aux, dt_next, cfl_hist, dt_hist)."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    aux = np.array(aux, dtype=np.float64, order="F", copy=True)
    mx, my = _dims(q, mbc)
    bcs = (ctypes.c_int * 4)(*bc)
    dt = ctypes.c_double(dt0)
    cfl_hist = np.zeros(max(nsteps, 1))
    dt_hist = np.zeros(max(nsteps, 1))
    lib().oracle_run(ctypes.byref(params), mx, my, mbc, aux.shape[0], dx, dy,
                     ctypes.cast(bcs, _ip), _ptr(q), _ptr(aux), nsteps, ctypes.byref(dt),
                     cfl_desired, dt_max, _ptr(cfl_hist), _ptr(dt_hist), 1)
    return q, aux, dt.value, cfl_hist[:nsteps], dt_hist[:nsteps]


def run_patches(params_list, qs, auxs, dxs, dys, nsteps, dt0, cfl_desired, dt_max,
                bc=(1, 1, 1, 1), nthreads=1, mbc=2):
    """nsteps of ``run`` on each of several independent patches at once, over
    nthreads OpenMP threads (the "omp" build of the library; the patches are
    spread with a dynamic schedule, as AMRClaw spreads grids).  Returns the
    final states and the next dt of each patch."""
    n = len(qs)
    qs = [np.array(q, dtype=np.float64, order="F", copy=True) for q in qs]
    auxs = [np.array(a, dtype=np.float64, order="F", copy=True) for a in auxs]
    prm = (OracleParams * n)(*params_list)
    mx = (ctypes.c_int * n)(*[q.shape[1] - 2 * mbc for q in qs])
    my = (ctypes.c_int * n)(*[q.shape[2] - 2 * mbc for q in qs])
    maux = (ctypes.c_int * n)(*[a.shape[0] for a in auxs])
    dx = np.array(dxs, dtype=np.float64)
    dy = np.array(dys, dtype=np.float64)
    dt = np.full(n, float(dt0))
    qp = (_dp * n)(*[_ptr(q) for q in qs])
    ap = (_dp * n)(*[_ptr(a) for a in auxs])
    bcs = (ctypes.c_int * 4)(*bc)
    lib().oracle_run_patches(prm, n, ctypes.cast(mx, _ip), ctypes.cast(my, _ip), mbc,
                             ctypes.cast(maux, _ip), _ptr(dx), _ptr(dy), ctypes.cast(bcs, _ip),
                             qp, ap, nsteps, _ptr(dt), cfl_desired, dt_max, int(nthreads))
    return qs, dt


def dt_rule(dt_used, cfl, t_before, cfl_desired, dt_max, tout=float("inf")):
    """(t_after, dt_next, hit) of the dt rule with the output-time hit."""
    ta, dn, hit = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    lib().oracle_dt_rule(dt_used, cfl, t_before, cfl_desired, dt_max, tout, ctypes.byref(ta),
                         ctypes.byref(dn), ctypes.byref(hit))
    return ta.value, dn.value, hit.value


def gauges(q, aux, xlower, ylower, dx, dy, dry_tolerance, xg, yg, mbc=2):
    """(ngauges, 8): h hu hv hm pb hchi bdif eta at each gauge."""
    q = np.asfortranarray(q, dtype=np.float64)
    aux = np.asfortranarray(aux, dtype=np.float64)
    mx, my = _dims(q, mbc)
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    yg = np.ascontiguousarray(yg, dtype=np.float64)
    out = np.zeros((len(xg), 8))
    lib().oracle_gauges(mx, my, mbc, aux.shape[0], _ptr(q), _ptr(aux), xlower, ylower, dx, dy,
                        dry_tolerance, _ptr(xg), _ptr(yg), len(xg), _ptr(out))
    return out


def amr_average(qc, auxc, qf, auxf, ci0, cj0, ni, nj, fi0, fj0, rx, ry, dry_tolerance):
    """update.f90's fine -> coarse average (a copy of qc with the covered
    cells replaced).  qc, qf: (7, nx, ny); auxc, auxf: (maux, nx, ny), all
    indices 0-based including ghosts."""
    qc = np.array(qc, dtype=np.float64, order="F", copy=True)
    qf = np.asfortranarray(qf, dtype=np.float64)
    auxc = np.asfortranarray(auxc, dtype=np.float64)
    auxf = np.asfortranarray(auxf, dtype=np.float64)
    assert ci0 >= 0 and cj0 >= 0 and ci0 + ni <= qc.shape[1] and cj0 + nj <= qc.shape[2]
    assert fi0 >= 0 and fj0 >= 0 and fi0 + ni * rx <= qf.shape[1] and fj0 + nj * ry <= qf.shape[2]
    lib().oracle_amr_average(_ptr(qc), _ptr(auxc), auxc.shape[0], qc.shape[1], _ptr(qf),
                             _ptr(auxf), auxf.shape[0], qf.shape[1], ci0, cj0, ni, nj, fi0, fj0,
                             rx, ry, dry_tolerance)
    return qc


def amr_interp(valc, auxc, val, auxf, needs_set, rx, ry, dry_tolerance, sea_level, rho_f, grav):
    """filval.f90's coarse -> fine interpolation loop (a copy of val with the
    fine cells of the coarse interior blocks updated).  valc: (7, mic, mjc)
    enlarged coarse patch; val: (7, mitot, mjtot) with 2 ghost cells;
    needs_set: (mitot, mjtot) bool."""
    valc = np.asfortranarray(valc, dtype=np.float64)
    auxc = np.asfortranarray(auxc, dtype=np.float64)
    val = np.array(val, dtype=np.float64, order="F", copy=True)
    auxf = np.asfortranarray(auxf, dtype=np.float64)
    ns = np.asfortranarray(needs_set, dtype=np.uint8)
    mic, mjc = valc.shape[1], valc.shape[2]
    assert (mic - 2) * rx + 4 <= val.shape[1] and (mjc - 2) * ry + 4 <= val.shape[2]
    assert ns.shape == val.shape[1:]
    lib().oracle_amr_interp(_ptr(valc), _ptr(auxc), auxc.shape[0], mic, mjc, _ptr(val), _ptr(auxf),
                            auxf.shape[0], val.shape[1], ns.ctypes.data, rx, ry, dry_tolerance,
                            sea_level, rho_f, grav)
    return val


# ------------------------------------------------------------ AMR region fill
# Regions: q (7, nx, ny) and aux (maux, nx, ny) Fortran-ordered arrays that
# hold cells ilo..ihi x jlo..jhi of a level from column r0, row c0; flag
# (ihi-ilo+1, jhi-jlo+1) uint8 Fortran-ordered.  These functions work in
# place on the arrays given (which must already be float64 / uint8 and
# Fortran-ordered).

def _fchk(a, dtype):
    assert a.dtype == dtype and a.flags.f_contiguous, "need a Fortran-ordered array"


def amr_coarse_box(isl, isr, jsb, jst, rx, ry):
    box = np.zeros(4, dtype=np.int32)
    lib().oracle_amr_coarse_box(isl, isr, jsb, jst, rx, ry, box.ctypes.data)
    return [int(v) for v in box]


def amr_filpatch(valc, auxc, iplo, jplo, q, aux, r0, c0, ilo, ihi, jlo, jhi, flag, rx, ry,
                 xlower, ylower, dxf, dyf, dxc, dyc, sea_level, dry_tolerance, rho_f, grav):
    for a in (valc, auxc, q, aux):
        _fchk(a, np.float64)
    _fchk(flag, np.uint8)
    rc = lib().oracle_amr_filpatch(_ptr(valc), _ptr(auxc), auxc.shape[0], valc.shape[1],
                                   valc.shape[2], iplo, jplo, _ptr(q), _ptr(aux), aux.shape[0],
                                   q.shape[1], r0, c0, ilo, ihi, jlo, jhi, flag.ctypes.data, rx,
                                   ry, xlower, ylower, dxf, dyf, dxc, dyc, sea_level,
                                   dry_tolerance, rho_f, grav)
    if rc != 0:
        raise ValueError("filpatch: an unset cell lies outside the coarse rectangle")


def amr_intfil(q, aux, r0, c0, ilo, ihi, jlo, jhi, flag, sq, saux, silo, sjlo, sq_old=None,
               w_old=0.0, copy_aux=True):
    """Copy from a source patch (sq: (7, smx+4, smy+4) with 2 ghost cells,
    interior cell (silo + a, sjlo + b) at column a + 2, row b + 2)."""
    for a in (q, aux, sq, saux):
        _fchk(a, np.float64)
    _fchk(flag, np.uint8)
    old = None
    if sq_old is not None:
        _fchk(sq_old, np.float64)
        old = sq_old.ctypes.data
    lib().oracle_amr_intfil(_ptr(q), _ptr(aux), aux.shape[0], q.shape[1], r0, c0, ilo, ihi, jlo,
                            jhi, flag.ctypes.data, _ptr(sq), old, w_old, _ptr(saux),
                            saux.shape[0], silo, sjlo, sq.shape[1] - 4, sq.shape[2] - 4,
                            1 if copy_aux else 0)


def amr_trimbd(flag, ilo, ihi, jlo, jhi):
    """Bounding box (isl, isr, jsb, jst) of the unset cells, or None."""
    _fchk(flag, np.uint8)
    box = np.zeros(4, dtype=np.int32)
    if lib().oracle_amr_trimbd(flag.ctypes.data, ilo, ihi, jlo, jhi, box.ctypes.data):
        return None
    return [int(v) for v in box]


def amr_bc2amr(q, aux, r0, c0, ilo, ihi, jlo, jhi, iregsz, jregsz, bc):
    for a in (q, aux):
        _fchk(a, np.float64)
    b = np.asarray(bc, dtype=np.int32)
    lib().oracle_amr_bc2amr(_ptr(q), _ptr(aux), aux.shape[0], q.shape[1], r0, c0, ilo, ihi, jlo,
                            jhi, iregsz, jregsz, b.ctypes.data)


def amr_setaux(aux, ilo, jlo, bed_set, bed, xlower, xupper, ylower, yupper, dx, dy, phi, theta):
    """setaux for a patch: aux (maux, mx+4, my+4); bed_set (mx+4, my+4)
    uint8, bed (mx+4, my+4) float64, both Fortran-ordered."""
    _fchk(aux, np.float64)
    _fchk(bed_set, np.uint8)
    _fchk(bed, np.float64)
    lib().oracle_amr_setaux(_ptr(aux), aux.shape[0], aux.shape[1] - 4, aux.shape[2] - 4, ilo, jlo,
                            bed_set.ctypes.data, _ptr(bed), xlower, xupper, ylower, yupper, dx,
                            dy, phi, theta)
