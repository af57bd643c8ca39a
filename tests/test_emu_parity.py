"""The product's strip kernel body, run on the CPU by the lane emulator,
against the oracle: bit-for-bit equality of q, aux(4..6) and cflgrid.

Both sides use the same elementary functions (dcmath.h) and IEEE double
arithmetic without FMA contraction, so any difference is a logic error in
the kernel (wrong neighbour, row lag, accumulation order, ownership)."""
import ctypes

import numpy as np
import pytest

from dclaw_amd import workloads as W
from helpers import Emu, advance_oracle, slope_patch


@pytest.fixture(scope="module")
def emu(emu_lib):
    return Emu(emu_lib)


def _check(O, emu, pt, q, aux, dt, seg_rows=32):
    qe, auxe, cfle = emu.stepgrid(pt.params, q, aux, pt.dx, pt.dy, dt, seg_rows=seg_rows)
    qo, auxo, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    # NaN (src2method 2's undefined branch) must sit in the same places
    assert np.array_equal(qe, qo, equal_nan=True), f"{(qe != qo).sum()} values differ"
    assert np.array_equal(auxe, auxo)
    assert cfle == cflo


def test_c0_bitwise_after_motion(oracle_mod, emu):
    pt = W.inclined_plane_mound()
    q, aux, dt = advance_oracle(oracle_mod, pt, 30)
    _check(oracle_mod, emu, pt, q, aux, dt)


@pytest.mark.parametrize("mx,my,seg,bc", [
    (130, 70, 8, (3, 3, 3, 3)),     # 3 strips, 9 segments, walls
    (59, 33, 4, (1, 3, 3, 1)),      # strip width + 1, short segments
    (1, 1, 32, (1, 1, 1, 1)),       # one cell
    (7, 90, 16, (3, 1, 1, 3)),
    (116, 5, 32, (1, 1, 1, 1)),     # exactly two strips
])
def test_ragged_and_walls(oracle_mod, emu, mx, my, seg, bc):
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx * my)
    q, aux, dt = advance_oracle(oracle_mod, pt, 6, bc=bc)
    _check(oracle_mod, emu, pt, q, aux, dt, seg_rows=seg)


@pytest.mark.parametrize("variant", [
    dict(order=1, transverse=0), dict(order=2, transverse=0), dict(order=2, transverse=1),
    dict(order=1, transverse=2), dict(mthlim=[1, 2, 3, 4, 5]), dict(mthlim=[0, 0, 0, 0, 0]),
    dict(src2method=-1), dict(src2method=1), dict(alphamethod=1), dict(rarecorrectortest=1),
    dict(segregation=1, delta_kr_order=1.5), dict(friction_forcing=1, manning_coefficient=0.06),
    dict(friction_forcing=1, manning_coefficient=0.06, dd_manning=1), dict(speed_limit=0.5),
    dict(src2method=2, alphamethod=1), dict(src2method=2, alphamethod=1, kref=1e-8),
])
def test_parameter_variants(oracle_mod, emu, variant):
    pt = slope_patch(70, 40, slope=0.9, phi_deg=22.0, seed=11)
    for k, v in variant.items():
        setattr(pt.params, k, v)
    if variant.get("segregation"):
        pt.q[5] = 0.4 * pt.q[0]
    if variant.get("src2method") == 2:
        # reach a flowing state with src2method 0 (src2method 2 leaves a
        # still wet cell undefined), then compare one step of each
        pt.params.src2method = 0
        q, aux, dt = advance_oracle(oracle_mod, pt, 5)
        pt.params.src2method = 2
    else:
        q, aux, dt = advance_oracle(oracle_mod, pt, 5)
    _check(oracle_mod, emu, pt, q, aux, dt, seg_rows=16)


def test_dry_patch_and_fast_path(oracle_mod, emu):
    pt = slope_patch(80, 40, h0=0.0)
    q = oracle_mod.bc_fill(pt.q, (1, 1, 1, 1))
    q[6] = np.abs(np.random.default_rng(2).standard_normal(q[6].shape))   # bdif passes through
    _check(oracle_mod, emu, pt, q, pt.aux, 0.01)


def test_chained_steps_bitwise(oracle_mod, emu):
    """Five chained steps (bc fill, step, dt rule) stay bit-identical."""
    O = oracle_mod
    pt = slope_patch(64, 64, slope=0.6, phi_deg=28.0, seed=3)
    P = O.params_from(pt.params)
    qo, auxo, dto, cflo, _ = O.run(P, pt.q, pt.aux, 1.0, 1.0, 5, 1e-3, 0.9, 1e9)
    q, aux, dt = pt.q, pt.aux, 1e-3
    for n in range(5):
        q = O.bc_fill(q, (1, 1, 1, 1))
        q, aux, c = emu.stepgrid(pt.params, q, aux, 1.0, 1.0, dt, seg_rows=8)
        assert c == cflo[n]
        dt = min(dt * 0.9 / c, 1e9)
    assert np.array_equal(q[:, 2:-2, 2:-2], qo[:, 2:-2, 2:-2])


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 3, 3), (3, 1, 1, 3)])
def test_bc_on_load_equals_fill(oracle_mod, emu, bc):
    """Boundary conditions applied while loading give the same step as filling
    the ghost cells first (stale ghosts in the input must not matter)."""
    O = oracle_mod
    pt = slope_patch(61, 37, slope=0.7, phi_deg=24.0, seed=9)
    q, aux, dt = advance_oracle(O, pt, 4, bc=bc)
    stale = q.copy(order="F")
    stale[:, :2, :] = 7.0
    stale[:, :, -2:] = -3.0
    qe, auxe, ce = emu.stepgrid(pt.params, stale, aux, 1.0, 1.0, dt, seg_rows=8, bc=bc)
    qo, auxo, co = O.stepgrid(O.params_from(pt.params), q, aux, 1.0, 1.0, dt)
    assert np.array_equal(qe, qo) and np.array_equal(auxe, auxo) and ce == co


def test_zero_dry_tolerance(oracle_mod, emu):
    """dry_tolerance = 0: cells with h = 0 exactly are the only dry ones, and
    rpn2 evaluates setvars on a dry side after resetting it in place."""
    pt = slope_patch(64, 36, slope=0.9, phi_deg=22.0, seed=5)
    thin = pt.q[0] < 0.05
    for m in range(6):
        pt.q[m][thin] = 0.0
    pt.params.dry_tolerance = 0.0
    q, aux, dt = advance_oracle(oracle_mod, pt, 4)
    assert (q[0][2:-2, 2:-2] == 0.0).any() and (q[0] > 0.0).any()
    _check(oracle_mod, emu, pt, q, aux, dt, seg_rows=8)


def _emu_step2(lib, params, q, aux, dx, dy, dt, seg):
    from dclaw_amd.params import to_c
    dp = ctypes.POINTER(ctypes.c_double)
    q = np.asfortranarray(q)
    aux = np.asfortranarray(aux)
    f = [np.zeros_like(q, order="F") for _ in range(4)]
    cfl = ctypes.c_double(0.0)
    cp = to_c(params)
    lib.emu_step2(ctypes.byref(cp), q.ctypes.data_as(dp), aux.ctypes.data_as(dp), aux.shape[0],
                  q.shape[1] - 4, q.shape[2] - 4, ctypes.c_double(dx), ctypes.c_double(dy),
                  ctypes.c_double(dt), seg, *[a.ctypes.data_as(dp) for a in f], ctypes.byref(cfl))
    return cfl.value, f


@pytest.mark.parametrize("mx,my,seg", [(70, 40, 16), (130, 33, 8), (59, 61, 4), (7, 9, 32),
                                       (1, 1, 32)])
def test_step2_flux_mode_bitwise(oracle_mod, emu_lib, mx, my, seg):
    """The strip body in step2 mode: every entry of fm, fp, gm, gp (border rows
    and columns included) and cflgrid equal the oracle's step2 bit for bit."""
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.9, phi_deg=22.0, seed=mx)
    q, aux, dt = advance_oracle(O, pt, 5)
    P = O.params_from(pt.params)
    qb, auxb = O.b4step2(P, q, aux, pt.dx, pt.dy)
    cflo, *fo = O.step2(P, qb, auxb, pt.dx, pt.dy, dt)
    cfle, fe = _emu_step2(emu_lib, pt.params, qb, auxb, pt.dx, pt.dy, dt, seg)
    assert cfle == cflo
    for name, a, b in zip(("fm", "fp", "gm", "gp"), fe, fo):
        assert np.array_equal(a, b), f"{name}: {(a != b).sum()} entries differ"
    if mx * my > 1:      # a single cell with extrapolated ghosts has no flux
        assert np.count_nonzero(fo[0]) > 0


@pytest.mark.parametrize("mx,my,seg,bc", [(130, 70, 8, (3, 3, 3, 3)), (59, 33, 4, (1, 3, 3, 1)),
                                          (116, 40, 16, (1, 1, 1, 1))])
def test_split_src2_bitwise(oracle_mod, emu_lib, mx, my, seg, bc):
    """The strip body without src2 plus the src2 pass over the flagged tiles
    equals the oracle's step (q and cflgrid) bit for bit."""
    from dclaw_amd.params import to_c
    O = oracle_mod
    dp = ctypes.POINTER(ctypes.c_double)
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + my)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc)
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    qi = np.array(q, order="F", copy=True)
    qe = np.zeros_like(qi, order="F")
    aux = np.asfortranarray(aux)
    cp = to_c(pt.params)
    cfl = ctypes.c_double(0.0)
    nfl = ctypes.c_int(0)
    rc = emu_lib.emu_stepgrid_split(ctypes.byref(cp), qi.ctypes.data_as(dp), qe.ctypes.data_as(dp),
                                    aux.ctypes.data_as(dp), aux.shape[0], mx, my,
                                    ctypes.c_double(pt.dx), ctypes.c_double(pt.dy),
                                    ctypes.c_double(dt), seg, (ctypes.c_int * 4)(*bc),
                                    ctypes.byref(cfl), ctypes.byref(nfl))
    assert rc == 0
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert np.array_equal(qe[I], qo[I]), f"{(qe[I] != qo[I]).sum()} values differ"
    assert cfl.value == cflo
    ntiles = -(-mx // 58) * -(-my // seg)
    assert 0 < nfl.value <= ntiles


@pytest.mark.parametrize("bc", [None, (1, 1, 1, 1), (3, 3, 3, 3)])
def test_dry_tiles_next_to_wet_cells(oracle_mod, emu, bc):
    """Wet cells two rows below a segment and two columns left of a strip:
    those tiles take the whole-segment dry path (their 3x3 neighbourhoods are
    dry) and must still equal the oracle, over several steps."""
    O = oracle_mod
    mx, my, seg = 130, 40, 8
    pt = slope_patch(mx, my, slope=0.6, phi_deg=25.0, seed=1, h0=0.0)
    # block of wet cells: rows 18..19 (segment [9,16] ends at 16), columns
    # 50..57 (the second strip starts at column 59); 0-based + 2 ghosts
    for (i0, i1, j0, j1) in [(50, 57, 18, 19), (1, 3, 1, 2), (128, 130, 39, 40)]:
        sl = (slice(i0 + 1, i1 + 2), slice(j0 + 1, j1 + 2))
        pt.q[0][sl] = 2.0
        pt.q[3][sl] = 1.2
        pt.q[4][sl] = 0.8 * 1000 * 9.81 * 2.0
    q = O.bc_fill(pt.q, bc or (1, 1, 1, 1))
    aux = pt.aux
    dt = 0.01
    for step in range(4):
        qe, auxe, cfle = emu.stepgrid(pt.params, q, aux, pt.dx, pt.dy, dt, seg_rows=seg, bc=bc)
        qo, auxo, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
        assert np.array_equal(qe, qo), f"step {step}: {(qe != qo).sum()} values differ"
        assert np.array_equal(auxe, auxo), f"step {step}: aux differs"
        assert cfle == cflo
        q, aux = O.bc_fill(qo, bc or (1, 1, 1, 1)), auxo


def _entrainment_patch(O):
    from dclaw_amd import workloads as W
    pt = slope_patch(90, 50, slope=0.9, phi_deg=22.0, seed=12)
    rng = np.random.default_rng(3)
    h_e = rng.uniform(0.0, 0.6, (90, 50))
    pe = W.make_patch(np.zeros((90, 50)), np.zeros((90, 50)), 0.6, 0.8, pt.params, 1.0, 1.0,
                      maux=7, h_e=h_e)
    pt.aux = np.concatenate([pt.aux, pe.aux[6:7]], axis=0)
    pt.params.entrainment, pt.params.entrainment_method = 1, 0
    pt.params.entrainment_rate = 0.3
    return pt


def test_entrainment_fused_and_split_bitwise(oracle_mod, emu, emu_lib):
    """Entrainment (method 0) through the fused strip body and the split
    step, against the oracle, over several steps."""
    from dclaw_amd.params import to_c
    O = oracle_mod
    pt = _entrainment_patch(O)
    q, aux, dt = advance_oracle(O, pt, 6)
    assert (q[6] > 0).any()                       # the bed has been eroded
    _check(O, emu, pt, q, aux, dt, seg_rows=8)
    # split
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    dp = ctypes.POINTER(ctypes.c_double)
    qi = np.array(q, order="F", copy=True)
    qe = np.zeros_like(qi, order="F")
    auxf = np.asfortranarray(aux)
    cp = to_c(pt.params)
    cfl, nfl = ctypes.c_double(0.0), ctypes.c_int(0)
    assert emu_lib.emu_stepgrid_split(ctypes.byref(cp), qi.ctypes.data_as(dp),
                                      qe.ctypes.data_as(dp), auxf.ctypes.data_as(dp),
                                      auxf.shape[0], 90, 50, ctypes.c_double(pt.dx),
                                      ctypes.c_double(pt.dy), ctypes.c_double(dt), 16, None,
                                      ctypes.byref(cfl), ctypes.byref(nfl)) == 0
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert np.array_equal(qe[I], qo[I])
    assert cfl.value == cflo


def _emu_split(lib, params, q, aux, dx, dy, dt, seg, bc, group=None):
    """The split step by the emulator: single strip launch (group None) or
    with the dry pass over groups of `group` segments.  q_out starts as NaN,
    so a cell the step does not write shows."""
    from dclaw_amd.params import to_c
    dp = ctypes.POINTER(ctypes.c_double)
    qi = np.array(q, order="F", copy=True)
    qe = np.full_like(qi, np.nan, order="F")
    auxf = np.array(aux, order="F", copy=True)
    mx, my = qi.shape[1] - 4, qi.shape[2] - 4
    cp = to_c(params)
    cfl, n = ctypes.c_double(0.0), ctypes.c_int(0)
    bca = None if bc is None else (ctypes.c_int * 4)(*bc)
    args = (ctypes.byref(cp), qi.ctypes.data_as(dp), qe.ctypes.data_as(dp),
            auxf.ctypes.data_as(dp), auxf.shape[0], mx, my, ctypes.c_double(dx),
            ctypes.c_double(dy), ctypes.c_double(dt), seg)
    if group is None:
        rc = lib.emu_stepgrid_split(*args, bca, ctypes.byref(cfl), ctypes.byref(n))
    else:
        rc = lib.emu_stepgrid_classified(*args, group, bca, ctypes.byref(cfl), ctypes.byref(n))
    assert rc == 0, f"emulator returned {rc}"
    return qe, cfl.value, n.value


@pytest.mark.parametrize("mx,my,seg,group,bc", [
    (130, 70, 8, 4, (3, 3, 3, 3)),
    (59, 33, 4, 5, (1, 3, 3, 1)),
    (116, 40, 16, 3, (1, 1, 1, 1)),
    (7, 90, 3, 11, (3, 1, 1, 3)),
    (1, 1, 32, 1, (1, 1, 1, 1)),
    (60, 200, 6, 8, None),
    (180, 64, 1, 32, (1, 1, 3, 3)),
])
def test_dry_pass_bitwise(oracle_mod, emu_lib, mx, my, seg, group, bc):
    """The split step with the dry pass (classify, strip body over the listed
    segments, src2 over the flagged ones) equals the single-launch split step
    on every cell, ghosts included, and the oracle's step on the interior."""
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + 3 * my)
    q, aux, dt = advance_oracle(O, pt, 6, bc=bc or (1, 1, 1, 1))
    qs, cfls, _ = _emu_split(emu_lib, pt.params, q, aux, pt.dx, pt.dy, dt, seg, bc)
    qc, cflc, nwet = _emu_split(emu_lib, pt.params, q, aux, pt.dx, pt.dy, dt, seg, bc, group)
    assert not np.isnan(qc).any(), f"{np.isnan(qc).sum()} values not written"
    assert np.array_equal(qc, qs), f"{(qc != qs).sum()} values differ from the split step"
    assert cflc == cfls
    qo, _, cflo = O.stepgrid(O.params_from(pt.params), q, aux, pt.dx, pt.dy, dt)
    I = (slice(None), slice(2, -2), slice(2, -2))
    assert np.array_equal(qc[I], qo[I])
    assert cflc == cflo
    ntiles = -(-mx // 58) * -(-my // seg)
    assert 0 < nwet <= ntiles
    if mx * my >= 60 * 200:
        assert nwet < ntiles          # some segments take the dry pass


@pytest.mark.parametrize("bc,group", [(None, 1), ((1, 1, 1, 1), 2), ((3, 3, 3, 3), 5),
                                      ((1, 3, 1, 3), 3)])
def test_dry_pass_edges_of_wet_blocks(oracle_mod, emu_lib, bc, group):
    """Wet blocks two rows below a segment, two columns left of a strip and in
    the patch corners, over several steps: the dry pass must list exactly the
    segments whose 1-cell neighbourhood holds a wet cell (a missed one shows
    as a difference, a spurious one costs only time)."""
    O = oracle_mod
    mx, my, seg = 130, 40, 8
    pt = slope_patch(mx, my, slope=0.6, phi_deg=25.0, seed=1, h0=0.0)
    for (i0, i1, j0, j1) in [(50, 57, 18, 19), (1, 3, 1, 2), (128, 130, 39, 40),
                             (100, 101, 25, 25)]:
        sl = (slice(i0 + 1, i1 + 2), slice(j0 + 1, j1 + 2))
        pt.q[0][sl] = 2.0
        pt.q[3][sl] = 1.2
        pt.q[4][sl] = 0.8 * 1000 * 9.81 * 2.0
    q = O.bc_fill(pt.q, bc or (1, 1, 1, 1))
    aux = pt.aux
    dt = 0.01
    for step in range(4):
        qs, cfls, _ = _emu_split(emu_lib, pt.params, q, aux, pt.dx, pt.dy, dt, seg, bc)
        qc, cflc, nwet = _emu_split(emu_lib, pt.params, q, aux, pt.dx, pt.dy, dt, seg, bc, group)
        assert not np.isnan(qc).any()
        assert np.array_equal(qc, qs), f"step {step}: {(qc != qs).sum()} values differ"
        assert cflc == cfls
        assert 0 < nwet < 3 * 5
        q = O.bc_fill(qc, bc or (1, 1, 1, 1))


def _offset_like(a, off):
    """A Fortran-ordered copy of `a` whose data starts `off` doubles into a
    16-byte aligned buffer (off = 1: 8-byte aligned only)."""
    buf = np.zeros(a.size + 2, dtype=np.float64)
    shift = (buf.ctypes.data // 8) % 2          # make buf + shift 16-byte aligned
    v = buf[shift + off:shift + off + a.size].reshape(a.shape, order="F")
    v[...] = a
    assert (v.ctypes.data % 16 == 0) == (off == 0)
    return v


@pytest.mark.parametrize("in_off,out_off", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_dry_scan_alignment(oracle_mod, emu_lib, in_off, out_off):
    """The dry paths read and write rows with 16-byte accesses aligned to
    q_in; q_in or q_out only 8-byte aligned (and of different parity) must
    give the same result, fused and with the dry pass."""
    from dclaw_amd.params import to_c
    O = oracle_mod
    dp = ctypes.POINTER(ctypes.c_double)
    mx, my, seg = 130, 40, 8
    pt = slope_patch(mx, my, slope=0.6, phi_deg=25.0, seed=1, h0=0.0)
    for (i0, i1, j0, j1) in [(50, 57, 18, 19), (1, 3, 1, 2), (128, 130, 39, 40)]:
        sl = (slice(i0 + 1, i1 + 2), slice(j0 + 1, j1 + 2))
        pt.q[0][sl] = 2.0
        pt.q[3][sl] = 1.2
        pt.q[4][sl] = 0.8 * 1000 * 9.81 * 2.0
    pt.q[6] = np.random.default_rng(5).uniform(-0.5, 0.5, pt.q[6].shape)   # bdif of both signs
    q = O.bc_fill(pt.q, (1, 1, 1, 1))
    aux = np.asfortranarray(pt.aux)
    dt = 0.01
    cp = to_c(pt.params)
    for bc in (None, (3, 1, 1, 3)):
        bca = None if bc is None else (ctypes.c_int * 4)(*bc)
        ref = None
        for group in (None, 3):
            qi = _offset_like(np.asfortranarray(q), in_off)
            qe = _offset_like(np.full(q.shape, np.nan, order="F"), out_off)
            cfl, n = ctypes.c_double(0.0), ctypes.c_int(0)
            args = (ctypes.byref(cp), qi.ctypes.data_as(dp), qe.ctypes.data_as(dp),
                    aux.ctypes.data_as(dp), aux.shape[0], mx, my, ctypes.c_double(pt.dx),
                    ctypes.c_double(pt.dy), ctypes.c_double(dt), seg)
            if group is None:
                rc = emu_lib.emu_stepgrid_split(*args, bca, ctypes.byref(cfl), ctypes.byref(n))
            else:
                rc = emu_lib.emu_stepgrid_classified(*args, group, bca, ctypes.byref(cfl),
                                                     ctypes.byref(n))
            assert rc == 0
            assert not np.isnan(qe).any()
            qa, _, _ = _emu_split(emu_lib, pt.params, q, aux, pt.dx, pt.dy, dt, seg, bc, group)
            assert np.array_equal(np.asarray(qe), qa), f"bc {bc} group {group}"
            if ref is None:
                ref = qa
            assert np.array_equal(qa, ref)


@pytest.mark.parametrize("mx,my,seg,bc", [
    (130, 70, 8, (3, 3, 3, 3)), (200, 90, 3, (1, 1, 1, 1)), (59, 33, 4, (1, 3, 3, 1)),
    (175, 61, 5, None), (120, 40, 2, (1, 1, 1, 1)),
])
def test_tile_flag_bands(oracle_mod, emu_lib, mx, my, seg, bc):
    """The split step's tile flag is 0 when no owned cell needs src2, else
    TF_ANY plus one bit per band (first / last 3 columns of the strip, first
    / last 3 rows of the segment) holding such a cell: checked against the
    wet mask of the strip pass's own output (before src2) on every tile."""
    from dclaw_amd.params import to_c
    O = oracle_mod
    pt = slope_patch(mx, my, slope=0.8, phi_deg=25.0, seed=mx + 3 * my)
    q, aux, dt = advance_oracle(O, pt, 8, bc=bc or (1, 1, 1, 1))
    dp = ctypes.POINTER(ctypes.c_double)
    qi = np.array(q, order="F", copy=True)
    qe = np.full_like(qi, np.nan, order="F")
    auxf = np.array(aux, order="F", copy=True)
    nstrips, nseg = -(-mx // 58), -(-my // seg)
    flags = np.full(nstrips * nseg, -7, dtype=np.int32)
    cp = to_c(pt.params)
    bca = None if bc is None else (ctypes.c_int * 4)(*bc)
    rc = emu_lib.emu_strip_flags(ctypes.byref(cp), qi.ctypes.data_as(dp), qe.ctypes.data_as(dp),
                                 auxf.ctypes.data_as(dp), auxf.shape[0], mx, my,
                                 ctypes.c_double(pt.dx), ctypes.c_double(pt.dy),
                                 ctypes.c_double(dt), seg, bca,
                                 flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0
    wet = qe[0][2:-2, 2:-2] > pt.params.dry_tolerance      # [i - 1, j - 1]
    assert wet.any() and not wet.all()
    nbands = 0
    for g in range(nseg):
        for s in range(nstrips):
            i0, j0 = s * 58, g * seg
            w = wet[i0:min(mx, i0 + 58), j0:min(my, j0 + seg)]
            want = 0
            if w.any():
                want = 1
                want |= 2 if w[:3].any() else 0
                want |= 4 if w[55:58].any() else 0
                want |= 8 if w[:, :3].any() else 0
                want |= 16 if w[:, -3:].any() else 0
                nbands += want != 31
            assert flags[g * nstrips + s] == want, (s, g, flags[g * nstrips + s], want)
    assert nbands > 0
