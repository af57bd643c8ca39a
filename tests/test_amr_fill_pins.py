"""AMR region fill: the oracle's restatement of filpatch.f90's interpolation
(coarse rectangle -> unset cells of a fine region) pinned by what the scheme
fixes (lake at rest, constant states, linear surfaces, momentum kept when a
fine cell comes out dry, hydrostatic pore pressure for water a refinement
creates), a mutation check of those pins, the AMRClaw steps around it
(copy, bounding box, boundary fill) and setaux's per-patch fields, and the
product's filpatch body (dclaw_amr.h, run on the host) against the oracle
bit for bit on random wet/dry regions."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

DRY = 1e-3
RHO_F, GRAV = 1000.0, 9.81
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


class Case:
    """A fine region (a patch of mx x my interior cells at (ilo0, jlo0) plus
    2 ghost cells) whose ghost cells are unset, and the coarse rectangle
    around them.  Fine level: dx = dxc / rx."""

    def __init__(self, rx, ry, mx=None, my=None, xlower=-3.0, ylower=1.5, dxc=2.0, dyc=1.0):
        self.rx, self.ry = rx, ry
        self.mx = mx or 3 * rx
        self.my = my or 2 * ry
        self.ilo0, self.jlo0 = 5 * rx, 4 * ry
        self.ilo, self.ihi = self.ilo0 - 2, self.ilo0 + self.mx + 1
        self.jlo, self.jhi = self.jlo0 - 2, self.jlo0 + self.my + 1
        self.xlower, self.ylower = xlower, ylower
        self.dxc, self.dyc = dxc, dyc
        self.dxf, self.dyf = dxc / rx, dyc / ry
        nx, ny = self.ihi - self.ilo + 1, self.jhi - self.jlo + 1
        self.q = np.zeros((7, nx, ny), order="F")
        self.aux = np.zeros((3, nx, ny), order="F")
        self.flag = np.zeros((nx, ny), dtype=np.uint8, order="F")
        self.flag[2:-2, 2:-2] = 1
        self.q[:, 2:-2, 2:-2] = 123.0          # interior: set, must stay untouched

    def coarse(self, O):
        box = O.amr_trimbd(self.flag, self.ilo, self.ihi, self.jlo, self.jhi)
        self.iplo, self.iphi, self.jplo, self.jphi = O.amr_coarse_box(*box, self.rx, self.ry)
        self.mxc, self.myc = self.iphi - self.iplo + 1, self.jphi - self.jplo + 1
        self.valc = np.zeros((7, self.mxc, self.myc), order="F")
        self.auxc = np.zeros((3, self.mxc, self.myc), order="F")

    def coarse_centres(self):
        x = self.xlower + (np.arange(self.iplo, self.iphi + 1) + 0.5) * self.dxc
        y = self.ylower + (np.arange(self.jplo, self.jphi + 1) + 0.5) * self.dyc
        return np.meshgrid(x, y, indexing="ij")

    def fine_centres(self):
        x = self.xlower + (np.arange(self.ilo, self.ihi + 1) + 0.5) * self.dxf
        y = self.ylower + (np.arange(self.jlo, self.jhi + 1) + 0.5) * self.dyf
        return np.meshgrid(x, y, indexing="ij")

    def run(self, O, sea_level=0.0, q=None):
        q = self.q.copy(order="F") if q is None else q
        O.amr_filpatch(self.valc, self.auxc, self.iplo, self.jplo, q, self.aux, 0, 0, self.ilo,
                       self.ihi, self.jlo, self.jhi, self.flag, self.rx, self.ry, self.xlower,
                       self.ylower, self.dxf, self.dyf, self.dxc, self.dyc, sea_level, DRY,
                       RHO_F, GRAV)
        return q

    def unset(self):
        return self.flag == 0

    def coarse_index(self):
        """Coarse (0-based) indices of every fine cell of the region."""
        gi = np.arange(self.ilo, self.ihi + 1)[:, None] + 0 * np.arange(self.jlo, self.jhi + 1)
        gj = 0 * gi + np.arange(self.jlo, self.jhi + 1)[None, :]
        return np.floor_divide(gi, self.rx) - self.iplo, np.floor_divide(gj, self.ry) - self.jplo


# --------------------------------------------------------------------- pins
def pin_filpatch_lake_at_rest(O):
    """A flat wet surface over any beds stays flat: fine depth = surface -
    fine bed, momenta 0 (filpatch.f90:274-376 with zero slopes)."""
    rng = np.random.default_rng(3)
    for rx, ry in ((2, 2), (4, 2), (3, 3)):
        C = Case(rx, ry)
        C.coarse(O)
        C.auxc[0] = rng.integers(0, 64, C.auxc[0].shape) / 64.0
        C.valc[0] = 4.0 - C.auxc[0]
        C.valc[3] = 0.5 * C.valc[0]
        C.aux[0] = rng.integers(0, 96, C.aux[0].shape) / 64.0
        q = C.run(O)
        u = C.unset()
        assert np.array_equal(q[0][u], (4.0 - C.aux[0])[u]), "lake at rest: depth"
        assert not q[1][u].any() and not q[2][u].any(), "lake at rest: momentum"
        assert (q[:, ~u] == 123.0).all(), "set cells touched"


def pin_filpatch_constant_state(O):
    """A uniform state on a flat bed comes back exactly in every unset cell."""
    vals = np.array([2.0, 0.5, -0.25, 1.0, 8000.0, 0.5, 0.125])
    for rx, ry in ((2, 2), (4, 4), (2, 4)):
        C = Case(rx, ry)
        C.coarse(O)
        C.valc[:] = vals[:, None, None]
        C.auxc[0] = 1.0
        C.aux[0] = 1.0
        q = C.run(O)
        u = C.unset()
        for m in range(7):
            assert (q[m][u] == vals[m]).all(), f"constant state: component {m}"


def pin_filpatch_linear_surface(O):
    """A linear wet surface on a flat bed is reproduced at the fine centres
    (the limiter keeps linear slopes; the offsets are the fine centres'
    positions in coarse cells), so each full fine block's mean depth is its
    coarse depth."""
    for rx, ry in ((2, 2), (4, 2), (3, 3)):
        C = Case(rx, ry)
        C.coarse(O)
        xc, yc = C.coarse_centres()
        a, b, c = 3.0, 0.125, -0.0625
        C.valc[0] = a + b * xc + c * yc
        q = C.run(O)
        xf, yf = C.fine_centres()
        u = C.unset()
        want = a + b * xf + c * yf
        assert np.abs(q[0][u] - want[u]).max() < 1e-13, "linear surface"
        ci, cj = C.coarse_index()
        for i in range(1, C.mxc - 1):
            for j in range(1, C.myc - 1):
                blk = (ci == i) & (cj == j)
                if blk.sum() == rx * ry and u[blk].all():
                    assert q[0][blk].mean() == pytest.approx(C.valc[0, i, j], rel=1e-14)


def pin_filpatch_momentum_kept_when_a_fine_cell_dries(O):
    """One dry fine cell resets its coarse cell's block to hv = h q_c /
    max(h_c, mean fine depth): with the mean above h_c the block holds rx ry
    times the coarse momentum, below it the block's momentum is q_c times
    the fine depths' sum over h_c (filpatch.f90:477-507)."""
    _momentum_case(O, 1.0, 2.0, 4.0)
    _momentum_case(O, 4.0, -1.0, 7.5 / 4.0)


def _momentum_case(O, hc, bc, factor):
    C = Case(2, 2)
    C.coarse(O)
    C.valc[0] = hc
    C.auxc[0] = bc                       # surface 3 everywhere: zero slopes
    C.valc[1] = 0.75
    C.valc[2] = -0.5
    C.aux[0] = 0.5                       # fine depth 2.5 ...
    ci, cj = C.coarse_index()
    u = C.unset()
    # ... except one fine cell of each coarse cell, whose bed is above the surface
    for i in range(1, C.mxc - 1):
        for j in range(1, C.myc - 1):
            blk = np.argwhere((ci == i) & (cj == j) & u)
            if len(blk):
                C.aux[0][tuple(blk[0])] = 3.5
    q = C.run(O)
    for i in range(1, C.mxc - 1):
        for j in range(1, C.myc - 1):
            blk = (ci == i) & (cj == j) & u
            if blk.sum() == 4:
                assert q[0][blk].sum() == 7.5
                assert q[1][blk].sum() == pytest.approx(factor * 0.75, rel=1e-14)
                assert q[2][blk].sum() == pytest.approx(factor * -0.5, rel=1e-14)


def pin_filpatch_new_water_is_hydrostatic(O):
    """Water a refinement creates under a dry coarse cell (sea level above
    the fine bed) starts at rest with pb = rho_f g h and m = 0
    (filpatch.f90:411-437: a dry coarse cell bounds pb / h at rho_f g)."""
    C = Case(2, 2)
    C.coarse(O)
    C.auxc[0] = 1.0
    C.aux[0] = 1.0
    q = C.run(O, sea_level=5.0)
    u = C.unset()
    assert (q[0][u] == 4.0).all()
    assert not q[1][u].any() and not q[2][u].any() and not q[3][u].any()
    assert np.abs(q[4][u] - RHO_F * GRAV * 4.0).max() < 1e-9


def pin_filpatch_all_dry(O):
    """Dry coarse cells below sea level give dry fine cells, every component 0."""
    C = Case(2, 2)
    C.coarse(O)
    C.auxc[0] = 1.0
    C.aux[0] = 1.0
    C.valc[1:] = 0.0
    q = C.run(O, sea_level=-100.0)
    u = C.unset()
    assert not q[:, u].any()


PINS = [pin_filpatch_lake_at_rest, pin_filpatch_constant_state, pin_filpatch_linear_surface,
        pin_filpatch_momentum_kept_when_a_fine_cell_dries, pin_filpatch_new_water_is_hydrostatic,
        pin_filpatch_all_dry]


@pytest.mark.parametrize("pin", PINS, ids=[p.__name__ for p in PINS])
def test_filpatch_pin(oracle_mod, pin):
    pin(oracle_mod)


# ----------------------------------------------------- mutation of the pins
MUTANTS = [
    ("y offset of the fine centre with -0.5",
     "const double ycent_fine = ylower + ((double)(jf - 1 + jlo) + 0.5) * dy_fine;\n"
     "        const double eta2",
     "const double ycent_fine = ylower + ((double)(jf - 1 + jlo) - 0.5) * dy_fine;\n"
     "        const double eta2"),
    ("x slope of the surface with the sign of the upwind difference",
     "dsign(1.0, CX(eta_coarse, i + 1, j) - CX(eta_coarse, i - 1, j));",
     "dsign(1.0, CX(eta_coarse, i, j) - CX(eta_coarse, i + 1, j));"),
    ("reset divides by the mean fine depth only",
     "const double divide_mass = dmax(h_coarse, h_fine_average);",
     "const double divide_mass = h_fine_average;"),
    ("dry coarse cell bounds pb / h at 0",
     "CX(vel_max, i, j) = rho_f * grav;\n                        CX(vel_min, i, j) = rho_f * grav;",
     "CX(vel_max, i, j) = 0.0;\n                        CX(vel_min, i, j) = 0.0;"),
    ("reset counts every fine cell of the block",
     "const double h_count = (double)CX(fine_cell_count, ic, jc);",
     "const double h_count = (double)(rx * ry + 1);"),
    ("fine depth not clipped at 0",
     "const double h_fine = dmax(eta_fine - BF(ifn, jf), 0.0);",
     "const double h_fine = eta_fine - BF(ifn, jf);"),
]


def _build_mutant(tmp_path, old, new):
    src = open(os.path.join(ROOT, "oracle", "dclaw_oracle.c")).read()
    assert src.count(old) == 1, f"mutation site not unique: {old!r}"
    d = tmp_path / "mut"
    d.mkdir()
    (d / "dclaw_oracle.c").write_text(src.replace(old, new))
    (d / "dclaw_oracle.h").write_text(open(os.path.join(ROOT, "oracle", "dclaw_oracle.h")).read())
    so = d / "libmut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                    "-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", str(so),
                    str(d / "dclaw_oracle.c"), "-lm"], check=True, cwd=os.path.join(ROOT, "oracle"))
    return str(so)


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_filpatch_mutant_is_rejected(oracle_mod, tmp_path, name, old, new):
    so = _build_mutant(tmp_path, old, new)
    rejected = None
    with oracle_mod.library(so):
        for pin in PINS:
            try:
                pin(oracle_mod)
            except (AssertionError, ValueError) as e:
                rejected = (pin.__name__, str(e)[:80])
                break
    assert rejected is not None, f"no pin rejects the mutant: {name}"


# ------------------------------------------------ the AMRClaw steps, setaux
def test_trimbd_and_coarse_box(oracle_mod):
    O = oracle_mod
    flag = np.ones((9, 7), dtype=np.uint8, order="F")
    assert O.amr_trimbd(flag, 10, 18, -3, 3) is None
    flag[2, 1] = 0
    flag[6, 5] = 0
    assert O.amr_trimbd(flag, 10, 18, -3, 3) == [12, 16, -2, 2]
    # one coarse cell beyond the unset cells' coarse cells on each side: fine
    # cell -2 lies in coarse cell -1 at ratio 2 and 4, fine cell 16 in 8
    assert O.amr_coarse_box(12, 16, -2, 2, 2, 2) == [5, 9, -2, 2]
    assert O.amr_coarse_box(-2, 5, 0, 7, 4, 4) == [-2, 2, -1, 2]


def test_intfil_copy_and_time_interpolation(oracle_mod):
    O = oracle_mod
    q = np.zeros((7, 8, 6), order="F")
    aux = np.zeros((2, 8, 6), order="F")
    flag = np.zeros((8, 6), dtype=np.uint8, order="F")
    sq = np.full((7, 5 + 4, 3 + 4), 2.0, order="F")
    so = np.full((7, 9, 7), 6.0, order="F")
    saux = np.full((3, 9, 7), 7.0, order="F")
    # region cells 0..7 x 10..15; source interior 5..9 x 12..14
    O.amr_intfil(q, aux, 0, 0, 0, 7, 10, 15, flag, sq, saux, 5, 12, sq_old=so, w_old=0.25)
    inside = np.zeros((8, 6), bool)
    inside[5:8, 2:5] = True
    assert (flag[inside] == 1).all() and (flag[~inside] == 0).all()
    assert (q[:, inside] == 0.75 * 2.0 + 0.25 * 6.0).all() and not q[:, ~inside].any()
    assert (aux[:, inside] == 7.0).all()


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 3, 3), (1, 3, 3, 1)])
def test_bc2amr_region(oracle_mod, bc):
    """Cells outside the domain: nearest domain cell (extrap) or mirror cell
    with the normal momentum negated (wall), corners from both rules."""
    O = oracle_mod
    rng = np.random.default_rng(5)
    iregsz, jregsz = 6, 5
    ilo, ihi, jlo, jhi = -2, 7, -2, 6           # 2 cells beyond every side
    nx, ny = ihi - ilo + 1, jhi - jlo + 1
    q = np.asfortranarray(rng.normal(size=(7, nx, ny)))
    aux = np.asfortranarray(rng.normal(size=(2, nx, ny)))
    dom = (slice(None), slice(2, 2 + iregsz), slice(2, 2 + jregsz))
    want_q = q[dom].copy()
    O.amr_bc2amr(q, aux, 0, 0, ilo, ihi, jlo, jhi, iregsz, jregsz, bc)
    for gi in range(ilo, ihi + 1):
        for gj in range(jlo, jhi + 1):
            si, sj, fu, fv = gi, gj, 1.0, 1.0
            if gi < 0:
                si, fu = (-1 - gi, -1.0) if bc[0] == 3 else (0, 1.0)
            elif gi >= iregsz:
                si, fu = (2 * iregsz - 1 - gi, -1.0) if bc[1] == 3 else (iregsz - 1, 1.0)
            if gj < 0:
                sj, fv = (-1 - gj, -1.0) if bc[2] == 3 else (0, 1.0)
            elif gj >= jregsz:
                sj, fv = (2 * jregsz - 1 - gj, -1.0) if bc[3] == 3 else (jregsz - 1, 1.0)
            got = q[:, gi - ilo, gj - jlo]
            w = want_q[:, si, sj].copy()
            w[1] *= fu
            w[2] *= fv
            assert np.array_equal(got, w), (gi, gj)


def test_setaux_fields_and_ghost_bed(oracle_mod):
    """A patch at the domain's lower-left corner: beds copied from an old
    patch stay, the others come from the topography integral, cells outside
    the domain take the nearest domain cell's bed, phi and theta everywhere,
    the other aux fields 0."""
    O = oracle_mod
    mx, my, maux = 6, 5, 7
    dx, dy = 0.5, 0.25
    xlower, ylower = 10.0, -2.0
    iregsz, jregsz = 40, 30
    aux = np.full((maux, mx + 4, my + 4), 9.0, order="F")
    bed = np.asfortranarray(np.arange((mx + 4) * (my + 4), dtype=float).reshape(mx + 4, my + 4))
    bed_set = np.zeros((mx + 4, my + 4), dtype=np.uint8, order="F")
    bed_set[4:6, 4:6] = 1
    aux[0, 4:6, 4:6] = -5.0
    O.amr_setaux(aux, 0, 0, bed_set, bed, xlower, xlower + iregsz * dx, ylower,
                 ylower + jregsz * dy, dx, dy, 0.6, 0.1)
    want = bed.copy()
    want[4:6, 4:6] = -5.0
    want[:2, :] = want[2:3, :]
    want[:, :2] = want[:, 2:3]
    assert np.array_equal(aux[0], want)
    assert (aux[1] == 0.6).all() and (aux[2] == 0.1).all() and not aux[3:].any()


# --------------------------------------- product body against the oracle
def _random_case(rng, rx, ry):
    C = Case(rx, ry, mx=4 * rx, my=3 * ry)
    # some ghost cells covered by neighbours (set), some interior cells unset
    C.flag[rng.random(C.flag.shape) < 0.1] = 1
    C.flag[2:-2, 2:-2][rng.random(C.flag[2:-2, 2:-2].shape) < 0.1] = 0
    return C


@pytest.mark.parametrize("seed", range(6))
def test_product_filpatch_body_equals_oracle(oracle_mod, emu_lib, seed):
    O = oracle_mod
    rng = np.random.default_rng(seed)
    rx, ry = [(2, 2), (4, 4), (2, 4), (3, 2), (4, 2), (2, 3)][seed]
    C = _random_case(rng, rx, ry)
    C.coarse(O)
    sh = C.valc[0].shape
    C.auxc[0] = rng.uniform(0.0, 3.0, sh)
    h = rng.uniform(0.0, 2.5, sh)
    h[rng.random(sh) < 0.25] = 0.0
    h[rng.random(sh) < 0.05] = 0.5 * DRY
    C.valc[0] = h
    for m in range(1, 7):
        C.valc[m] = rng.normal(size=sh) * h * (1.0 if m < 3 else 0.3)
    C.valc[4] = h * RHO_F * GRAV * rng.uniform(0.3, 1.2, sh)
    C.aux[0] = rng.uniform(0.0, 4.0, C.aux[0].shape)
    q0 = np.asfortranarray(rng.normal(size=C.q.shape))
    for sea in (0.0, 2.0):
        want = C.run(O, sea_level=sea, q=q0.copy(order="F"))
        got = q0.copy(order="F")
        emu_lib.emu_amr_filpatch(
            _p(C.valc), _p(C.auxc), C.auxc.shape[0], C.mxc, C.myc, C.iplo, C.jplo, _p(got),
            _p(C.aux), C.aux.shape[0], got.shape[1], 0, 0, C.ilo, C.ihi, C.jlo, C.jhi,
            C.flag.ctypes.data_as(ctypes.c_void_p), rx, ry, ctypes.c_double(C.xlower),
            ctypes.c_double(C.ylower), ctypes.c_double(C.dxf), ctypes.c_double(C.dyf),
            ctypes.c_double(C.dxc), ctypes.c_double(C.dyc), ctypes.c_double(sea),
            ctypes.c_double(DRY), ctypes.c_double(RHO_F), ctypes.c_double(GRAV))
        assert np.array_equal(got, want), f"seed {seed} sea {sea}"
        assert not np.array_equal(want[:, C.flag == 0], q0[:, C.flag == 0])


def test_library_coarse_box_equals_oracle(oracle_mod):
    """dclaw_amr_coarse_box (host function of the C-ABI library) against the
    oracle over signed unset boxes and ratios."""
    from dclaw_amd import _lib
    L = _lib.lib()
    out = (ctypes.c_int * 4)()
    for r in (1, 2, 3, 4, 6):
        for a in range(-9, 10):
            for b in (a, a + 1, a + 7):
                L.dclaw_amr_coarse_box(a, b, a - 1, b + 2, r, r + 1, out)
                assert list(out) == oracle_mod.amr_coarse_box(a, b, a - 1, b + 2, r, r + 1)
