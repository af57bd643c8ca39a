"""Device transfers between AMR levels (the numerical part of the coupling an
AMR driver performs around the per-patch step): fine -> coarse averaging
(update.f90:112-203) and coarse -> fine interpolation (filval.f90:190-381).

Tensors are device float64 in the library's layout: q as (ny, nx, 7) and aux
as (ny, nx, maux) -- the bytes of Fortran q(7, nx, ny) -- and needs_set as a
(ny, nx) uint8 tensor.  Indices are 0-based over the whole arrays (ghost
cells included)."""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _chk(t: torch.Tensor, dtype, ndim: int, name: str) -> None:
    if not (t.is_cuda and t.dtype == dtype and t.dim() == ndim and t.is_contiguous()):
        raise _lib.DClawError(f"{name} must be a contiguous {ndim}-d {dtype} device tensor")


def average_to_coarse(qc: torch.Tensor, auxc: torch.Tensor, qf: torch.Tensor,
                      auxf: torch.Tensor, ci0: int, cj0: int, ni: int, nj: int, fi0: int,
                      fj0: int, rx: int, ry: int, dry_tolerance: float) -> None:
    """Overwrite coarse cells (ci0 + a, cj0 + b), a < ni, b < nj, of qc with
    the wet-surface average of their rx x ry fine blocks of qf starting at
    (fi0 + a rx, fj0 + b ry)."""
    for t, n in ((qc, "qc"), (auxc, "auxc"), (qf, "qf"), (auxf, "auxf")):
        _chk(t, torch.float64, 3, n)
    if qc.shape[2] != 7 or qf.shape[2] != 7:
        raise _lib.DClawError("q tensors must have 7 components")
    _lib.check(_lib.lib().dclaw_amr_average_dev(
        qc.data_ptr(), auxc.data_ptr(), auxc.shape[2], qc.shape[1], qc.shape[0], qf.data_ptr(),
        auxf.data_ptr(), auxf.shape[2], qf.shape[1], qf.shape[0], ci0, cj0, ni, nj, fi0, fj0,
        rx, ry, float(dry_tolerance), torch.cuda.current_stream().cuda_stream))


def interpolate_to_fine(valc: torch.Tensor, auxc: torch.Tensor, val: torch.Tensor,
                        auxf: torch.Tensor, needs_set: torch.Tensor, rx: int, ry: int,
                        dry_tolerance: float, sea_level: float, rho_f: float,
                        grav: float) -> None:
    """Fill the fine patch val (2 ghost cells) from the coarse patch valc
    enlarged by one coarse cell all round: depth where needs_set, the other
    components over every interior coarse cell's fine block."""
    for t, n in ((valc, "valc"), (auxc, "auxc"), (val, "val"), (auxf, "auxf")):
        _chk(t, torch.float64, 3, n)
    _chk(needs_set, torch.uint8, 2, "needs_set")
    if tuple(needs_set.shape) != tuple(val.shape[:2]):
        raise _lib.DClawError("needs_set must have the fine patch's (ny, nx) shape")
    _lib.check(_lib.lib().dclaw_amr_interp_dev(
        valc.data_ptr(), auxc.data_ptr(), auxc.shape[2], valc.shape[1], valc.shape[0],
        val.data_ptr(), auxf.data_ptr(), auxf.shape[2], val.shape[1], val.shape[0],
        needs_set.data_ptr(), rx, ry, float(dry_tolerance), float(sea_level), float(rho_f),
        float(grav), torch.cuda.current_stream().cuda_stream))


# ----------------------------------------------------------------- region fill
# Device mirrors of the patches of an AMR hierarchy and the fills an AMR
# driver performs with them: a patch's ghost cells from its neighbours and
# the coarser level (filpatch.f90's filrecur), new patches after a regrid
# (filval.f90 with setaux.f90), and the fine -> coarse average (update.f90).
# Index boxes are in each level's index space (cell i covers [xlower + i dx,
# xlower + (i+1) dx)); patches carry 2 ghost cells.

class AmrPatch:
    """One patch's device arrays: q (my+4, mx+4, 7) and aux (my+4, mx+4,
    maux) float64, interior cells ilo..ilo+mx-1 x jlo..jlo+my-1.  q_old and
    t_old, when set, hold the patch's state at its previous time (used to
    interpolate in time when a finer level fills from it between the two)."""

    def __init__(self, q: torch.Tensor, aux: torch.Tensor, ilo: int, jlo: int, t: float = 0.0):
        _chk(q, torch.float64, 3, "q")
        _chk(aux, torch.float64, 3, "aux")
        if q.shape[2] != 7 or tuple(aux.shape[:2]) != tuple(q.shape[:2]):
            raise _lib.DClawError("q must be (my+4, mx+4, 7) and aux (my+4, mx+4, maux)")
        self.q, self.aux, self.ilo, self.jlo, self.t = q, aux, int(ilo), int(jlo), float(t)
        self.q_old = None
        self.t_old = None

    @property
    def mx(self) -> int:
        return self.q.shape[1] - 4

    @property
    def my(self) -> int:
        return self.q.shape[0] - 4

    @property
    def box(self):
        return self.ilo, self.ilo + self.mx - 1, self.jlo, self.jlo + self.my - 1


class AmrLevel:
    """The patches of one level, its cell size, its domain size in cells and
    its refinement ratio (rx, ry) over the next coarser level."""

    def __init__(self, dx: float, dy: float, iregsz: int, jregsz: int, rx: int = 1, ry: int = 1):
        self.dx, self.dy, self.iregsz, self.jregsz = float(dx), float(dy), int(iregsz), int(jregsz)
        self.rx, self.ry = int(rx), int(ry)
        self.patches: list = []


class AmrHierarchy:
    """Levels 0 (covers the domain) .. L-1, domain origin (xlower, ylower),
    boundary codes bc (xlo, xhi, ylo, yhi: 1 extrap, 3 wall) and the
    constants the fills read (geoclaw_module / digclaw_module values)."""

    def __init__(self, levels, xlower: float, ylower: float, bc=(1, 1, 1, 1),
                 dry_tolerance: float = 1e-3, sea_level: float = 0.0, rho_f: float = 1000.0,
                 grav: float = 9.81, phi: float = 0.0, theta: float = 0.0):
        self.levels = list(levels)
        self.xlower, self.ylower = float(xlower), float(ylower)
        self.bc = (ctypes.c_int * 4)(*bc)
        self.dry_tolerance, self.sea_level = float(dry_tolerance), float(sea_level)
        self.rho_f, self.grav, self.phi, self.theta = float(rho_f), float(grav), float(phi), float(theta)

    # -- one region -------------------------------------------------------
    def _region(self, q, aux, flag, r0, c0, box):
        ilo, ihi, jlo, jhi = box
        return _lib.CAmrRegion(q.data_ptr(), aux.data_ptr(), flag.data_ptr(), q.shape[1],
                               q.shape[0], aux.shape[2], r0, c0, ilo, ihi, jlo, jhi)

    @staticmethod
    def _weight_old(p: AmrPatch, t: float) -> float:
        """w_old of a source patch at time t: 0 at its current time (or with
        no previous state), 1 at its previous time, linear in between."""
        tol = 1e-12 * max(1.0, abs(t))
        if p.q_old is None or abs(t - p.t) <= tol:
            return 0.0
        if abs(t - p.t_old) <= tol:
            return 1.0
        w = (p.t - t) / (p.t - p.t_old)
        if not 0.0 <= w <= 1.0:
            raise _lib.DClawError(f"time {t} outside the source patch's [{p.t_old}, {p.t}]")
        return w

    def fill_region(self, level: int, box, q: torch.Tensor, aux: torch.Tensor, r0: int, c0: int,
                    t: float, patch_only: bool, own: AmrPatch = None) -> None:
        """filrecur: fill cells box = (ilo, ihi, jlo, jhi) of `level`, held
        from (r0, c0) of q/aux, from the level's patches (own: the patch whose
        arrays these are, whose interior is already in place), what remains
        from the coarser level by filpatch's interpolation, and, with
        patch_only, the cells outside the domain by bc2amr."""
        L = _lib.lib()
        st = torch.cuda.current_stream().cuda_stream
        lev = self.levels[level]
        ilo, ihi, jlo, jhi = box
        flag = torch.empty((jhi - jlo + 1, ihi - ilo + 1), dtype=torch.uint8, device=q.device)
        R = self._region(q, aux, flag, r0, c0, box)
        rp = ctypes.byref(R)
        _lib.check(L.dclaw_amr_region_init_dev(rp, lev.iregsz, lev.jregsz, st))
        for p in lev.patches:
            if p is own:
                _lib.check(L.dclaw_amr_copy_dev(rp, q.data_ptr(), None, 0.0, None, 0, p.ilo, p.jlo,
                                                p.mx, p.my, 0, st))
                continue
            w = self._weight_old(p, t)
            old = p.q_old.data_ptr() if (w != 0.0) else None
            _lib.check(L.dclaw_amr_copy_dev(rp, p.q.data_ptr(), old, w, p.aux.data_ptr(),
                                            p.aux.shape[2], p.ilo, p.jlo, p.mx, p.my, 1, st))
        ub = torch.empty(4, dtype=torch.int32, device=q.device)
        _lib.check(L.dclaw_amr_unset_box_dev(rp, ub.data_ptr(), st))
        isl, isr, jsb, jst = ub.tolist()
        if isl <= isr:
            if level == 0:
                raise _lib.DClawError("level 0 does not cover the region (filrecur: level 1 not set)")
            if own is None:
                raise _lib.DClawError(
                    f"level {level} does not cover the coarse rectangle {box}: the finer level "
                    "is not properly nested")
            cb = (ctypes.c_int * 4)()
            L.dclaw_amr_coarse_box(isl, isr, jsb, jst, lev.rx, lev.ry, cb)
            iplo, iphi, jplo, jphi = list(cb)
            crs = self.levels[level - 1]
            valc = torch.zeros((jphi - jplo + 1, iphi - iplo + 1, 7), dtype=torch.float64,
                               device=q.device)
            auxc = torch.zeros((jphi - jplo + 1, iphi - iplo + 1, aux.shape[2]),
                               dtype=torch.float64, device=q.device)
            self.fill_region(level - 1, (iplo, iphi, jplo, jphi), valc, auxc, 0, 0, t, True)
            _lib.check(L.dclaw_amr_filpatch_dev(
                rp, valc.data_ptr(), auxc.data_ptr(), auxc.shape[2], iplo, iphi, jplo, jphi,
                lev.rx, lev.ry, self.xlower, self.ylower, lev.dx, lev.dy, crs.dx, crs.dy,
                self.sea_level, self.dry_tolerance, self.rho_f, self.grav, st))
        if patch_only:
            _lib.check(L.dclaw_amr_region_bc_dev(rp, lev.iregsz, lev.jregsz, self.bc, st))

    def fill_ghosts(self, level: int, k: int, t: float) -> None:
        """AMRClaw bound for patch k of `level` at time t: its 2 ghost cells
        from the level's other patches, the coarser level and the boundary
        conditions."""
        p = self.levels[level].patches[k]
        ilo, ihi, jlo, jhi = p.box
        self.fill_region(level, (ilo - 2, ihi + 2, jlo - 2, jhi + 2), p.q, p.aux, 0, 0, t, True,
                         own=p)

    # -- regrid -----------------------------------------------------------
    def regrid(self, level: int, boxes, bed_fn, t: float) -> None:
        """Replace the patches of `level` (>= 1) by new ones over `boxes`
        (ilo, ihi, jlo, jhi), aligned to the coarser level's cells, filled as
        filval.f90 does: the coarse rectangle from the coarser level, the
        cells the old patches of this level cover copied from them (q and the
        bed), setaux (bed_fn(level, ilo, jlo, mx, my) -> (my+4, mx+4) float64
        device tensor: the cell averages of the topography, ghost cells
        included; used where nothing was copied), then the interpolation of
        the remaining interior cells."""
        L = _lib.lib()
        st = torch.cuda.current_stream().cuda_stream
        lev = self.levels[level]
        if level < 1:
            raise _lib.DClawError("level 0 is not regridded")
        rx, ry = lev.rx, lev.ry
        old = lev.patches
        maux = old[0].aux.shape[2] if old else self.levels[level - 1].patches[0].aux.shape[2]
        dev = self.levels[level - 1].patches[0].q.device
        new = []
        for (ilo, ihi, jlo, jhi) in boxes:
            if ilo % rx or (ihi + 1) % rx or jlo % ry or (jhi + 1) % ry:
                raise _lib.DClawError(f"box {(ilo, ihi, jlo, jhi)} is not aligned to coarse cells")
            if ilo < 0 or jlo < 0 or ihi >= lev.iregsz or jhi >= lev.jregsz:
                raise _lib.DClawError(f"box {(ilo, ihi, jlo, jhi)} outside the domain")
            mx, my = ihi - ilo + 1, jhi - jlo + 1
            # filval.f90:84-92: the coarse rectangle, one coarse cell all round
            iclo, ichi = ilo // rx - 1, (ihi + 1) // rx
            jclo, jchi = jlo // ry - 1, (jhi + 1) // ry
            valc = torch.zeros((jchi - jclo + 1, ichi - iclo + 1, 7), dtype=torch.float64,
                               device=dev)
            auxc = torch.zeros((jchi - jclo + 1, ichi - iclo + 1, maux), dtype=torch.float64,
                               device=dev)
            self.fill_region(level - 1, (iclo, ichi, jclo, jchi), valc, auxc, 0, 0, t, True)
            q = torch.zeros((my + 4, mx + 4, 7), dtype=torch.float64, device=dev)
            aux = torch.zeros((my + 4, mx + 4, maux), dtype=torch.float64, device=dev)
            flag = torch.zeros((my + 4, mx + 4), dtype=torch.uint8, device=dev)
            R = self._region(q, aux, flag, 0, 0, (ilo - 2, ihi + 2, jlo - 2, jhi + 2))
            for p in old:
                _lib.check(L.dclaw_amr_copy_dev(ctypes.byref(R), p.q.data_ptr(), None, 0.0,
                                                p.aux.data_ptr(), p.aux.shape[2], p.ilo, p.jlo,
                                                p.mx, p.my, 1, st))
            bed = bed_fn(level, ilo, jlo, mx, my)
            _chk(bed, torch.float64, 2, "bed")
            if tuple(bed.shape) != (my + 4, mx + 4):
                raise _lib.DClawError("bed_fn must return (my+4, mx+4)")
            _lib.check(L.dclaw_amr_setaux_dev(aux.data_ptr(), maux, mx, my, ilo, jlo,
                                              flag.data_ptr(), bed.data_ptr(), lev.iregsz,
                                              lev.jregsz, self.phi, self.theta, st))
            needs = (flag == 0).to(torch.uint8)
            interpolate_to_fine(valc, auxc, q, aux, needs, rx, ry, self.dry_tolerance,
                                self.sea_level, self.rho_f, self.grav)
            new.append(AmrPatch(q, aux, ilo, jlo, t))
        lev.patches = new

    # -- fine -> coarse ---------------------------------------------------
    def average_down(self, level: int) -> None:
        """update.f90: every coarse cell of level-1 covered by a patch of
        `level` takes the average of its fine block."""
        lev, crs = self.levels[level], self.levels[level - 1]
        rx, ry = lev.rx, lev.ry
        for fp in lev.patches:
            fi0, fi1, fj0, fj1 = fp.box
            ci0, ci1 = fi0 // rx, (fi1 + 1) // rx - 1
            cj0, cj1 = fj0 // ry, (fj1 + 1) // ry - 1
            for cp in crs.patches:
                a0, a1 = max(ci0, cp.ilo), min(ci1, cp.ilo + cp.mx - 1)
                b0, b1 = max(cj0, cp.jlo), min(cj1, cp.jlo + cp.my - 1)
                if a0 > a1 or b0 > b1:
                    continue
                average_to_coarse(cp.q, cp.aux, fp.q, fp.aux, a0 - cp.ilo + 2, b0 - cp.jlo + 2,
                                  a1 - a0 + 1, b1 - b0 + 1, a0 * rx - fp.ilo + 2,
                                  b0 * ry - fp.jlo + 2, rx, ry, self.dry_tolerance)
