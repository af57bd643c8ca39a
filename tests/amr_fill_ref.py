"""The oracle's composition of an AMR region fill (filpatch.f90's filrecur
with the AMRClaw steps around it) and of filval.f90's new-patch fill, over
numpy mirrors of a hierarchy: the checker of dclaw_amd.amr's device fills.
Arrays are Fortran-ordered: q (7, nx, ny), aux (maux, nx, ny)."""
import numpy as np


class RefPatch:
    def __init__(self, q, aux, ilo, jlo, t=0.0, q_old=None, t_old=None):
        self.q, self.aux, self.ilo, self.jlo, self.t = q, aux, ilo, jlo, t
        self.q_old, self.t_old = q_old, t_old

    @property
    def mx(self):
        return self.q.shape[1] - 4

    @property
    def my(self):
        return self.q.shape[2] - 4


class RefLevel:
    def __init__(self, dx, dy, iregsz, jregsz, rx=1, ry=1, patches=()):
        self.dx, self.dy, self.iregsz, self.jregsz = dx, dy, iregsz, jregsz
        self.rx, self.ry = rx, ry
        self.patches = list(patches)


def from_device(t):
    """(ny, nx, c) device tensor -> Fortran (c, nx, ny) numpy array."""
    return np.asfortranarray(t.detach().cpu().numpy().transpose(2, 1, 0))


def weight_old(p, t):
    tol = 1e-12 * max(1.0, abs(t))
    if p.q_old is None or abs(t - p.t) <= tol:
        return 0.0
    if abs(t - p.t_old) <= tol:
        return 1.0
    return (p.t - t) / (p.t - p.t_old)


class RefHierarchy:
    def __init__(self, O, levels, xlower, ylower, bc, dry_tolerance, sea_level, rho_f, grav,
                 phi, theta):
        self.O, self.levels, self.xlower, self.ylower, self.bc = O, levels, xlower, ylower, bc
        self.dry, self.sea, self.rho_f, self.grav = dry_tolerance, sea_level, rho_f, grav
        self.phi, self.theta = phi, theta

    def fill_region(self, level, box, q, aux, r0, c0, t, patch_only, own=None):
        O = self.O
        lev = self.levels[level]
        ilo, ihi, jlo, jhi = box
        gi = np.arange(ilo, ihi + 1)[:, None]
        gj = np.arange(jlo, jhi + 1)[None, :]
        inside = (gi >= 0) & (gi < lev.iregsz) & (gj >= 0) & (gj < lev.jregsz)
        flag = np.asfortranarray((~inside).astype(np.uint8))
        for p in lev.patches:
            if p is own:
                a0, a1 = max(ilo, p.ilo), min(ihi, p.ilo + p.mx - 1)
                b0, b1 = max(jlo, p.jlo), min(jhi, p.jlo + p.my - 1)
                flag[a0 - ilo:a1 - ilo + 1, b0 - jlo:b1 - jlo + 1] = 1
                continue
            w = weight_old(p, t)
            O.amr_intfil(q, aux, r0, c0, ilo, ihi, jlo, jhi, flag, p.q, p.aux, p.ilo, p.jlo,
                         sq_old=p.q_old if w != 0.0 else None, w_old=w, copy_aux=True)
        ub = O.amr_trimbd(flag, ilo, ihi, jlo, jhi)
        if ub is not None:
            assert level > 0 and own is not None
            iplo, iphi, jplo, jphi = O.amr_coarse_box(*ub, lev.rx, lev.ry)
            crs = self.levels[level - 1]
            valc = np.zeros((7, iphi - iplo + 1, jphi - jplo + 1), order="F")
            auxc = np.zeros((aux.shape[0], iphi - iplo + 1, jphi - jplo + 1), order="F")
            self.fill_region(level - 1, (iplo, iphi, jplo, jphi), valc, auxc, 0, 0, t, True)
            O.amr_filpatch(valc, auxc, iplo, jplo, q, aux, r0, c0, ilo, ihi, jlo, jhi, flag,
                           lev.rx, lev.ry, self.xlower, self.ylower, lev.dx, lev.dy, crs.dx,
                           crs.dy, self.sea, self.dry, self.rho_f, self.grav)
        if patch_only:
            O.amr_bc2amr(q, aux, r0, c0, ilo, ihi, jlo, jhi, lev.iregsz, lev.jregsz, self.bc)

    def fill_ghosts(self, level, k, t):
        p = self.levels[level].patches[k]
        box = (p.ilo - 2, p.ilo + p.mx + 1, p.jlo - 2, p.jlo + p.my + 1)
        self.fill_region(level, box, p.q, p.aux, 0, 0, t, True, own=p)

    def regrid(self, level, boxes, bed_fn, t):
        O = self.O
        lev = self.levels[level]
        rx, ry = lev.rx, lev.ry
        maux = self.levels[level - 1].patches[0].aux.shape[0]
        new = []
        for (ilo, ihi, jlo, jhi) in boxes:
            mx, my = ihi - ilo + 1, jhi - jlo + 1
            iclo, ichi = ilo // rx - 1, (ihi + 1) // rx
            jclo, jchi = jlo // ry - 1, (jhi + 1) // ry
            valc = np.zeros((7, ichi - iclo + 1, jchi - jclo + 1), order="F")
            auxc = np.zeros((maux, ichi - iclo + 1, jchi - jclo + 1), order="F")
            self.fill_region(level - 1, (iclo, ichi, jclo, jchi), valc, auxc, 0, 0, t, True)
            q = np.zeros((7, mx + 4, my + 4), order="F")
            aux = np.zeros((maux, mx + 4, my + 4), order="F")
            flag = np.zeros((mx + 4, my + 4), dtype=np.uint8, order="F")
            for p in lev.patches:
                O.amr_intfil(q, aux, 0, 0, ilo - 2, ihi + 2, jlo - 2, jhi + 2, flag, p.q, p.aux,
                             p.ilo, p.jlo)
            bed = np.asfortranarray(bed_fn(level, ilo, jlo, mx, my))
            O.amr_setaux(aux, ilo, jlo, flag, bed, self.xlower,
                         self.xlower + lev.iregsz * lev.dx, self.ylower,
                         self.ylower + lev.jregsz * lev.dy, lev.dx, lev.dy, self.phi, self.theta)
            q = O.amr_interp(valc, auxc, q, aux, flag == 0, rx, ry, self.dry, self.sea,
                             self.rho_f, self.grav)
            new.append(RefPatch(np.asfortranarray(q), aux, ilo, jlo, t))
        lev.patches = new

    def average_down(self, level):
        O = self.O
        lev, crs = self.levels[level], self.levels[level - 1]
        rx, ry = lev.rx, lev.ry
        for fp in lev.patches:
            ci0, ci1 = fp.ilo // rx, (fp.ilo + fp.mx) // rx - 1
            cj0, cj1 = fp.jlo // ry, (fp.jlo + fp.my) // ry - 1
            for cp in crs.patches:
                a0, a1 = max(ci0, cp.ilo), min(ci1, cp.ilo + cp.mx - 1)
                b0, b1 = max(cj0, cp.jlo), min(cj1, cp.jlo + cp.my - 1)
                if a0 > a1 or b0 > b1:
                    continue
                cp.q = O.amr_average(cp.q, cp.aux, fp.q, fp.aux, a0 - cp.ilo + 2, b0 - cp.jlo + 2,
                                     a1 - a0 + 1, b1 - b0 + 1, a0 * rx - fp.ilo + 2,
                                     b0 * ry - fp.jlo + 2, rx, ry, self.dry)
