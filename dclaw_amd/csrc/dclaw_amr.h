// dclaw_amr.h -- the two numerical transfers between refinement levels that
// the AMR driver performs around the per-patch step, as per-coarse-cell
// device functions (one thread per coarse cell; every coarse cell owns its
// rx x ry block of fine cells, so cells are independent).
//
// Layout as everywhere in this library: q(7, nx, ny) and aux(maux, nx, ny),
// component fastest, indices 0-based over the whole array including ghost
// cells; aux component 0 is the bed.  coordinate_system 1 (capacity 1).
#pragma once

#include "dclaw_physics.h"

namespace dclaw {

// update.f90:112-203 (fine -> coarse) for coarse cell (ic, jc) and its fine
// block from (if0, jf0).  The coarse depth is the mean surface of the wet
// fine cells (h > dry_tolerance; nwet of them) over the coarse bed, capped by
// their mean depth hsum / nwet, where hsum adds the depth of every fine cell,
// dry ones included; components 2..7 are the sums over the wet fine cells
// scaled by hc / hsum; an all-dry block gives zeros.
DC_HD void amr_average_cell(double* qc, const double* auxc, int mauxc, int mitc,
                            const double* qf, const double* auxf, int mauxf, int mitf, int ic,
                            int jc, int if0, int jf0, int rx, int ry, double dry_tolerance)
{
    const double bc = auxc[(size_t)mauxc * ((size_t)ic + (size_t)mitc * jc)];
    double etasum = 0.0, hsum = 0.0, hus[7];
    for (int m = 1; m < 7; m++) hus[m] = 0.0;
    int nwet = 0;
    // fine cells in the reference's order: jco outer, ico inner
    for (int jco = 0; jco < ry; jco++) {
        for (int ico = 0; ico < rx; ico++) {
            const size_t cf = (size_t)(if0 + ico) + (size_t)mitf * (size_t)(jf0 + jco);
            const double* f = qf + 7 * cf;
            const double h = f[0];
            const bool wet = h > dry_tolerance;
            if (wet) nwet++;
            hsum = hsum + h;
            for (int m = 1; m < 7; m++) hus[m] = hus[m] + (wet ? f[m] : 0.0);
            etasum = etasum + (wet ? h + auxf[(size_t)mauxf * cf] : 0.0);
        }
    }
    double* c = qc + 7 * ((size_t)ic + (size_t)mitc * jc);
    if (nwet > 0) {
        const double etaav = etasum / (double)nwet;
        const double hav = hsum / (double)nwet;
        const double hc = dmin(hav, dmax(etaav - bc, 0.0));
        const double fac = dmin(hav, hc) / hsum;
        c[0] = hc;
        for (int m = 1; m < 7; m++) c[m] = fac * hus[m];
    } else {
        for (int m = 0; m < 7; m++) c[m] = 0.0;
    }
}

// Limited (minmod) slope of the three values a, b, c: the smaller one-sided
// difference with the sign of the central one, 0 at an extremum
// (filval.f90:196-200).
DC_HD double amr_slope(double a, double b, double c)
{
    const double sp = c - b, sm = b - a;
    if (sm * sp <= 0.0) return 0.0;
    return dmin(fabs(sp), fabs(sm)) * dsign(1.0, c - a);
}

// filval.f90:190-381 (coarse -> fine, the interpolation loop) for coarse cell
// (i, j), 1 <= i <= mic-2, 1 <= j <= mjc-2, of the coarse patch valc
// enlarged by one cell all round, into the fine block from
// ((i-1) rx + 2, (j-1) ry + 2) of val (2 ghost cells).
//  * Depth, in fine cells with needs_set != 0: the surface h + b (sea_level
//    where the coarse cell is dry) at the fine centre by minmod-limited
//    slopes, minus the fine bed, clipped at 0; a set cell that comes out dry
//    (<= dry_tolerance) has its other components zeroed and marks the block.
//  * If the set depths add up to at least dry_tolerance, each component 2..7
//    (over the whole block, set or not): the limited-linear value hvf when
//    every fine hvf / h lies within [velmin, velmax] -- the range of
//    component / h over the coarse cell (0, or rho_f g for pb, when it is
//    dry) and its wet x and y neighbours -- and the block is not marked;
//    otherwise Vnew * h with Vnew = rx ry q_c / max(fine depth sum, h_c),
//    clamped to [velmin, velmax] for components 4..7.  The reference writes
//    hvf cell by cell and stops at the first value out of range, then
//    overwrites the whole block, so only the final state is produced here.
DC_HD void amr_interp_cell(const double* valc, const double* auxc, int mauxc, int mic,
                           double* val, const double* auxf, int mauxf, int mitot,
                           const unsigned char* needs_set, int rx, int ry, int i, int j,
                           double dry_tolerance, double sea_level, double rho_f, double grav)
{
    auto vc = [&](int m, int ii, int jj) -> double {
        return valc[m + 7 * ((size_t)ii + (size_t)mic * jj)];
    };
    auto eta = [&](int ii, int jj) -> double {
        const double h = vc(0, ii, jj);
        return h <= dry_tolerance ? sea_level
                                  : h + auxc[(size_t)mauxc * ((size_t)ii + (size_t)mic * jj)];
    };
    const int fi0 = (i - 1) * rx + 2, fj0 = (j - 1) * ry + 2;
    const double ec = eta(i, j);
    const double sx = amr_slope(eta(i - 1, j), ec, eta(i + 1, j));
    const double sy = amr_slope(eta(i, j - 1), ec, eta(i, j + 1));
    bool marked = false;
    double finemass = 0.0;
    for (int jco = 1; jco <= ry; jco++) {
        for (int ico = 1; ico <= rx; ico++) {
            const size_t cf = (size_t)(fi0 + ico - 1) + (size_t)mitot * (size_t)(fj0 + jco - 1);
            if (!needs_set[cf]) continue;
            const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
            const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
            double* f = val + 7 * cf;
            const double h = dmax(0.0, ((ec + xoff * sx) + yoff * sy) - auxf[(size_t)mauxf * cf]);
            f[0] = h;
            finemass = finemass + h;
            if (h <= dry_tolerance) {
                marked = true;
                for (int m = 1; m < 7; m++) f[m] = 0.0;
            }
        }
    }
    if (!(finemass >= dry_tolerance)) return;
    const double hcc = vc(0, i, j);
    for (int m = 1; m < 7; m++) {
        const double qcm = vc(m, i, j);
        const double slx = amr_slope(vc(m, i - 1, j), qcm, vc(m, i + 1, j));
        const double sly = amr_slope(vc(m, i, j - 1), qcm, vc(m, i, j + 1));
        double vmax, vmin;
        if (hcc > dry_tolerance) {
            vmax = vmin = qcm / hcc;
        } else {
            vmax = vmin = (m == 4) ? rho_f * grav : 0.0;
        }
        // neighbours in the reference's order: x-1, y-1, x+1, y+1
        const int ni[4] = {i - 1, i, i + 1, i}, nj[4] = {j, j - 1, j, j + 1};
        for (int k = 0; k < 4; k++) {
            const double hn = vc(0, ni[k], nj[k]);
            if (hn > dry_tolerance) {
                const double vel = vc(m, ni[k], nj[k]) / hn;
                vmax = dmax(vel, vmax);
                vmin = dmin(vel, vmin);
            }
        }
        bool out = marked;
        for (int ico = 1; ico <= rx && !out; ico++) {
            for (int jco = 1; jco <= ry; jco++) {
                const size_t cf =
                    (size_t)(fi0 + ico - 1) + (size_t)mitot * (size_t)(fj0 + jco - 1);
                const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
                const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
                const double vf = ((qcm + xoff * slx) + yoff * sly) / val[7 * cf];
                if (vf > vmax || vf < vmin) {
                    out = true;
                    break;
                }
            }
        }
        double vnew = 0.0;
        if (out) {
            vnew = (double)(rx * ry) * qcm / dmax(finemass, hcc);
            if (m >= 3) vnew = dmin(vmax, dmax(vmin, vnew));
        }
        for (int jco = 1; jco <= ry; jco++) {
            for (int ico = 1; ico <= rx; ico++) {
                const size_t cf =
                    (size_t)(fi0 + ico - 1) + (size_t)mitot * (size_t)(fj0 + jco - 1);
                const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
                const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
                val[m + 7 * cf] = out ? vnew * val[7 * cf] : (qcm + xoff * slx) + yoff * sly;
            }
        }
    }
}

// ---------------------------------------------------------------- filpatch
// The coarse rectangle around the unset cells isl..isr x jsb..jst of a fine
// region (level index space, refinement rx, ry; nghost = 2), with Fortran's
// truncating integer division: filpatch.f90:182-187.
DC_HD void amr_coarse_box(int isl, int isr, int jsb, int jst, int rx, int ry, int* box)
{
    box[0] = (isl - rx + 2 * rx) / rx - 2;
    box[1] = (isr + rx) / rx;
    box[2] = (jsb - ry + 2 * ry) / ry - 2;
    box[3] = (jst + ry) / ry;
}

// Geometry of one filpatch interpolation: the fine region ilo..ihi x
// jlo..jhi (level index space) stored from column r0, row c0 of the
// destination (mx columns), the coarse rectangle of mxc x myc cells from
// (iplo, jplo), the refinement ratios and the cell sizes.
struct AmrFill {
    int ilo, ihi, jlo, jhi, r0, c0, mx, maux;
    int iplo, jplo, mxc, myc, mauxc, rx, ry;
    double xlower, ylower, dxf, dyf, dxc, dyc;
    double sea_level, drytol, rho_f, grav;
};

// filpatch.f90:256-520 for one interior coarse cell (ic, jc), 1-based in the
// coarse rectangle (2 <= ic <= mxc-1, 2 <= jc <= myc-1), over the unset cells
// (flaguse == 0) of its fine block within the region.  The reference runs
// each stage over the whole region; a coarse cell's fine cells depend only on
// that coarse cell and its 4 neighbours, and its sums (fine mass and count)
// are taken in the reference's order (fine rows ascending, columns ascending
// within a row), so the result is the reference's.  Depth: the limited-linear
// surface (sea_level where the coarse cell is dry, h < dry_tolerance) at the
// fine centre minus the fine bed, clipped at 0; a fine depth below
// dry_tolerance flags the coarse cell.  Components 2..7: the limited-linear
// value where the coarse cell is not flagged and every unset fine cell's
// value / depth lies within the bounds of component / depth over the coarse
// cell and its wet neighbours (0, or rho_f g for pb, from a dry coarse cell);
// otherwise every unset fine cell gets depth * q_c / max(h_c, mean fine
// depth), clamped to the bounds for components 4..7, or 0 when the fine mass
// is at most dry_tolerance.
DC_HD void amr_filpatch_cell(const double* valc, const double* auxc, double* val,
                             const double* aux, const unsigned char* flaguse, const AmrFill& G,
                             int ic, int jc)
{
    auto vc = [&](int m, int i, int j) -> double {
        return valc[m + 7 * ((size_t)(i - 1) + (size_t)G.mxc * (size_t)(j - 1))];
    };
    auto eta = [&](int i, int j) -> double {
        const double h = vc(0, i, j);
        return h < G.drytol ? G.sea_level
                            : h + auxc[(size_t)G.mauxc * ((size_t)(i - 1) + (size_t)G.mxc * (size_t)(j - 1))];
    };
    auto slope = [](double lo, double c, double hi) -> double {
        const double down = c - lo, up = hi - c;
        return up * down > 0.0 ? dmin(fabs(up), fabs(down)) * dsign(1.0, hi - lo) : 0.0;
    };
    const int mxp = G.ihi - G.ilo + 1;
    // fine cells of this coarse cell within the region (level index space)
    const int gx = G.iplo + ic - 1, gy = G.jplo + jc - 1;
    const int ia = max(G.ilo, gx * G.rx), ib = min(G.ihi, gx * G.rx + G.rx - 1);
    const int ja = max(G.jlo, gy * G.ry), jb = min(G.jhi, gy * G.ry + G.ry - 1);
    if (ia > ib || ja > jb) return;
    const double xlow_c = G.xlower + G.iplo * G.dxc, ylow_c = G.ylower + G.jplo * G.dyc;
    const double xcent_c = xlow_c + ((double)ic - 0.5) * G.dxc;
    const double ycent_c = ylow_c + ((double)jc - 0.5) * G.dyc;
    auto fine = [&](int gi, int gj) -> size_t {
        return (size_t)(G.r0 + gi - G.ilo) + (size_t)G.mx * (size_t)(G.c0 + gj - G.jlo);
    };
    auto unset = [&](int gi, int gj) -> bool {
        return flaguse[(size_t)(gi - G.ilo) + (size_t)mxp * (size_t)(gj - G.jlo)] == 0;
    };
    auto off1 = [&](int gi) -> double {
        return ((G.xlower + ((double)gi + 0.5) * G.dxf) - xcent_c) / G.dxc;
    };
    auto off2 = [&](int gj) -> double {
        return ((G.ylower + ((double)gj + 0.5) * G.dyf) - ycent_c) / G.dyc;
    };
    const double ec = eta(ic, jc);
    const double s1 = slope(eta(ic - 1, jc), ec, eta(ic + 1, jc));
    const double s2 = slope(eta(ic, jc - 1), ec, eta(ic, jc + 1));
    double fine_mass = 0.0;
    int count = 0;
    bool flag1 = false;
    for (int gj = ja; gj <= jb; gj++) {
        const double e2 = off2(gj);
        for (int gi = ia; gi <= ib; gi++) {
            if (!unset(gi, gj)) continue;
            count++;
            const double ef = (ec + off1(gi) * s1) + e2 * s2;
            const size_t f = fine(gi, gj);
            const double hf = dmax(ef - aux[(size_t)G.maux * f], 0.0);
            val[7 * f] = hf;
            fine_mass = fine_mass + hf;
            if (hf < G.drytol) flag1 = true;
        }
    }
    if (count == 0) return;
    const double hc = vc(0, ic, jc);
    for (int m = 1; m < 7; m++) {
        const double qc = vc(m, ic, jc);
        const double t1 = slope(vc(m, ic - 1, jc), qc, vc(m, ic + 1, jc));
        const double t2 = slope(vc(m, ic, jc - 1), qc, vc(m, ic, jc + 1));
        double vmax, vmin;
        if (hc > G.drytol) {
            vmax = qc / hc;
            vmin = qc / hc;
        } else {
            vmax = vmin = (m == 4) ? G.rho_f * G.grav : 0.0;
        }
        // neighbours in the reference's order: x-1, y-1, x+1, y+1
        const int ni[4] = {ic - 1, ic, ic + 1, ic}, nj[4] = {jc, jc - 1, jc, jc + 1};
        for (int k = 0; k < 4; k++) {
            const double hn = vc(0, ni[k], nj[k]);
            if (hn > G.drytol) {
                vmax = dmax(vmax, vc(m, ni[k], nj[k]) / hn);
                vmin = dmin(vmin, vc(m, ni[k], nj[k]) / hn);
            }
        }
        bool reset = flag1;
        for (int gj = ja; gj <= jb && !reset; gj++) {
            const double e2 = off2(gj);
            for (int gi = ia; gi <= ib; gi++) {
                if (!unset(gi, gj)) continue;
                const size_t f = fine(gi, gj);
                const double hv = (qc + off1(gi) * t1) + e2 * t2;
                const double v = hv / val[7 * f];
                if (v < vmin || v > vmax) {
                    reset = true;
                    break;
                }
            }
        }
        double vnew = 0.0;
        if (reset && fine_mass > G.drytol) {
            vnew = qc / dmax(hc, fine_mass / (double)count);
            if (m >= 3) {
                vnew = dmax(vmin, vnew);
                vnew = dmin(vmax, vnew);
            }
        }
        for (int gj = ja; gj <= jb; gj++) {
            const double e2 = off2(gj);
            for (int gi = ia; gi <= ib; gi++) {
                if (!unset(gi, gj)) continue;
                const size_t f = fine(gi, gj);
                if (!reset) val[m + 7 * f] = (qc + off1(gi) * t1) + e2 * t2;
                else if (fine_mass > G.drytol) val[m + 7 * f] = vnew * val[7 * f];
                else val[m + 7 * f] = 0.0;
            }
        }
    }
}

// ---------------------------------------------- region fill (the AMRClaw steps)
DC_HD bool amr_in_domain(int gi, int gj, int iregsz, int jregsz)
{
    return gi >= 0 && gi < iregsz && gj >= 0 && gj < jregsz;
}

// Source cell (level index space) and momentum signs of a cell outside the
// domain under the bc2amr rules (extrap: nearest domain cell; wall: mirror
// cell, normal momentum negated), x then y as the patch ghost fill does.
DC_HD void amr_domain_source(const int* bc, int iregsz, int jregsz, int gi, int gj, int& si,
                             int& sj, double& fu, double& fv)
{
    si = gi; sj = gj; fu = 1.0; fv = 1.0;
    if (gi < 0) {
        if (bc[0] == DCLAW_BC_WALL) { si = -1 - gi; fu = -1.0; } else si = 0;
    } else if (gi >= iregsz) {
        if (bc[1] == DCLAW_BC_WALL) { si = 2 * iregsz - 1 - gi; fu = -1.0; } else si = iregsz - 1;
    }
    if (gj < 0) {
        if (bc[2] == DCLAW_BC_WALL) { sj = -1 - gj; fv = -1.0; } else sj = 0;
    } else if (gj >= jregsz) {
        if (bc[3] == DCLAW_BC_WALL) { sj = 2 * jregsz - 1 - gj; fv = -1.0; } else sj = jregsz - 1;
    }
}

// Storage index (cell, not component) of region cell (gi, gj).
DC_HD size_t amr_region_cell(const dclaw_amr_region& R, int gi, int gj)
{
    return (size_t)(R.r0 + gi - R.ilo) + (size_t)R.mx * (size_t)(R.c0 + gj - R.jlo);
}
DC_HD size_t amr_region_flag(const dclaw_amr_region& R, int gi, int gj)
{
    return (size_t)(gi - R.ilo) + (size_t)(R.ihi - R.ilo + 1) * (size_t)(gj - R.jlo);
}

// Same-level copy for region cell (gi, gj), inside the source's interior.
DC_HD void amr_copy_cell(const dclaw_amr_region& R, const double* sq, const double* sq_old,
                         double w_old, const double* saux, int smaux, int silo, int sjlo,
                         int smx, bool copy_aux, int gi, int gj)
{
    const size_t d = amr_region_cell(R, gi, gj);
    R.flag[amr_region_flag(R, gi, gj)] = 1;
    if (sq == R.q) return;
    const size_t s = (size_t)(gi - silo + 2) + (size_t)(smx + 4) * (size_t)(gj - sjlo + 2);
    for (int m = 0; m < 7; m++) {
        const double v = sq[7 * s + m];
        R.q[7 * d + m] = (sq_old && w_old != 0.0) ? v * (1.0 - w_old) + sq_old[7 * s + m] * w_old
                                                  : v;
    }
    if (copy_aux) {
        const int na = R.maux < smaux ? R.maux : smaux;
        for (int k = 0; k < na; k++) R.aux[(size_t)R.maux * d + k] = saux[(size_t)smaux * s + k];
    }
}

// bc2amr for region cell (gi, gj) outside the domain.
DC_HD void amr_region_bc_cell(const dclaw_amr_region& R, const int* bc, int iregsz, int jregsz,
                              int gi, int gj)
{
    int si, sj;
    double fu, fv;
    amr_domain_source(bc, iregsz, jregsz, gi, gj, si, sj, fu, fv);
    const size_t d = amr_region_cell(R, gi, gj), s = amr_region_cell(R, si, sj);
    for (int m = 0; m < 7; m++) R.q[7 * d + m] = R.q[7 * s + m];
    R.q[7 * d + 1] = fu * R.q[7 * d + 1];
    R.q[7 * d + 2] = fv * R.q[7 * d + 2];
    for (int k = 0; k < R.maux; k++) R.aux[(size_t)R.maux * d + k] = R.aux[(size_t)R.maux * s + k];
}

}  // namespace dclaw
