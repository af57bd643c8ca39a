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

}  // namespace dclaw
