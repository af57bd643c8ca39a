// dclaw_strip.h -- fused stepgrid for one strip segment of one patch.
//
// One wavefront of 64 lanes owns a strip of 58 interior columns
// [i0, i0+57] (lane l holds column c = i0 - 3 + l, so the strip carries a
// 3-column halo on each side) and marches down the rows of a segment
// [jA, jB], keeping a rolling window of 4 rows in registers.  At loop index
// R (the row being loaded) it:
//   S1  loads row R and applies b4step2's pointwise part (qfix, bdif >= 0,
//       speed limit: b4step2.f90:56-100), then the qfix chain rpn2 applies
//       to its slice (F1 = qfix(Q0), F2 = qfix(F1): rpn2_dclaw.f:164-168);
//   S2  calc_taudir for row R-1 (digclaw_module.f90:500-719);
//   S3  solves the y-edge between rows R-2 and R-1 (rpn2 with ixy = 2);
//   S4  runs the whole x-sweep of row R-2 (flux2fw.f with ixy = 1: rpn2,
//       limiter, second-order correction, transverse rpt2 split);
//   S5  limits the y-edge between rows R-3 and R-2 and splits it
//       transversally (flux2fw.f with ixy = 2);
//   S6  assembles fm, fp, gm, gp for row R-3 in GeoClaw step2's summation
//       order, updates it (stepgrid) and applies src2 (src2.f90:92-303).
// Neighbouring columns are reached with lane shifts (prev: value of lane l-1,
// next: of lane l+1; the end lanes keep their own); no LDS tile is held.
// The code is a template on the lane policy so the same body runs on the GPU
// (one wavefront) and in a host emulator (64 threads in lockstep).
//
// With FLUX = true the same body is GeoClaw's step2 instead of stepgrid: the
// input is already b4step2's output (no qfix / speed limit on load), fsphi
// and taudir are read from aux(4..6) instead of recomputed, and S6 stores
// fm, fp, gm, gp (every entry the reference's sweeps accumulate, including
// the border rows and columns no update reads) instead of updating q.
#pragma once

#include "dclaw_physics.h"

namespace dclaw {

// Modes of the strip body: the fused step (stepgrid), the same without src2
// (a separate pass applies it to the tiles flagged wet), and step2 alone.
constexpr int MODE_STEP = 0;
constexpr int MODE_STEP_NOSRC2 = 1;
constexpr int MODE_FLUX = 2;

constexpr int STRIP_LANES = 64;
constexpr int DRY_CHUNK = 8;       // rows per load batch of the whole-segment dry check
constexpr int CLASSIFY_CHUNK = 4;  // rows per load batch of the split step's dry pass
constexpr int STRIP_HALO = 3;
constexpr int STRIP_W = STRIP_LANES - 2 * STRIP_HALO;  // 58 interior columns

struct StripArgs {
    const double* q_in;
    double* q_out;
    const double* aux;
    double* aux_out;       // if non-null: write aux(4..6) as calc_taudir does (maux >= 6)
    int maux, mx, my;
    double dx, dy, dt;
    int i0, jA, jB;        // interior strip start column, segment rows [jA, jB]
    int bc[4];             // bc codes (xlo, xhi, ylo, yhi) applied on load; 0 = use stored ghosts
    double *fm, *fp, *gm, *gp;   // step2 outputs (FLUX mode; same shape as q, zero-filled)
    int* tile_flag;              // MODE_STEP_NOSRC2: set to 1 if src2 has wet cells here, else 0
    bool known_wet = false;      // skip the whole-segment dry check (strip_classify found the
                                 // segment not dry and already wrote the dry state)
};

// Source cell and momentum signs of a ghost cell under the bc2amr rules
// (extrap: nearest interior cell; wall: mirror, normal momentum negated).
// Identical to what k_bc_fill writes, so loading through this map equals
// filling the ghosts first.
DC_HD void bc_source(const int* bc, int mx, int my, int i, int j, int& si, int& sj, double& fu,
                     double& fv)
{
    si = i; sj = j; fu = 1.0; fv = 1.0;
    if (i < 1) {
        if (bc[0] == DCLAW_BC_WALL) { si = 1 - i; fu = -1.0; } else si = 1;
    } else if (i > mx) {
        if (bc[1] == DCLAW_BC_WALL) { si = 2 * mx + 1 - i; fu = -1.0; } else si = mx;
    }
    if (j < 1) {
        if (bc[2] == DCLAW_BC_WALL) { sj = 1 - j; fv = -1.0; } else sj = 1;
    } else if (j > my) {
        if (bc[3] == DCLAW_BC_WALL) { sj = 2 * my + 1 - j; fv = -1.0; } else sj = my;
    }
}

// Row state of one column in the rolling window.
struct RowCell {
    double h, hu, hv, bdif, b;       // Q0 (b4step2 output); b = aux(1)
    double hm0, p0, hchi0;           // Q0 remaining components
    double hm1, p1, hchi1, m1, chi1; // F1 = qfix(Q0)
    double u, v;                     // hu/h, hv/h (same for F1 and F2)
    double hm2, p2, hchi2, m2, chi2; // F2 = qfix(F1)
    double rho1, tau1;               // setvars of F1 (wet cells; set when the row is R-1)
};

DC_HD void zero_row(RowCell& r)
{
    r.h = r.hu = r.hv = r.bdif = r.b = 0.0;
    r.hm0 = r.p0 = r.hchi0 = 0.0;
    r.hm1 = r.p1 = r.hchi1 = r.m1 = r.chi1 = 0.0;
    r.u = r.v = 0.0;
    r.hm2 = r.p2 = r.hchi2 = r.m2 = r.chi2 = 0.0;
    r.rho1 = r.tau1 = 0.0;
}

// b4step2 pointwise part on a loaded cell, then F1 and F2.
DC_HD void prepare_cell(const dclaw_params& P, const double* q7, double b, RowCell& r)
{
    double h = q7[0], hu = q7[1], hv = q7[2], hm = q7[3], p = q7[4], hchi = q7[5];
    double u, v, m, chi, rho;
    const double gz = P.grav;
    qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, gz);
    r.bdif = dmax(q7[6], 0.0);
    if (h > 0.0) {
        const double s = sqrt((hu * hu + hv * hv)) / h;
        if (s > P.speed_limit) {
            const double sratio = P.speed_limit / s;
            hu = hu * sratio;
            hv = hv * sratio;
        }
    }
    r.h = h; r.hu = hu; r.hv = hv; r.b = b;
    r.hm0 = hm; r.p0 = p; r.hchi0 = hchi;
    // F1
    double h1 = h, hu1 = hu, hv1 = hv, hm1 = hm, p1 = p, hchi1 = hchi;
    qfix(P, h1, hu1, hv1, hm1, p1, hchi1, u, v, m, chi, rho, gz);
    r.hm1 = hm1; r.p1 = p1; r.hchi1 = hchi1; r.m1 = m; r.chi1 = chi; r.u = u; r.v = v;
    // F2
    double hm2 = hm1, p2 = p1, hchi2 = hchi1;
    qfix(P, h1, hu1, hv1, hm2, p2, hchi2, u, v, m, chi, rho, gz);
    r.hm2 = hm2; r.p2 = p2; r.hchi2 = hchi2; r.m2 = m; r.chi2 = chi;
}

// step2 input: the cell is already b4step2's output (Q0 as stored), then F1
// and F2 as rpn2 applies them.
DC_HD void prepare_cell_q0(const dclaw_params& P, const double* q7, double b, RowCell& r)
{
    double u, v, m, chi, rho;
    const double gz = P.grav;
    r.h = q7[0]; r.hu = q7[1]; r.hv = q7[2]; r.b = b;
    r.hm0 = q7[3]; r.p0 = q7[4]; r.hchi0 = q7[5];
    r.bdif = q7[6];
    double h1 = q7[0], hu1 = q7[1], hv1 = q7[2], hm1 = q7[3], p1 = q7[4], hchi1 = q7[5];
    qfix(P, h1, hu1, hv1, hm1, p1, hchi1, u, v, m, chi, rho, gz);
    // rpn2 reads h, hu, hv after its qfix (a dry cell comes back zeroed)
    r.h = h1; r.hu = hu1; r.hv = hv1;
    r.hm1 = hm1; r.p1 = p1; r.hchi1 = hchi1; r.m1 = m; r.chi1 = chi; r.u = u; r.v = v;
    double hm2 = hm1, p2 = p1, hchi2 = hchi1;
    qfix(P, h1, hu1, hv1, hm2, p2, hchi2, u, v, m, chi, rho, gz);
    r.hm2 = hm2; r.p2 = p2; r.hchi2 = hchi2; r.m2 = m; r.chi2 = chi;
}

// rpn2 side state from a row cell: F2 (left side, when the cell is not the
// slice's first cell) or F1 (right side, or the slice's first cell); sv is
// setvars of that state.
DC_HD Side side_of(const RowCell& r, bool second, double taudir, double fs, const RhoTau& sv)
{
    Side s;
    s.h = r.h; s.hu = r.hu; s.hv = r.hv;
    s.hm = second ? r.hm2 : r.hm1;
    s.p = second ? r.p2 : r.p1;
    s.hchi = second ? r.hchi2 : r.hchi1;
    s.m = second ? r.m2 : r.m1;
    s.chi = second ? r.chi2 : r.chi1;
    s.u = r.u;
    s.v = r.v;
    s.b = r.b - r.bdif;
    s.taudir = taudir;
    s.fs = fs;
    s.rho = sv.rho;
    s.tau = sv.tau;
    return s;
}

// A whole wave record from the next lane.
template <class Lane>
DC_HD EdgeWaves next_waves(Lane& L, const EdgeWaves& w)
{
    EdgeWaves o;
    for (int k = 0; k < 3; k++) o.s[k] = L.next(w.s[k]);
    for (int k = 0; k < 6; k++) o.w1[k] = L.next(w.w1[k]);
    for (int k = 0; k < 6; k++) o.w5[k] = L.next(w.w5[k]);
    o.w2h = L.next(w.w2h);
    o.w2n = L.next(w.w2n);
    return o;
}

// Limiter sums of inlinelimiter.f over the 7 components in q-index order
// (structural zeros dropped).  Family 0 = wave 1, 1 = wave 2, 2 = wave 5.
DC_HD void wave_dots(const EdgeWaves& a, const EdgeWaves& b, int ixy, double* dot3)
{
    double d = 0.0;
    for (int k = 0; k < 6; k++) d = d + a.w1[k] * b.w1[k];
    dot3[0] = d;
    // wave 2 components in q-index order: h, then normal (hu for x, hv for y)
    d = 0.0;
    d = d + a.w2h * b.w2h;
    d = d + a.w2n * b.w2n;
    dot3[1] = d;
    d = 0.0;
    for (int k = 0; k < 6; k++) d = d + a.w5[k] * b.w5[k];
    dot3[2] = d;
    (void)ixy;
}

// Second-order correction and fluctuations of one edge (flux2fw.f:149-243):
// amdq/apdq from the unlimited waves (rpn2_dclaw.f:347-379), limited waves
// (inlinelimiter.f) with dotl/dotr against unlimited neighbours, cqxx with
// the wet-dry gate, fadd = (0 +- a{m|p}dq) + cqxx/2, and the cqxx fold into
// amdq/apdq when transverse = 2.  Vectors in q-index order (6 components).
struct EdgeFlux {
    double faddm[6], faddp[6], amdq[6], apdq[6];
};

DC_HD EdgeFlux edge_flux(const dclaw_params& P, int ixy, const EdgeWaves& w, const double* dotl,
                         const double* dotr, bool gate_wet, double dtdx)
{
    EdgeFlux o;
    const int nrm = (ixy == 1) ? 1 : 2;
    double amdq[6], apdq[6];
    for (int k = 0; k < 6; k++) { amdq[k] = 0.0; apdq[k] = 0.0; }
    // rpn2 fluctuations, waves in order 1, 2 (3, 4 are zero), 5
    if (w.s[0] < 0.0) { for (int k = 0; k < 6; k++) amdq[k] = amdq[k] + w.w1[k]; }
    else if (w.s[0] > 0.0) { for (int k = 0; k < 6; k++) apdq[k] = apdq[k] + w.w1[k]; }
    if (w.s[1] < 0.0) { amdq[0] = amdq[0] + w.w2h; amdq[nrm] = amdq[nrm] + w.w2n; }
    else if (w.s[1] > 0.0) { apdq[0] = apdq[0] + w.w2h; apdq[nrm] = apdq[nrm] + w.w2n; }
    if (w.s[2] < 0.0) { for (int k = 0; k < 6; k++) amdq[k] = amdq[k] + w.w5[k]; }
    else if (w.s[2] > 0.0) { for (int k = 0; k < 6; k++) apdq[k] = apdq[k] + w.w5[k]; }

    double cq[6];
    for (int k = 0; k < 6; k++) cq[k] = 0.0;
    if (P.order != 1) {
        // limited waves
        double phi[3] = {1.0, 1.0, 1.0};
        const int fam_mw[3] = {0, 1, 4};
        double wn2[3];
        {
            double d = 0.0;
            for (int k = 0; k < 6; k++) d = d + w.w1[k] * w.w1[k];
            wn2[0] = d;
            d = 0.0;
            d = d + w.w2h * w.w2h;
            d = d + w.w2n * w.w2n;
            wn2[1] = d;
            d = 0.0;
            for (int k = 0; k < 6; k++) d = d + w.w5[k] * w.w5[k];
            wn2[2] = d;
        }
        for (int f = 0; f < 3; f++) {
            const int lim = P.mthlim[fam_mw[f]];
            if (lim == 0 || wn2[f] == 0.0) continue;
            double r;
            if (w.s[f] > 0.0) r = dotl[f] / wn2[f];
            else r = dotr[f] / wn2[f];
            phi[f] = limiter_phi(lim, r);
        }
        if (gate_wet) {
            const double c0 = dsign(1.0, w.s[0]) * (1.0 - fabs(w.s[0]) * dtdx);
            const double c1 = dsign(1.0, w.s[1]) * (1.0 - fabs(w.s[1]) * dtdx);
            const double c2 = dsign(1.0, w.s[2]) * (1.0 - fabs(w.s[2]) * dtdx);
            // waves limited only when the family's limiter applied
            double l1[6], l5[6];
            const bool a0 = (P.mthlim[0] != 0 && wn2[0] != 0.0);
            const bool a1 = (P.mthlim[1] != 0 && wn2[1] != 0.0);
            const bool a2 = (P.mthlim[4] != 0 && wn2[2] != 0.0);
            for (int k = 0; k < 6; k++) {
                l1[k] = a0 ? phi[0] * w.w1[k] : w.w1[k];
                l5[k] = a2 ? phi[2] * w.w5[k] : w.w5[k];
            }
            const double l2h = a1 ? phi[1] * w.w2h : w.w2h;
            const double l2n = a1 ? phi[1] * w.w2n : w.w2n;
            for (int k = 0; k < 6; k++) {
                double c = 0.0;
                c = c + c0 * l1[k];
                if (k == 0) c = c + c1 * l2h;
                if (k == nrm) c = c + c1 * l2n;
                c = c + c2 * l5[k];
                cq[k] = c;
            }
        }
    }
    for (int k = 0; k < 6; k++) {
        o.faddm[k] = (0.0 + amdq[k]) + 0.5 * cq[k];
        o.faddp[k] = (0.0 - apdq[k]) + 0.5 * cq[k];
    }
    if (P.order > 1 && P.transverse == 2) {
        for (int k = 0; k < 6; k++) {
            o.amdq[k] = amdq[k] + cq[k];
            o.apdq[k] = apdq[k] - cq[k];
        }
    } else {
        for (int k = 0; k < 6; k++) { o.amdq[k] = amdq[k]; o.apdq[k] = apdq[k]; }
    }
    return o;
}

// G(k, side) of one cell: (0 - g_imp1) - g_imp2, g = (0.5*dtdx)*b{m|p}asdq.
DC_HD void trans_combine(const double* bm1, const double* bp1, const double* bm2,
                         const double* bp2, double dtdx, double* G1, double* G2)
{
    for (int k = 0; k < 6; k++) {
        G1[k] = (0.0 - 0.5 * dtdx * bm1[k]) - 0.5 * dtdx * bm2[k];
        G2[k] = (0.0 - 0.5 * dtdx * bp1[k]) - 0.5 * dtdx * bp2[k];
    }
}

DC_HD size_t qidx(int m, int i, int j, int mx)
{
    // q(m, i, j) with m 0-based, i, j Fortran indices, mbc = 2
    return (size_t)m + 7u * ((size_t)(i + 1) + (size_t)(mx + 4) * (size_t)(j + 1));
}
DC_HD size_t aidx(int i, int j, int mx, int maux)
{
    return (size_t)maux * ((size_t)(i + 1) + (size_t)(mx + 4) * (size_t)(j + 1));
}

// Raw q and bed of column c, row R (zeros outside the array; boundary
// conditions applied through bc_source when A.bc is set).
DC_HD void load_row(const StripArgs& A, int c, int R, bool col_in_array, double* q7, double& b)
{
    for (int m = 0; m < 7; m++) q7[m] = 0.0;
    b = 0.0;
    const int mx = A.mx, my = A.my;
    if (!(R >= -1 && R <= my + 2) || !col_in_array) return;
    const bool ghost = (c < 1 || c > mx || R < 1 || R > my);
    if (ghost && A.bc[0] != 0) {
        int si, sj;
        double fu, fv;
        bc_source(A.bc, mx, my, c, R, si, sj, fu, fv);
        const size_t o = qidx(0, si, sj, mx);
        for (int m = 0; m < 7; m++) q7[m] = A.q_in[o + m];
        q7[1] = fu * q7[1];
        q7[2] = fv * q7[2];
    } else {
        const size_t o = qidx(0, c, R, mx);
        for (int m = 0; m < 7; m++) q7[m] = A.q_in[o + m];
    }
    b = A.aux[aidx(c, R, mx, A.maux)];
}

// The whole-segment dry check of rows [jA, jB] of a strip.  An owned cell's
// output differs from the dry state (zeros and bdif >= 0) only if a cell of
// its 3x3 neighbourhood is wet (h > dry_tolerance before qfix): a wave needs
// a wet side, the second-order correction needs four wet cells in a row
// (flux2's gate), and the transverse terms reaching the cell come from edges
// touching its neighbours; an owned ghost cell only needs itself dry
// (b4step2 is pointwise).  So if rows jA-1..jB+1 and columns i0-1..i0+58
// (plus the owned ghost rows and columns) are dry, every output of the
// segment is the dry state.  With the boundary conditions applied on load
// only interior cells are checked (the ghost sources are interior cells in
// range).
//
// The strip body scans one cell per lane (dry_scan_cell).  The split step's
// dry pass reads each checked row as one contiguous run of doubles (the cells
// [kl, kh] of row R are q words [7(kl+1+(mx+4)(R+1)), 7(kh+2+(mx+4)(R+1))))
// with 16-byte accesses, and writes the dry state of the owned cells [wl, wh]
// from the same run: word m of a cell becomes 0 for m < 6 and max(bdif, 0)
// for m = 6, an elementwise map of the run.  Owned ghost cells whose values
// come through the boundary map are written one by one.
struct DryScan {
    bool own;       // this lane's column is written by the strip (interior or owned ghost)
    bool bcl;       // boundary conditions applied on load
    bool vec_out;   // q_out has q_in's 16-byte alignment (16-byte stores)
    int lo, hi;     // rows scanned: jA-1..jB+1, or the owned ghost rows at the patch edges
    int r0, r1;     // rows read as runs (their h checked)
    int kl, kh;     // columns of the run: the checked columns
    int wl, wh;     // owned columns written from the run
};

DC_HD DryScan dry_scan(const StripArgs& A, bool own)
{
    DryScan D;
    const int mx = A.mx, my = A.my;
    D.bcl = (A.bc[0] != 0);
    const bool first_s = (A.i0 == 1), last_s = (A.i0 + STRIP_W - 1 >= mx);
    const int clo = D.bcl ? 1 : -1, chi = D.bcl ? mx : mx + 2;
    const int rlo = D.bcl ? 1 : -1, rhi = D.bcl ? my : my + 2;
    const int ca = first_s ? -1 : A.i0 - 1;
    const int cb = last_s ? mx + 2 : A.i0 + STRIP_W;
    const int ol = first_s ? -1 : A.i0;
    const int oh = last_s ? mx + 2 : A.i0 + STRIP_W - 1;
    D.kl = ca > clo ? ca : clo;
    D.kh = cb < chi ? cb : chi;
    D.wl = ol > clo ? ol : clo;
    D.wh = oh < chi ? oh : chi;
    D.own = own;
    D.vec_out = A.q_out != nullptr &&
                ((((uintptr_t)A.q_in) ^ ((uintptr_t)A.q_out)) & (uintptr_t)15) == 0;
    D.lo = (A.jA == 1) ? -1 : A.jA - 1;
    D.hi = (A.jB == my) ? my + 2 : A.jB + 1;
    D.r0 = D.lo > rlo ? D.lo : rlo;
    D.r1 = D.hi < rhi ? D.hi : rhi;
    return D;
}

// Row R of the scan on this lane's own column, one cell at a time (the strip
// body's check): h if the cell is checked (else 0), whether the lane writes
// the cell (wr) and then its raw bdif (through the boundary map for a ghost
// cell when the boundary conditions are applied on load).
DC_HD void dry_scan_cell(const StripArgs& A, const DryScan& D, int c, int R, double& h, bool& wr,
                         double& bd)
{
    const int mx = A.mx, my = A.my;
    const double* __restrict__ qi = A.q_in;
    h = 0.0;
    if (c >= D.kl && c <= D.kh && R >= D.r0 && R <= D.r1) h = qi[qidx(0, c, R, mx)];
    const bool row_in = (R >= -1 && R <= my + 2) && R <= D.hi;
    const bool row_ghost_w = (R >= -1 && R <= 0 && A.jA == 1) ||
                             (R >= my + 1 && R <= my + 2 && A.jB == my);
    const bool row_interior = (R >= 1 && R <= my) && R >= A.jA && R <= A.jB;
    wr = D.own && row_in && (row_ghost_w || row_interior);
    bd = 0.0;
    if (wr) {
        if ((c < 1 || c > mx || R < 1 || R > my) && D.bcl) {
            int si, sj;
            double fu, fv;
            bc_source(A.bc, mx, my, c, R, si, sj, fu, fv);
            bd = qi[qidx(6, si, sj, mx)];
        } else {
            bd = qi[qidx(6, c, R, mx)];
        }
    }
}

// One row run as this lane holds it: the 16-byte chunks lane, lane+64,
// lane+128, lane+192 of the run aligned down to a 16-byte boundary of q_in
// (62 cells at most: 218 chunks), each as its two doubles (0 outside the run).
constexpr int VROW_CHUNKS = 4;
struct VRow {
    double x[VROW_CHUNKS], y[VROW_CHUNKS];
};

// Word positions of row R's run: rb = first word of the row, [wb, we) the
// run, a0 its 16-byte aligned start.
struct VRun {
    long rb, wb, we, a0;
};

DC_HD VRun vrun(const StripArgs& A, const DryScan& D, int R)
{
    VRun r;
    r.rb = 7L * (long)(A.mx + 4) * (long)(R + 1);
    r.wb = r.rb + 7L * (D.kl + 1);
    r.we = r.rb + 7L * (D.kh + 2);
    r.a0 = r.wb - (long)((((uintptr_t)A.q_in >> 3) + (uintptr_t)r.wb) & (uintptr_t)1);
    return r;
}

DC_HD void vrow_load(const double* __restrict__ q, const VRun& r, int lane, VRow& v)
{
#pragma unroll
    for (int u = 0; u < VROW_CHUNKS; u++) {
        const long w = r.a0 + 2L * (lane + STRIP_LANES * u);
        if (w >= r.wb && w + 1 < r.we) {
            const double2 t = *reinterpret_cast<const double2*>(q + w);
            v.x[u] = t.x;
            v.y[u] = t.y;
        } else {
            v.x[u] = (w >= r.wb && w < r.we) ? q[w] : 0.0;
            v.y[u] = (w + 1 >= r.wb && w + 1 < r.we) ? q[w + 1] : 0.0;
        }
    }
}

// Any h > dry_tolerance in this lane's part of the run (the run holds only
// checked cells).
DC_HD bool vrow_wet(const VRun& r, const VRow& v, int lane, double dry_tolerance)
{
    bool wet = false;
#pragma unroll
    for (int u = 0; u < VROW_CHUNKS; u++) {
        const long w = r.a0 + 2L * (lane + STRIP_LANES * u);
        const int m0 = ((int)(w - r.rb) + 7) % 7;   // component of the first double
        if (m0 == 0 && w >= r.wb && w < r.we && v.x[u] > dry_tolerance) wet = true;
        if (m0 == 6 && w + 1 >= r.wb && w + 1 < r.we && v.y[u] > dry_tolerance) wet = true;
    }
    return wet;
}

// The dry state of cell (c, R): qfix zeroes the dry cell and bdif is clipped
// at 0 (b4step2.f90:56-63); a cell with a dry neighbourhood keeps that state.
DC_HD void dry_write(const StripArgs& A, int c, int R, double bd)
{
    double* __restrict__ qo = A.q_out;
    const size_t o = qidx(0, c, R, A.mx);
    for (int z = 0; z < 6; z++) qo[o + z] = 0.0;
    qo[o + 6] = dmax(bd, 0.0);
}

// Dry state of the cells row R of the segment owns: the run's owned columns
// from the run (v, when R was read as a run), and on this lane's own column
// an owned ghost cell whose bdif comes through the boundary map.
DC_HD void dry_row_write(const StripArgs& A, const DryScan& D, int c, int lane, int R,
                         const VRow& v)
{
    const int mx = A.mx, my = A.my;
    if (R > D.hi) return;
    const bool row_own = (R >= -1 && R <= 0 && A.jA == 1) ||
                         (R >= my + 1 && R <= my + 2 && A.jB == my) ||
                         (R >= 1 && R <= my && R >= A.jA && R <= A.jB);
    if (!row_own) return;
    if (R >= D.r0 && R <= D.r1) {
        const VRun r = vrun(A, D, R);
        const long ob = r.rb + 7L * (D.wl + 1), oe = r.rb + 7L * (D.wh + 2);
        double* __restrict__ qo = A.q_out;
#pragma unroll
        for (int u = 0; u < VROW_CHUNKS; u++) {
            const long w = r.a0 + 2L * (lane + STRIP_LANES * u);
            const int m0 = ((int)(w - r.rb) + 7) % 7;
            const double ox = (m0 == 6) ? dmax(v.x[u], 0.0) : 0.0;
            const double oy = (m0 == 5) ? dmax(v.y[u], 0.0) : 0.0;
            const bool in0 = (w >= ob && w < oe), in1 = (w + 1 >= ob && w + 1 < oe);
            if (in0 && in1 && D.vec_out) {
                double2 t;
                t.x = ox;
                t.y = oy;
                *reinterpret_cast<double2*>(qo + w) = t;
            } else {
                if (in0) qo[w] = ox;
                if (in1) qo[w + 1] = oy;
            }
        }
    }
    if (D.bcl && D.own && (c < 1 || c > mx || R < 1 || R > my)) {
        int si, sj;
        double fu, fv;
        bc_source(A.bc, mx, my, c, R, si, sj, fu, fv);
        dry_write(A, c, R, A.q_in[qidx(6, si, sj, mx)]);
    }
}

// The strip body.  Returns this lane's CFL contribution (max |dt/dx * s|).
template <int MODE, class Lane>
DC_HD double strip_step(Lane& L, const dclaw_params& P, const StripArgs& A)
{
    constexpr bool FLUX = (MODE == MODE_FLUX);
    constexpr bool NOSRC2 = (MODE == MODE_STEP_NOSRC2);
    const int lane = L.id();
    const int c = A.i0 - STRIP_HALO + lane;   // this lane's column
    const int mx = A.mx, my = A.my;
    const double dtdx = A.dt / A.dx, dtdy = A.dt / A.dy;
    const bool col_in_array = (c >= -1 && c <= mx + 2);
    const bool col_taudir = (c >= 0 && c <= mx + 1);
    const int lane_prev = lane > 0 ? lane - 1 : 0;
    const int lane_next = lane < STRIP_LANES - 1 ? lane + 1 : STRIP_LANES - 1;
    // interior cells this lane writes
    const bool col_out = (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx);
    // ghost columns written by this strip (first strip: -1, 0; last: mx+1, mx+2)
    const bool col_ghost_w = (c >= -1 && c <= 0 && A.i0 == 1) ||
                             (c >= mx + 1 && c <= mx + 2 && A.i0 + STRIP_W - 1 >= mx);

    double cfl = 0.0;
    bool src2_wet = false;   // MODE_STEP_NOSRC2: a cell of this tile needs src2
    // calc_taudir of a cell whose whole neighbourhood is dry: 0 when
    // dry_tolerance > 0 (early return), else the h = 0 fall-through values
    const bool drytol_pos = P.dry_tolerance > 0.0;
    const double dry_fs = drytol_pos ? 0.0 : 1.0;
    const double dry_tx = drytol_pos ? 0.0 : A.dx * 1.0;
    const double dry_ty = drytol_pos ? 0.0 : A.dy * 1.0;

    RowCell q3, q2, q1, q0;   // rows R-3, R-2, R-1, R
    zero_row(q3); zero_row(q2); zero_row(q1); zero_row(q0);
    double h_r4 = 0.0;        // h of row R-4
    Taudir ta2, ta1;          // rows R-2, R-1
    ta2.fsphi = ta2.tx = ta2.ty = 0.0;
    ta1 = ta2;
    EdgeWaves yw_pend;        // y-edge R-2 (unlimited)
    for (int k = 0; k < 3; k++) yw_pend.s[k] = 0.0;
    for (int k = 0; k < 6; k++) { yw_pend.w1[k] = 0.0; yw_pend.w5[k] = 0.0; }
    yw_pend.w2h = yw_pend.w2n = 0.0;
    double ydotl[3] = {0.0, 0.0, 0.0};   // dot(w(R-3), w(R-2))
    double gy_imp2_m[6], gy_imp2_p[6];    // imp2 of y-edge R-3 with cell (c, R-3)
    double fxm_prev[6], fxp_prev[6];      // x faddm/faddp of row R-3
    double gx_prev2[6];                   // Gx_{R-3}(c, 2)
    double gp_prev[6];                    // gp(c, R-3)
    for (int k = 0; k < 6; k++) {
        gy_imp2_m[k] = gy_imp2_p[k] = 0.0;
        fxm_prev[k] = fxp_prev[k] = 0.0;
        gx_prev2[k] = gp_prev[k] = 0.0;
    }

    // Whole-segment dry check (DryScan).  Rows are scanned in chunks of
    // DRY_CHUNK with h and bdif loaded together (one pass over the cache
    // lines), and the dry state of each chunk's own rows is written as soon as
    // the chunk is known dry; if a later chunk is wet, the marching loop
    // rewrites every row this segment owns, so those early writes are
    // harmless.  Skipped when the split step's dry pass has already found the
    // segment not dry and written its dry state.
    if (!A.known_wet) {
        const DryScan D = dry_scan(A, !FLUX && col_in_array && (col_out || col_ghost_w));
        bool seg_wet = false;
        for (int R0 = D.lo; R0 <= D.hi; R0 += DRY_CHUNK) {
            double hh[DRY_CHUNK], bd[DRY_CHUNK];
            bool wr[DRY_CHUNK];
#pragma unroll
            for (int k = 0; k < DRY_CHUNK; k++) dry_scan_cell(A, D, c, R0 + k, hh[k], wr[k], bd[k]);
            bool w = false;
#pragma unroll
            for (int k = 0; k < DRY_CHUNK; k++) w |= (hh[k] > P.dry_tolerance);
            if (L.any(w)) {
                seg_wet = true;
                break;
            }
#pragma unroll
            for (int k = 0; k < DRY_CHUNK; k++)
                if (wr[k]) dry_write(A, c, R0 + k, bd[k]);
        }
        if (!seg_wet) {
            if (NOSRC2 && L.id() == 0) *A.tile_flag = 0;
            if (!FLUX && A.aux_out) {
                for (int r = A.jA - 2; r <= A.jB + 2; r++) {
                    if (r < 0 || r > my + 1 || !col_taudir) continue;
                    const bool row_own = (r >= A.jA && r <= A.jB) || (r == 0 && A.jA == 1) ||
                                         (r == my + 1 && A.jB == my);
                    const bool col_own =
                        (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx) ||
                        (c == 0 && A.i0 == 1) || (c == mx + 1 && A.i0 + STRIP_W - 1 >= mx);
                    if (row_own && col_own) {
                        const size_t o = aidx(c, r, mx, A.maux);
                        A.aux_out[o + 3] = dry_fs;
                        A.aux_out[o + 4] = dry_tx;
                        A.aux_out[o + 5] = dry_ty;
                    }
                }
            }
            return 0.0;
        }
    }

    // Row loads run one iteration ahead of their use (software prefetch):
    // the raw values of row R+1 are requested before row R is processed.
    // step2's top segment runs one more row to assemble the fluxes of row
    // my+1 (and gm/gp of row my+2)
    const int R_end = A.jB + 3 + ((FLUX && A.jB == my) ? 1 : 0);
    double nq7[7], nb;
    load_row(A, c, A.jA - 3, col_in_array, nq7, nb);
    for (int R = A.jA - 3; R <= R_end; R++) {
        // ---------------- S1: row R (prefetched), b4step2 pointwise, F1, F2
        {
            double q7[7];
            for (int m = 0; m < 7; m++) q7[m] = nq7[m];
            const double b = nb;
            if (R + 1 <= R_end) load_row(A, c, R + 1, col_in_array, nq7, nb);
            const bool row_in = (R >= -1 && R <= my + 2);
            if (FLUX) prepare_cell_q0(P, q7, b, q0);
            else prepare_cell(P, q7, b, q0);
            const bool row_ghost_w = (R >= -1 && R <= 0 && A.jA == 1) ||
                                     (R >= my + 1 && R <= my + 2 && A.jB == my);
            const bool row_interior = (R >= 1 && R <= my);
            if (!FLUX && row_in && col_in_array &&
                ((row_ghost_w && (col_out || col_ghost_w)) || (row_interior && col_ghost_w &&
                                                              R >= A.jA && R <= A.jB))) {
                const size_t o = qidx(0, c, R, mx);
                A.q_out[o + 0] = q0.h; A.q_out[o + 1] = q0.hu; A.q_out[o + 2] = q0.hv;
                A.q_out[o + 3] = q0.hm0; A.q_out[o + 4] = q0.p0; A.q_out[o + 5] = q0.hchi0;
                A.q_out[o + 6] = q0.bdif;
            }
        }

        // A window whose rows R-4..R are dry in every lane of the strip has
        // every flux, wave, transverse term and taudir exactly +0 (both-dry
        // edges are skipped by rpn2, dry cells by rpt2 and src2, and the
        // wet-dry gate zeroes the corrections), so the row reduces to
        // writing the dry state; the decision is uniform across the wave.
        const bool wet_here = (q0.h > 0.0) || (q1.h > 0.0) || (q2.h > 0.0) || (q3.h > 0.0) ||
                              (h_r4 > 0.0);
        if (!L.any(wet_here)) {
            if (A.aux_out) {
                const int r = R - 1;
                if (r >= 0 && r <= my + 1 && col_taudir && r >= A.jA - 2 && r <= A.jB + 2) {
                    const bool row_own = (r >= A.jA && r <= A.jB) || (r == 0 && A.jA == 1) ||
                                         (r == my + 1 && A.jB == my);
                    const bool col_own =
                        (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx) ||
                        (c == 0 && A.i0 == 1) || (c == mx + 1 && A.i0 + STRIP_W - 1 >= mx);
                    if (row_own && col_own) {
                        const size_t o = aidx(c, r, mx, A.maux);
                        A.aux_out[o + 3] = dry_fs;
                        A.aux_out[o + 4] = dry_tx;
                        A.aux_out[o + 5] = dry_ty;
                    }
                }
            }
            ta2 = ta1;
            ta1.fsphi = ta1.tx = ta1.ty = 0.0;
            const int j = R - 3;
            if (!FLUX && j >= A.jA && j <= A.jB && col_out) {
                const size_t o = qidx(0, c, j, mx);
                for (int z = 0; z < 6; z++) A.q_out[o + z] = 0.0;
                A.q_out[o + 6] = q3.bdif;
            }
            for (int z = 0; z < 3; z++) { ydotl[z] = 0.0; yw_pend.s[z] = 0.0; }
            for (int z = 0; z < 6; z++) {
                yw_pend.w1[z] = 0.0;
                yw_pend.w5[z] = 0.0;
                gy_imp2_m[z] = gy_imp2_p[z] = 0.0;
                fxm_prev[z] = fxp_prev[z] = 0.0;
                gx_prev2[z] = gp_prev[z] = 0.0;
            }
            yw_pend.w2h = yw_pend.w2n = 0.0;
            h_r4 = q3.h;
            q3 = q2;
            q2 = q1;
            q1 = q0;
            continue;
        }
        // ---------------- S2: setvars of F1 and calc_taudir for row R-1
        {
            const int r = R - 1;
            // setvars(F1) of a wet cell serves calc_taudir here and the right
            // side of both sweeps' edges (S3 now, S4 when the row is R-2)
            q1.rho1 = 0.0;
            q1.tau1 = 0.0;
            if (r >= A.jA - 2 && r <= A.jB + 2 && q1.h > P.dry_tolerance) {
                const RhoTau v1 = rho_tau(P, q1.h, q1.u, q1.v, q1.m1, q1.p1, q1.chi1);
                q1.rho1 = v1.rho;
                q1.tau1 = v1.tau;
            }
            RhoTau sv1;
            sv1.rho = q1.rho1;
            sv1.tau = q1.tau1;
            // neighbour columns of row R-1 (F1 states), all lanes participate
            const double lh = L.prev(q1.h), lhu = L.prev(q1.hu);
            const double lhv = L.prev(q1.hv), lhm = L.prev(q1.hm0);
            const double lp = L.prev(q1.p0), lhc = L.prev(q1.hchi0);
            const double lb = L.prev(q1.b), lbd = L.prev(q1.bdif);
            const double rh = L.next(q1.h), rhu = L.next(q1.hu);
            const double rhv = L.next(q1.hv), rhm = L.next(q1.hm0);
            const double rp = L.next(q1.p0), rhc = L.next(q1.hchi0);
            const double rb = L.next(q1.b), rbd = L.next(q1.bdif);
            Taudir t;
            t.fsphi = t.tx = t.ty = 0.0;
            if (FLUX) {
                // step2 reads fsphi and taudir as b4step2 left them in aux
                if (r >= -1 && r <= my + 2 && col_in_array && r >= A.jA - 2 && r <= A.jB + 2) {
                    const size_t o = aidx(c, r, mx, A.maux);
                    t.fsphi = A.aux[o + 3];
                    t.tx = A.aux[o + 4];
                    t.ty = A.aux[o + 5];
                }
            } else if (r >= 0 && r <= my + 1 && col_taudir && r >= A.jA - 2 && r <= A.jB + 2) {
                Cell C = {q1.h, q1.hu, q1.hv, q1.hm0, q1.p0, q1.hchi0};
                Cell Lc = {lh, lhu, lhv, lhm, lp, lhc};
                Cell Rc = {rh, rhu, rhv, rhm, rp, rhc};
                Cell B = {q2.h, q2.hu, q2.hv, q2.hm0, q2.p0, q2.hchi0};
                Cell T = {q0.h, q0.hu, q0.hv, q0.hm0, q0.p0, q0.hchi0};
                t = calc_taudir_cell(P, A.dx, A.dy, C, q1.b - q1.bdif, Lc, lb - lbd, Rc, rb - rbd,
                                     B, q2.b - q2.bdif, T, q0.b - q0.bdif, sv1);
            }
            if (!FLUX && A.aux_out && r >= 0 && r <= my + 1 && col_taudir) {
                const bool row_own = (r >= A.jA && r <= A.jB) || (r == 0 && A.jA == 1) ||
                                     (r == my + 1 && A.jB == my);
                const bool col_own = (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx) ||
                                     (c == 0 && A.i0 == 1) ||
                                     (c == mx + 1 && A.i0 + STRIP_W - 1 >= mx);
                if (row_own && col_own) {
                    const size_t o = aidx(c, r, mx, A.maux);
                    A.aux_out[o + 3] = t.fsphi;
                    A.aux_out[o + 4] = t.tx;
                    A.aux_out[o + 5] = t.ty;
                }
            }
            ta2 = ta1;
            ta1 = t;
        }

        // setvars of F2 for row R-2: the left side of the y-edge R-1 and of
        // the x-edges of row R-2
        RhoTau v2;
        v2.rho = 0.0;
        v2.tau = 0.0;
        if (R - 2 >= A.jA - 2 && R - 2 <= A.jB + 1 && q2.h > P.dry_tolerance)
            v2 = rho_tau(P, q2.h, q2.u, q2.v, q2.m2, q2.p2, q2.chi2);
        RhoTau v1_q2, v1_q1;
        v1_q2.rho = q2.rho1;
        v1_q2.tau = q2.tau1;
        v1_q1.rho = q1.rho1;
        v1_q1.tau = q1.tau1;

        // ---------------- S3: y-edge R-1 (cells (c, R-2), (c, R-1))
        EdgeWaves yw_new;
        double ydotr[3] = {0.0, 0.0, 0.0};
        {
            const int k = R - 1;
            for (int z = 0; z < 3; z++) yw_new.s[z] = 0.0;
            for (int z = 0; z < 6; z++) { yw_new.w1[z] = 0.0; yw_new.w5[z] = 0.0; }
            yw_new.w2h = yw_new.w2n = 0.0;
            if (k >= A.jA - 1 && k <= A.jB + 2 && col_taudir) {
                const bool second = (k - 1) >= 0;
                const Side sL = side_of(q2, second, ta2.ty, ta2.fsphi, second ? v2 : v1_q2);
                const Side sR = side_of(q1, false, ta1.ty, ta1.fsphi, v1_q1);
                // y sweep: primitives u = hv/h (normal), v = hu/h (transverse)
                Side l2 = sL, r2 = sR;
                l2.u = q2.v; l2.v = q2.u;
                r2.u = q1.v; r2.v = q1.u;
                yw_new = rpn2_edge(P, 2, l2, r2);
            }
            wave_dots(yw_pend, yw_new, 2, ydotr);
        }

        // ---------------- S4: x-sweep of row R-2
        double fxm[6], fxp[6], Gx1[6], Gx2[6];
        for (int z = 0; z < 6; z++) { fxm[z] = fxp[z] = Gx1[z] = Gx2[z] = 0.0; }
        {
            const int r = R - 2;
            const bool row_act = (r >= A.jA - 1 && r <= A.jB + 1);
            // left neighbour's F2 (or F1 for the slice's first cell, column
            // -1) state: each lane selects its own left-role state, then the
            // lane to its right reads it
            const bool sec_me = (c >= 0);
            const double lh = L.prev(q2.h), lhu = L.prev(q2.hu);
            const double lhv = L.prev(q2.hv);
            const double lhm = L.prev(sec_me ? q2.hm2 : q2.hm1);
            const double lp = L.prev(sec_me ? q2.p2 : q2.p1);
            const double lhc = L.prev(sec_me ? q2.hchi2 : q2.hchi1);
            const double lm = L.prev(sec_me ? q2.m2 : q2.m1);
            const double lc = L.prev(sec_me ? q2.chi2 : q2.chi1);
            const double lrho = L.prev(sec_me ? v2.rho : q2.rho1);
            const double ltau = L.prev(sec_me ? v2.tau : q2.tau1);
            const double lu = L.prev(q2.u), lv = L.prev(q2.v);
            const double lb = L.prev(q2.b), lbd = L.prev(q2.bdif);
            const double ltx = L.prev(ta2.tx), lfs = L.prev(ta2.fsphi);
            EdgeWaves xw;
            for (int z = 0; z < 3; z++) xw.s[z] = 0.0;
            for (int z = 0; z < 6; z++) { xw.w1[z] = 0.0; xw.w5[z] = 0.0; }
            xw.w2h = xw.w2n = 0.0;
            const bool edge_ok = row_act && c >= 0 && c <= mx + 2 && lane >= 1;
            if (edge_ok) {
                Side sL;
                sL.h = lh; sL.hu = lhu; sL.hv = lhv;
                sL.hm = lhm;
                sL.p = lp;
                sL.hchi = lhc;
                sL.m = lm;
                sL.chi = lc;
                sL.u = lu; sL.v = lv;
                sL.b = lb - lbd;
                sL.taudir = ltx; sL.fs = lfs;
                sL.rho = lrho; sL.tau = ltau;
                const Side sR = side_of(q2, false, ta2.tx, ta2.fsphi, v1_q2);
                xw = rpn2_edge(P, 1, sL, sR);
            }
            // limiter neighbours
            const EdgeWaves xn = next_waves(L, xw);
            double dr[3];
            wave_dots(xw, xn, 1, dr);
            double dl[3];
            for (int f = 0; f < 3; f++) dl[f] = L.prev(dr[f]);
            // wet-dry gate: h(c-2) h(c-1) h(c) h(c+1) > 0, in flux2's product order
            const double hm2c = L.prev(lh);
            const double hm1c = lh;
            const double hp1c = L.next(q2.h);
            const bool wet = (hm1c * q2.h * hm2c * hp1c > 0.0);
            const EdgeFlux ef = edge_flux(P, 1, xw, dl, dr, wet, dtdx);
            const bool lim_ok = row_act && c >= 1 && c <= mx + 1 && lane >= 2 && lane <= 62;
            if (lim_ok) {
                for (int z = 0; z < 6; z++) { fxm[z] = ef.faddm[z]; fxp[z] = ef.faddp[z]; }
                for (int f = 0; f < 3; f++) {
                    const double a = fabs(dtdx * xw.s[f]);
                    cfl = dmax(cfl, a);
                }
            }
            // transverse split in cell c: imp1 with amdq of edge c+1, imp2 with apdq of edge c
            double amdq_next[6];
            for (int z = 0; z < 6; z++) amdq_next[z] = L.next(ef.amdq[z]);
            if (P.transverse != 0) {
                // imp1 (amdq of edge c+1) reaches cells 0..mx, imp2 (apdq of edge c)
                // cells 1..mx+1 (flux2's loops over edges 1..mx+1); the update
                // reads cells 1..mx, step2's gm/gp also cell mx+1
                const bool cell_ok1 = row_act && c >= 1 && c <= mx && lane >= 2 && lane <= 61;
                const bool cell_ok2 = row_act && c >= 1 && c <= (FLUX ? mx + 1 : mx) &&
                                      lane >= 2 && lane <= 61;
                double bm1[6], bp1[6], bm2[6], bp2[6];
                rpt2_cell_split(P, 1, q2.h, q2.hu, q2.hv, q2.hm2, q2.hchi2, q2.b, q2.bdif, q3.b,
                                q1.b, amdq_next, cell_ok1, bm1, bp1);
                rpt2_cell_split(P, 1, q2.h, q2.hu, q2.hv, q2.hm2, q2.hchi2, q2.b, q2.bdif, q3.b,
                                q1.b, ef.apdq, cell_ok2, bm2, bp2);
                if (cell_ok1 || cell_ok2) trans_combine(bm1, bp1, bm2, bp2, dtdx, Gx1, Gx2);
            }
        }

        // ---------------- S5: y-edge R-2 limited and split
        double fym[6], fyp[6], Gy1[6], Gy2[6];
        for (int z = 0; z < 6; z++) { fym[z] = fyp[z] = Gy1[z] = Gy2[z] = 0.0; }
        {
            const int k = R - 2;
            const bool act = (k >= A.jA && k <= A.jB + 1) && col_taudir;
            // wet gate along the column: h(k-2) h(k-1) h(k) h(k+1) = rows R-4..R-1
            const bool wet = (q3.h * q2.h * h_r4 * q1.h > 0.0);
            const EdgeFlux ef = edge_flux(P, 2, yw_pend, ydotl, ydotr, wet, dtdy);
            if (act) {
                for (int z = 0; z < 6; z++) { fym[z] = ef.faddm[z]; fyp[z] = ef.faddp[z]; }
                for (int f = 0; f < 3; f++) cfl = dmax(cfl, fabs(dtdy * yw_pend.s[f]));
            }
            // topo of the neighbouring columns for rows R-3 and R-2
            const double b3l = L.prev(q3.b), b3r = L.next(q3.b);
            const double b2l = L.prev(q2.b), b2r = L.next(q2.b);
            if (P.transverse != 0) {
                double bm1[6], bp1[6], bm2[6], bp2[6];
                // imp1: amdq of edge R-2 split in cell (c, R-3)
                rpt2_cell_split(P, 2, q3.h, q3.hu, q3.hv, q3.hm2, q3.hchi2, q3.b, q3.bdif, b3l,
                                b3r, ef.amdq, act, bm1, bp1);
                for (int z = 0; z < 6; z++) {
                    Gy1[z] = (0.0 - 0.5 * dtdy * bm1[z]) - gy_imp2_m[z];
                    Gy2[z] = (0.0 - 0.5 * dtdy * bp1[z]) - gy_imp2_p[z];
                }
                // imp2: apdq of edge R-2 split in cell (c, R-2), kept for the next row
                rpt2_cell_split(P, 2, q2.h, q2.hu, q2.hv, q2.hm2, q2.hchi2, q2.b, q2.bdif, b2l,
                                b2r, ef.apdq, act, bm2, bp2);
                for (int z = 0; z < 6; z++) {
                    gy_imp2_m[z] = 0.5 * dtdy * bm2[z];
                    gy_imp2_p[z] = 0.5 * dtdy * bp2[z];
                }
            }
        }

        // ---------------- S6: assemble, update and src2 for row R-3
        {
            const int j = R - 3;
            // fm(c, j) = ((0 + faddm_x(c, j)) + Gy_{c-1}(j, 2)) + Gy_c(j, 1); same for fp
            double gyl2[6];
            for (int z = 0; z < 6; z++) gyl2[z] = L.prev(Gy2[z]);
            double fmc[6], fpc[6];
            for (int z = 0; z < 6; z++) {
                fmc[z] = ((0.0 + fxm_prev[z]) + gyl2[z]) + Gy1[z];
                fpc[z] = ((0.0 + fxp_prev[z]) + gyl2[z]) + Gy1[z];
            }
            double fm_next[6];
            for (int z = 0; z < 6; z++) fm_next[z] = L.next(fmc[z]);
            // gm(c, j+1) = ((0 + Gx_j(c,2)) + Gx_{j+1}(c,1)) + faddm_y(c, j+1)
            double gm_next[6], gp_next[6];
            for (int z = 0; z < 6; z++) {
                const double t = (0.0 + gx_prev2[z]) + Gx1[z];
                gm_next[z] = t + fym[z];
                gp_next[z] = t + fyp[z];
            }
            if (FLUX) {
                // fm/fp(c, j): rows 1..my (+ 0 and my+1 at the patch edges), columns
                // 1..mx (+ 0, mx+1, mx+2); gm/gp(c, j+1): rows 1..my (+ 0, my+1,
                // my+2), columns 1..mx (+ 0, mx+1).  Row 0 of fm/fp has no y-sweep
                // part (step2 accumulates the y-sweeps' transverse terms for rows
                // 1..my+1 only).
                const bool first_s = (A.i0 == 1), last_s = (A.i0 + STRIP_W - 1 >= mx);
                const bool in_s = (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx);
                const bool col_f = in_s || (first_s && c == 0) ||
                                   (last_s && (c == mx + 1 || c == mx + 2));
                const bool col_g = in_s || (first_s && c == 0) || (last_s && c == mx + 1);
                const bool row_f = (j >= A.jA && j <= A.jB) || (j == 0 && A.jA == 1) ||
                                   (j == my + 1 && A.jB == my);
                const int jg = j + 1;
                const bool row_g = (jg >= A.jA && jg <= A.jB) || (jg == 0 && A.jA == 1) ||
                                   ((jg == my + 1 || jg == my + 2) && A.jB == my);
                if (row_f && col_f) {
                    const size_t o = qidx(0, c, j, mx);
                    for (int z = 0; z < 6; z++) {
                        A.fm[o + z] = (j == 0) ? 0.0 + fxm_prev[z] : fmc[z];
                        A.fp[o + z] = (j == 0) ? 0.0 + fxp_prev[z] : fpc[z];
                    }
                }
                if (row_g && col_g) {
                    const size_t o = qidx(0, c, jg, mx);
                    for (int z = 0; z < 6; z++) {
                        A.gm[o + z] = gm_next[z];
                        A.gp[o + z] = gp_next[z];
                    }
                }
            } else if (j >= A.jA && j <= A.jB && col_out) {
                double qn[6] = {q3.h, q3.hu, q3.hv, q3.hm0, q3.p0, q3.hchi0};
                double bdif_out = q3.bdif;
                for (int z = 0; z < 6; z++)
                    qn[z] = qn[z] - dtdx * (fm_next[z] - fpc[z]) - dtdy * (gm_next[z] - gp_prev[z]);
                if (NOSRC2) {
                    // src2 zeroes a cell with h <= dry_tolerance (src2.f90:103-111);
                    // the others are left for the src2 pass over flagged tiles
                    if (qn[0] <= P.dry_tolerance) {
                        for (int z = 0; z < 6; z++) qn[z] = 0.0;
                    } else {
                        src2_wet = true;
                    }
                } else {
                    // aux(i_ent) = aux(7): the erodible depth (maux >= 7 with entrainment on)
                    const double h_e = P.entrainment ? A.aux[aidx(c, j, mx, A.maux) + 6] : 0.0;
                    src2_cell(P, A.dt, qn[0], qn[1], qn[2], qn[3], qn[4], qn[5], bdif_out, h_e);
                }
                const size_t o = qidx(0, c, j, mx);
                for (int z = 0; z < 6; z++) A.q_out[o + z] = qn[z];
                A.q_out[o + 6] = bdif_out;
            }
            for (int z = 0; z < 6; z++) {
                gp_prev[z] = gp_next[z];
                gx_prev2[z] = Gx2[z];
                fxm_prev[z] = fxm[z];
                fxp_prev[z] = fxp[z];
            }
        }

        // ---------------- rotate the window
        h_r4 = q3.h;
        q3 = q2;
        q2 = q1;
        q1 = q0;
        yw_pend = yw_new;
        for (int f = 0; f < 3; f++) ydotl[f] = ydotr[f];
    }
    if (NOSRC2) {
        const bool w = L.any(src2_wet);
        if (L.id() == 0) *A.tile_flag = w ? 1 : 0;
    }
    return cfl;
}

// src2 on the interior cells of one (strip, segment) tile: cell k of the
// tile's 58 x (jB - jA + 1) block, k = column offset + 58 * row offset.
DC_HD void src2_tile_cell(const dclaw_params& P, double dt, double* q, const double* aux,
                          int maux, int mx, int i0, int jA, int jB, int k)
{
    const int i = i0 + k % STRIP_W, j = jA + k / STRIP_W;
    if (i > mx || j > jB) return;
    double* c = q + qidx(0, i, j, mx);
    double h = c[0], hu = c[1], hv = c[2], hm = c[3], p = c[4], hchi = c[5], bdif = c[6];
    const double h_e = P.entrainment ? aux[aidx(i, j, mx, maux) + 6] : 0.0;
    src2_cell(P, dt, h, hu, hv, hm, p, hchi, bdif, h_e);
    c[0] = h; c[1] = hu; c[2] = hv; c[3] = hm; c[4] = p; c[5] = hchi; c[6] = bdif;
}

// Dry pass of the split step over a group of consecutive segments of one
// strip: A.jA..A.jB span the group, cut into segments of seg_rows rows from
// A.jA (the last possibly shorter).  One scan of rows A.jA-1..A.jB+1 answers
// the whole-segment dry check of strip_step for every segment of the group
// (row by row, each wet row marking the segments whose scan range holds it)
// and writes the dry state to every cell the group owns.  Returns the mask of
// the segments that are not dry (bit s: segment s; at most 32 segments);
// strip_step with known_wet must then run on each of them, and it rewrites
// every cell its segment owns.  Only the scan and the stores touch memory, so
// this runs at far higher occupancy than the strip body.
template <class Lane>
DC_HD unsigned strip_classify(Lane& L, double dry_tolerance, const StripArgs& A, int seg_rows)
{
    const int c = A.i0 - STRIP_HALO + L.id();
    const int mx = A.mx, my = A.my;
    const bool col_in_array = (c >= -1 && c <= mx + 2);
    const bool col_out = (c >= A.i0 && c <= A.i0 + STRIP_W - 1 && c >= 1 && c <= mx);
    const bool col_ghost_w = (c >= -1 && c <= 0 && A.i0 == 1) ||
                             (c >= mx + 1 && c <= mx + 2 && A.i0 + STRIP_W - 1 >= mx);
    const int lane = L.id();
    const DryScan D = dry_scan(A, col_in_array && (col_out || col_ghost_w));
    const int nseg = (A.jB - A.jA) / seg_rows + 1;
    unsigned wet = 0u;
    for (int R0 = D.lo; R0 <= D.hi; R0 += CLASSIFY_CHUNK) {
        VRow vr[CLASSIFY_CHUNK];
#pragma unroll
        for (int k = 0; k < CLASSIFY_CHUNK; k++)
            if (R0 + k >= D.r0 && R0 + k <= D.r1)
                vrow_load(A.q_in, vrun(A, D, R0 + k), lane, vr[k]);
#pragma unroll
        for (int k = 0; k < CLASSIFY_CHUNK; k++) {
            const bool rw = R0 + k >= D.r0 && R0 + k <= D.r1 &&
                            vrow_wet(vrun(A, D, R0 + k), vr[k], lane, dry_tolerance);
            if (!L.any(rw)) continue;
            // segments s whose scan range [jA_s - 1, jB_s + 1] (extended to
            // the ghost rows at the patch edges) holds row R
            const int R = R0 + k;
            const int sr = R >= A.jA ? (R - A.jA) / seg_rows : -1;
            for (int s = sr - 1; s <= sr + 1; s++) {
                if (s < 0 || s >= nseg) continue;
                const int sa = A.jA + s * seg_rows;
                const int sb = sa + seg_rows - 1 < A.jB ? sa + seg_rows - 1 : A.jB;
                const int slo = (sa == 1) ? -1 : sa - 1;
                const int shi = (sb == my) ? my + 2 : sb + 1;
                if (R >= slo && R <= shi) wet |= 1u << s;
            }
        }
#pragma unroll
        for (int k = 0; k < CLASSIFY_CHUNK; k++) dry_row_write(A, D, c, lane, R0 + k, vr[k]);
    }
    return wet;
}

// The default method switches of a D-Claw run (second order, transverse
// = 2, MC limiter on every wave family, src2method 0, alphamethod 0, no
// rare-corrector test, segregation and Manning friction off).
DC_HD bool scheme_is_std(const dclaw_params& P)
{
    return P.order == 2 && P.transverse == 2 && P.mthlim[0] == 4 && P.mthlim[1] == 4 &&
           P.mthlim[2] == 4 && P.mthlim[3] == 4 && P.mthlim[4] == 4 && P.src2method == 0 &&
           P.alphamethod == 0 && P.rarecorrectortest == 0 && P.segregation == 0 &&
           P.friction_forcing == 0 && P.entrainment == 0;
}

// The strip body with the method switches folded to constants when they are
// the defaults (a copy of the record whose switches are literal constants, so
// the inlined body drops the branches of the other methods), or read from
// the record otherwise.  Both copies compute the same thing.
template <int MODE, class Lane>
DC_HD double strip_step_dispatch(Lane& L, const dclaw_params& P, const StripArgs& A)
{
    if (scheme_is_std(P)) {
        dclaw_params Q = P;
        Q.order = 2;
        Q.transverse = 2;
        for (int k = 0; k < 5; k++) Q.mthlim[k] = 4;
        Q.src2method = 0;
        Q.alphamethod = 0;
        Q.rarecorrectortest = 0;
        Q.segregation = 0;
        Q.friction_forcing = 0;
        Q.entrainment = 0;
        return strip_step<MODE>(L, Q, A);
    }
    return strip_step<MODE>(L, P, A);
}

}  // namespace dclaw
