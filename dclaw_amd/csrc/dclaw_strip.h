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

// Tile flag of the split step (MODE_STEP_NOSRC2): 0 when no output cell of
// the tile needs src2, else TF_ANY plus a bit for each band of the tile
// that holds such a cell -- its first and last STRIP_HALO columns (west:
// i0..i0+2, east: i0+55..i0+57) and rows (south: jA..jA+2, north:
// jB-2..jB), the cells a neighbouring tile's march reads.
constexpr int TF_ANY = 1;
constexpr int TF_WEST = 2;
constexpr int TF_EAST = 4;
constexpr int TF_SOUTH = 8;
constexpr int TF_NORTH = 16;
constexpr int TF_ALL = TF_ANY | TF_WEST | TF_EAST | TF_SOUTH | TF_NORTH;
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
    int* tile_flag;              // MODE_STEP_NOSRC2: the tile flag (TF_* bits), 0 if no cell
                                 // here needs src2
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

DC_HD size_t qidx(int m, int i, int j, int mx)
{
    // q(m, i, j) with m 0-based, i, j Fortran indices, mbc = 2
    return (size_t)m + 7u * ((size_t)(i + 1) + (size_t)(mx + 4) * (size_t)(j + 1));
}
DC_HD size_t aidx(int i, int j, int mx, int maux)
{
    return (size_t)maux * ((size_t)(i + 1) + (size_t)(mx + 4) * (size_t)(j + 1));
}

// Memory access of the strip body (dclaw_strip_body.inc): plain loads and
// stores here; the tangent-linear namespace reads and writes value and
// tangent arrays through its own overloads.
DC_HD double ld(const double* __restrict__ p, size_t i) { return p[i]; }
DC_HD void st(double* __restrict__ p, size_t i, double v) { p[i] = v; }

#include "dclaw_strip_body.inc"

// ---- the split step's dry pass and src2 pass (double only)

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

// src2 on the interior cells of one (strip, segment) tile: cell k of the
// tile's 58 x (jB - jA + 1) block, k = column offset + 58 * row offset.  q
// is the strip pass's output, in which a cell with h <= dry_tolerance is
// already what src2 makes of it (h ... hchi zeroed, bdif kept: src2.f90:
// 103-111), so such a cell is neither read further nor stored.
DC_HD void src2_tile_cell(const dclaw_params& P, double dt, double* q, const double* aux,
                          int maux, int mx, int i0, int jA, int jB, int k)
{
    const int i = i0 + k % STRIP_W, j = jA + k / STRIP_W;
    if (i > mx || j > jB) return;
    double* c = q + qidx(0, i, j, mx);
    double h = c[0];
    if (h <= P.dry_tolerance) return;
    double hu = c[1], hv = c[2], hm = c[3], p = c[4], hchi = c[5], bdif = c[6];
    const double h_e = P.entrainment ? aux[aidx(i, j, mx, maux) + 6] : 0.0;
    src2_cell(P, dt, h, hu, hv, hm, p, hchi, bdif, h_e);
    c[0] = h; c[1] = hu; c[2] = hv; c[3] = hm; c[4] = p; c[5] = hchi; c[6] = bdif;
}

// Curvature acceleration of interior cell (i, j) at src2's entry state
// (src2.f90:133-156): centred second differences of the bed b = aux(1) -
// bdif, the squared velocity along them and the bed-angle gradient of
// aux(3), clipped at 0.  0 for a cell src2 skips (dry, or hm < 0), so
// adding it to the sweep's running gz changes nothing there, as the skip
// does.  maux >= 3.
DC_HD double curvature_gacc(const dclaw_params& P, const double* q, const double* aux,
                            int maux, int mx, int i, int j, double dx, double dy)
{
    const double* c = q + qidx(0, i, j, mx);
    const double h = c[0];
    if (h <= P.dry_tolerance) return 0.0;
    if (c[3] < 0.0) return 0.0;
#define DC_BD(ii, jj) (aux[aidx(ii, jj, mx, maux)] - q[qidx(6, ii, jj, mx)])
    const double b = DC_BD(i, j), bL = DC_BD(i - 1, j), bR = DC_BD(i + 1, j);
    const double bT = DC_BD(i, j + 1), bB = DC_BD(i, j - 1);
    const double bTR = DC_BD(i + 1, j + 1), bTL = DC_BD(i - 1, j + 1);
    const double bBR = DC_BD(i + 1, j - 1), bBL = DC_BD(i - 1, j - 1);
#undef DC_BD
    const double b_xx = (bR - 2.0 * b + bL) / (dx * dx);
    const double b_yy = (bT - 2.0 * b + bB) / (dy * dy);
    const double b_xy = ((bTR - bTL) - (bBR - bBL)) / (4.0 * dx * dy);
    const double u = c[1] / h, v = c[2] / h;
    const double dtheta = -((aux[aidx(i + 1, j, mx, maux) + 2] -
                             aux[aidx(i - 1, j, mx, maux) + 2]) / (2.0 * dx));
    return dmax(u * u * b_xx + v * v * b_yy + 2.0 * u * v * b_xy + u * u * dtheta, 0.0);
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
    DryScan D = dry_scan(A, col_in_array && (col_out || col_ghost_w));
    D.vec_out = A.q_out != nullptr &&
                ((((uintptr_t)A.q_in) ^ ((uintptr_t)A.q_out)) & (uintptr_t)15) == 0;
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

}  // namespace dclaw
