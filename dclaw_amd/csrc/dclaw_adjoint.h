// dclaw_adjoint.h -- the reverse-mode (adjoint) step.
//
// Given lambda_{n+1} (the gradient of a loss with respect to q_{n+1}, same
// layout as q), the adjoint step returns lambda_n = J_q^T lambda_{n+1} and,
// for each parameter direction dP_k, g_k = lambda_{n+1} . J_P dP_k, where J is
// the Jacobian of the fused step (b4step2, step2, update, src2) at q_n with
// dt fixed.  J is exactly the tangent-linear step's (dclaw_tangent.h): at a
// branch of the scheme the adjoint is the transpose of the active branch's
// derivative, and the dot-product identity
//     <lambda_{n+1}, J dq> + sum_k g_k a_k = <lambda_n, dq>  (dP = sum a_k dP_k)
// holds to rounding.
//
// The step is laid out over the ghosted grid as stages, each a function of a
// few values of its own cell or edge (the order and index ranges of the
// strip body, dclaw_strip_body.inc):
//   F1 prepare   cell  q (boundary map on load) -> b4step2 state Q0, the two
//                      qfix passes of rpn2 (F1, F2) and setvars of both
//   F2 taudir    cell  calc_taudir from the prepared states of the cell and
//                      its 4 neighbours
//   F3 waves     edge  rpn2 (x-edges between (i-1,j) and (i,j), y-edges
//                      between (i,k-1) and (i,k))
//   F4 flux      edge  limiter, second-order correction, fluctuations
//                      (flux2's edge body, neighbours' waves through the
//                      limiter's dot products)
//   F5 split     cell  transverse rpt2 splits of the neighbouring edges'
//                      fluctuations
//   F6 update    cell  assembly of fm, fp, gm, gp in step2's order, the
//                      conservative update and src2
// The record pass runs F1..F6 and keeps every stage's inputs in a
// workspace; its q_out is the step's result bit for bit.  The reverse pass
// runs the stages backwards (B1 src2, B2 splits, B3 flux, B4 waves, B5
// taudir, B6 prepare, B7 the boundary map), each computing the
// vector-Jacobian product of its stage.  waves (rpn2), flux and split have
// hand-written reverses (rpn2_vjp, flux_vjp, split_vjp); prepare, taudir and
// src2 take theirs from the stage's local Jacobian: tangent-linear
// evaluations of the stage seeded with each input and each parameter
// direction, four directions per evaluation (vjpk).  Every stage output has exactly
// one consumer slot, so each reverse kernel gathers from its neighbours'
// slots and writes only its own cell: no atomics, and the result does not
// depend on scheduling.
#pragma once

#include "dclaw_tangent.h"
#include "dclaw_tangent_k.h"
#include "dclaw_adjoint_src2.h"

namespace dclaw_adj {

using dclaw_tl::Dual;
using TP = dclaw_tl::Params;
// the step's own (value-only) bodies, for the record pass over double
using dclaw::rho_tau;
using dclaw::src2_cell;
using dclaw::aidx;
using dclaw::bc_source;
using dclaw::qidx;

// prepared-cell fields
enum {
    PC_H, PC_HU, PC_HV, PC_BDIF, PC_HM0, PC_P0, PC_HCHI0,
    PC_HM1, PC_P1, PC_HCHI1, PC_M1, PC_CHI1, PC_U, PC_V,
    PC_HM2, PC_P2, PC_HCHI2, PC_M2, PC_CHI2,
    PC_RHO1, PC_TAU1, PC_RHO2, PC_TAU2,
    NPC
};
// rpn2 side inputs (side's own orientation: u normal)
enum {
    SD_H, SD_HU, SD_HV, SD_HM, SD_P, SD_HCHI, SD_U, SD_V, SD_M, SD_CHI, SD_BDIF, SD_TAUDIR,
    SD_FS, SD_RHO, SD_TAU,
    NSD
};
constexpr int NW = 17;      // wave record: s[3], w1[6], w5[6], w2h, w2n
constexpr int NEF = 24;     // faddm[6], faddp[6], amdq[6], apdq[6]
constexpr int NEI = 23;     // flux inputs: waves, dotl[3], dotr[3]
constexpr int NTI = 37;     // taudir inputs: 5 x (Q0 state 6 + bdif), rho1, tau1 of the centre
constexpr int NSI = 12;     // split inputs: h, hu, hv, hm2, hchi2, bdif, asdq[6]
constexpr int MAXP = 4;     // parameter directions per call

// Workspace fields (each a plane of N = ld (my+4) doubles, cell (i, j) at
// (i+1) + ld (j+1) with the row stride ld = mx+4 rounded up to 16 doubles, so
// that every row of a plane, and every 16-cell row of a tile, starts on a
// 128-byte line; an edge's values sit at its right/top cell).  A
// reverse-pass field shares planes with record fields that are dead by the
// time it is written (no stage reads a plane it or a neighbour writes in the
// same pass):
//   TD, XW, YW    read through the waves' reverse pass  -> then LTD
//   XE, YE        read through the splits' reverse pass -> then LXE, LYE
//   G             read by the update                    -> then LXA, LYA,
//                 and after the flux body's reverse pass   LQ7
//   QN            read by src2's reverse pass           -> then LS
enum {
    W_PC = 0,                 // prepared cell (NPC)
    W_TD = W_PC + NPC,        // fsphi, taudir_x, taudir_y (3)
    W_XW = W_TD + 3,          // x-edge waves (NW)
    W_YW = W_XW + NW,         // y-edge waves (NW)
    W_XE = W_YW + NW,         // x-edge flux outputs (NEF)
    W_YE = W_XE + NEF,        // y-edge flux outputs (NEF)
    W_G = W_YE + NEF,         // Gx1, Gx2, Gy1, Gy2 (24)
    W_QN = W_G + 24,          // updated state before src2 (7)
    W_LQN = W_QN + 7,         // adjoint of that state (7)
    W_LR = W_LQN + 7,         // adjoint of the cell's F1 fields from the right (top)
                              // sides of its x- and y-edge, summed (NLR)
    W_LXL = W_LR + 16,        // adjoint of the x-edge's left side (NSD)
    W_LYL = W_LXL + NSD,      // adjoint of the y-edge's bottom side (NSD)
    W_GP = W_LYL + NSD,       // parameter gradient per cell (MAXP)
    W_FIELDS = W_GP + MAXP,
    // aliases
    W_LTD = W_TD,             // adjoint of the taudir inputs that enter it (NLT planes)
    W_LXE = W_XE,             // adjoint of the x-edge flux inputs (NEI)
    W_LYE = W_XE + NEI,       // y-edges (NEI; 2 NEI <= 2 NEF)
    W_LXA = W_G,              // adjoint of amdq of x-edge i+1, apdq of x-edge i (12)
    W_LYA = W_G + 12,         // same for y-edges j+1, j (12)
    W_LQ7 = W_G,              // adjoint of the loaded q of the cell (7)
    W_LS = W_QN               // adjoint of the split inputs of the cell state (6)
};
// The fields of W_LR: the F1 prepared fields a right (top) side reads, then
// the side's fsphi and the taudir of each direction.
enum {
    LR_H, LR_HU, LR_HV, LR_HM1, LR_P1, LR_HCHI1, LR_U, LR_V, LR_M1, LR_CHI1, LR_BDIF,
    LR_RHO1, LR_TAU1, LR_FS, LR_TDX, LR_TDY, NLR
};
static_assert(NLR == 16, "W_LR holds NLR planes");
// The taudir inputs with a derivative (taudir_vjp): depth and bdif of each of
// the five cells, the centre's hu, hv, rho1 and tau1.
enum { LT_HU = 10, LT_HV = 11, LT_RHO = 12, LT_TAU = 13, NLT = 14 };
DC_HD int lt_h(int s) { return 2 * s; }
DC_HD int lt_bdif(int s) { return 2 * s + 1; }
static_assert(NLT <= 3 + 2 * NW, "taudir adjoint must fit the taudir and wave planes");
static_assert(2 * NEI <= 2 * NEF, "flux adjoint must fit the flux planes");

struct AdjArgs {
    const double* q_in;      // q_n (ghosts read only when bc[0] == 0)
    const double* aux;
    int maux, mx, my;
    double dx, dy, dt;
    int bc[4];               // boundary conditions applied on load (0: stored ghosts)
    dclaw_params P;
    const dclaw_params* dP;  // np parameter directions
    int np;
    const double* lam_out;   // lambda_{n+1}, layout of q
    double* lam_in;          // lambda_n, layout of q
    double* q_out;           // q_{n+1} (the record pass; optional)
    double* W;               // workspace, W_FIELDS planes
    size_t N;                // plane size, plane_size(mx, my)
    int ld;                  // plane row stride, plane_ld(mx)
    const dclaw_params* Pd;  // device copies of P and dt read by each GPU thread (or null)
    const double* dt_dev;
    // Tile-sparse stepping (the GPU's tile passes): the marks of the tiles
    // the passes run this call (TILE_W x TILE_H cells, tntx per row; a
    // cell (i, j) is in tile ((j + 1) / TILE_H) * tntx + (i + 1) / TILE_W),
    // or null when every cell runs every stage (the host emulator).
    const int* tflags = nullptr;
    int tntx = 0;
};

constexpr int TILE_W = 16;   // tile of the sparse passes: one wavefront of cells
constexpr int TILE_H = 4;

DC_HD int plane_ld(int mx) { return (mx + 4 + 15) / 16 * 16; }
DC_HD size_t plane_size(int mx, int my) { return (size_t)plane_ld(mx) * (size_t)(my + 4); }

DC_HD size_t cix(const AdjArgs& A, int i, int j)
{
    return (size_t)(i + 1) + (size_t)A.ld * (size_t)(j + 1);
}
DC_HD double* plane(const AdjArgs& A, int f) { return A.W + (size_t)f * A.N; }
DC_HD double wv(const AdjArgs& A, int f, size_t n) { return A.W[(size_t)f * A.N + n]; }
DC_HD void ws(const AdjArgs& A, int f, size_t n, double v) { A.W[(size_t)f * A.N + n] = v; }
DC_HD bool in_grid(const AdjArgs& A, int i, int j)
{
    return i >= -1 && i <= A.mx + 2 && j >= -1 && j <= A.my + 2;
}
DC_HD bool interior(const AdjArgs& A, int i, int j)
{
    return i >= 1 && i <= A.mx && j >= 1 && j <= A.my;
}
DC_HD double bed(const AdjArgs& A, int i, int j) { return A.aux[aidx(i, j, A.mx, A.maux)]; }

DC_HD TP zero_params(const dclaw_params& P)
{
    dclaw_params z = {};
    return dclaw_tl::make_params(P, z);
}

// ---------------------------------------------------------------- stages

// F1: q of the cell (after the boundary map) -> prepared fields.
// The record types of the namespace that evaluates a stage over R (dclaw_tl
// for one tangent, dclaw_tlk for K).
template <class R> struct PhysOf;
template <> struct PhysOf<double> {
    using RowCell = dclaw::RowCell;
    using Cell = dclaw::Cell;
    using RhoTau = dclaw::RhoTau;
};
template <> struct PhysOf<Dual> {
    using RowCell = dclaw_tl::RowCell;
    using Cell = dclaw_tl::Cell;
    using RhoTau = dclaw_tl::RhoTau;
};
template <> struct PhysOf<dclaw_tlk::DualK> {
    using RowCell = dclaw_tlk::RowCell;
    using Cell = dclaw_tlk::Cell;
    using RhoTau = dclaw_tlk::RhoTau;
};

struct PrepF {
    double bed;
    template <class PP, class R>
    DC_HD void operator()(const PP& P, const R* x, R* y) const
    {
        typename PhysOf<R>::RowCell r;
        prepare_cell(P, x, R(bed), r);
        y[PC_H] = r.h; y[PC_HU] = r.hu; y[PC_HV] = r.hv; y[PC_BDIF] = r.bdif;
        y[PC_HM0] = r.hm0; y[PC_P0] = r.p0; y[PC_HCHI0] = r.hchi0;
        y[PC_HM1] = r.hm1; y[PC_P1] = r.p1; y[PC_HCHI1] = r.hchi1; y[PC_M1] = r.m1;
        y[PC_CHI1] = r.chi1; y[PC_U] = r.u; y[PC_V] = r.v;
        y[PC_HM2] = r.hm2; y[PC_P2] = r.p2; y[PC_HCHI2] = r.hchi2; y[PC_M2] = r.m2;
        y[PC_CHI2] = r.chi2;
        y[PC_RHO1] = y[PC_TAU1] = y[PC_RHO2] = y[PC_TAU2] = R(0.0);
        if (r.h > P.dry_tolerance) {
            const auto v1 = rho_tau(P, r.h, r.u, r.v, r.m1, r.p1, r.chi1);
            const auto v2 = rho_tau(P, r.h, r.u, r.v, r.m2, r.p2, r.chi2);
            y[PC_RHO1] = v1.rho; y[PC_TAU1] = v1.tau;
            y[PC_RHO2] = v2.rho; y[PC_TAU2] = v2.tau;
        }
    }
};

// F2: calc_taudir of the centre from 5 Q0 states (each 6 values and bdif) and
// setvars(F1) of the centre.  Slot s (0 centre, 1 left, 2 right, 3 bottom,
// 4 top) holds inputs 7s..7s+6; 35, 36 are rho1, tau1.
struct TaudirF {
    double dx, dy;
    double b[5];
    template <class PP, class R>
    DC_HD void operator()(const PP& P, const R* x, R* y) const
    {
        typename PhysOf<R>::Cell c[5];
        R bb[5];
        for (int s = 0; s < 5; s++) {
            c[s].h = x[7 * s + 0]; c[s].hu = x[7 * s + 1]; c[s].hv = x[7 * s + 2];
            c[s].hm = x[7 * s + 3]; c[s].p = x[7 * s + 4]; c[s].hchi = x[7 * s + 5];
            bb[s] = R(b[s]) - x[7 * s + 6];
        }
        typename PhysOf<R>::RhoTau sv;
        sv.rho = x[35];
        sv.tau = x[36];
        const auto t = calc_taudir_cell(P, R(dx), R(dy), c[0], bb[0], c[1], bb[1], c[2], bb[2],
                                        c[3], bb[3], c[4], bb[4], sv);
        y[0] = t.fsphi;
        y[1] = t.tx;
        y[2] = t.ty;
    }
};

DC_HD void side_from(const Dual* x, double bedv, dclaw_tl::Side& s)
{
    s.h = x[SD_H]; s.hu = x[SD_HU]; s.hv = x[SD_HV]; s.hm = x[SD_HM]; s.p = x[SD_P];
    s.hchi = x[SD_HCHI]; s.u = x[SD_U]; s.v = x[SD_V]; s.m = x[SD_M]; s.chi = x[SD_CHI];
    s.b = Dual(bedv) - x[SD_BDIF];
    s.taudir = x[SD_TAUDIR]; s.fs = x[SD_FS]; s.rho = x[SD_RHO]; s.tau = x[SD_TAU];
}

DC_HD void side_from_d(const double* x, double bedv, dclaw::Side& s)
{
    s.h = x[SD_H]; s.hu = x[SD_HU]; s.hv = x[SD_HV]; s.hm = x[SD_HM]; s.p = x[SD_P];
    s.hchi = x[SD_HCHI]; s.u = x[SD_U]; s.v = x[SD_V]; s.m = x[SD_M]; s.chi = x[SD_CHI];
    s.b = bedv - x[SD_BDIF];
    s.taudir = x[SD_TAUDIR]; s.fs = x[SD_FS]; s.rho = x[SD_RHO]; s.tau = x[SD_TAU];
}

// F3: rpn2 of one edge from its two sides (inputs 0..14 left, 15..29 right).
struct Rpn2F {
    int ixy;
    double bedL, bedR;
    DC_HD void operator()(const TP& P, const Dual* x, Dual* y) const
    {
        dclaw_tl::Side L, R;
        side_from(x, bedL, L);
        side_from(x + NSD, bedR, R);
        const dclaw_tl::EdgeWaves w = dclaw_tl::rpn2_edge(P, ixy, L, R);
        for (int k = 0; k < 3; k++) y[k] = w.s[k];
        for (int k = 0; k < 6; k++) y[3 + k] = w.w1[k];
        for (int k = 0; k < 6; k++) y[9 + k] = w.w5[k];
        y[15] = w.w2h;
        y[16] = w.w2n;
    }
    DC_HD void operator()(const dclaw_params& P, const double* x, double* y) const
    {
        dclaw::Side L, R;
        side_from_d(x, bedL, L);
        side_from_d(x + NSD, bedR, R);
        const dclaw::EdgeWaves w = dclaw::rpn2_edge(P, ixy, L, R);
        for (int k = 0; k < 3; k++) y[k] = w.s[k];
        for (int k = 0; k < 6; k++) y[3 + k] = w.w1[k];
        for (int k = 0; k < 6; k++) y[9 + k] = w.w5[k];
        y[15] = w.w2h;
        y[16] = w.w2n;
    }
};

template <class WavesT, class R>
DC_HD void waves_from(const R* x, WavesT& w)
{
    for (int k = 0; k < 3; k++) w.s[k] = x[k];
    for (int k = 0; k < 6; k++) w.w1[k] = x[3 + k];
    for (int k = 0; k < 6; k++) w.w5[k] = x[9 + k];
    w.w2h = x[15];
    w.w2n = x[16];
}

// F4: flux2's edge body from the edge's waves and the limiter's dot products.
struct FluxF {
    int ixy;
    bool gate;
    double dtdx;
    DC_HD void operator()(const TP& P, const Dual* x, Dual* y) const
    {
        dclaw_tl::EdgeWaves w;
        waves_from(x, w);
        const dclaw_tl::EdgeFlux e = dclaw_tl::edge_flux(P, ixy, w, x + NW, x + NW + 3, gate,
                                                         Dual(dtdx));
        for (int k = 0; k < 6; k++) {
            y[k] = e.faddm[k];
            y[6 + k] = e.faddp[k];
            y[12 + k] = e.amdq[k];
            y[18 + k] = e.apdq[k];
        }
    }
    DC_HD void operator()(const dclaw_params& P, const double* x, double* y) const
    {
        dclaw::EdgeWaves w;
        waves_from(x, w);
        const dclaw::EdgeFlux e = dclaw::edge_flux(P, ixy, w, x + NW, x + NW + 3, gate, dtdx);
        for (int k = 0; k < 6; k++) {
            y[k] = e.faddm[k];
            y[6 + k] = e.faddp[k];
            y[12 + k] = e.amdq[k];
            y[18 + k] = e.apdq[k];
        }
    }
};

// F5: one rpt2 split of a fluctuation in a cell.
struct SplitF {
    int ixy;
    double bedc, topo1, topo3;
    bool active;
    DC_HD void operator()(const TP& P, const Dual* x, Dual* y) const
    {
        dclaw_tl::rpt2_cell_split(P, ixy, x[0], x[1], x[2], x[3], x[4], Dual(bedc), x[5],
                                  Dual(topo1), Dual(topo3), x + 6, active, y, y + 6);
    }
    DC_HD void operator()(const dclaw_params& P, const double* x, double* y) const
    {
        dclaw::rpt2_cell_split(P, ixy, x[0], x[1], x[2], x[3], x[4], bedc, x[5], topo1, topo3,
                               x + 6, active, y, y + 6);
    }
};

// F6: src2 of the updated state (6 values and bdif).
struct Src2F {
    double dt, h_e;
    template <class PP, class R>
    DC_HD void operator()(const PP& P, const R* x, R* y) const
    {
        for (int k = 0; k < 7; k++) y[k] = x[k];
        src2_cell(P, R(dt), y[0], y[1], y[2], y[3], y[4], y[5], y[6], R(h_e));
    }
};

// Values of a stage: the step's own double arithmetic (the tangent types
// compute these values operation for operation).
template <int NI, int NO, class F>
DC_HD void evald(const F& f, const dclaw_params& P, const double* x, double* y)
{
    f(P, x, y);
}

// Values of a stage over the tangent type (tangents zero).
template <int NI, int NO, class F>
DC_HD void eval(const F& f, const TP& Pz, const double* x, double* y)
{
    Dual X[NI], Y[NO];
    for (int a = 0; a < NI; a++) X[a] = Dual(x[a], 0.0);
    f(Pz, X, Y);
    for (int o = 0; o < NO; o++) y[o] = Y[o].v;
}

template <int N>
DC_HD bool all_zero(const double* v)
{
    for (int k = 0; k < N; k++)
        if (v[k] != 0.0) return false;
    return true;
}

// lx = J^T ly for the stage's inputs (one seeded evaluation per input) and
// gp[k] += ly . dY/dP_k for the parameter directions.  Nothing is evaluated
// when ly is zero.
template <int NI, int NO, class F>
DC_HD void vjp(const AdjArgs& A, const F& f, const TP& Pz, const double* x, const double* ly,
               double* lx, double* gp)
{
    for (int k = 0; k < NI; k++) lx[k] = 0.0;
    if (all_zero<NO>(ly)) return;
#pragma unroll 1
    for (int k = 0; k < NI; k++) {
        Dual X[NI], Y[NO];
#pragma unroll
        for (int a = 0; a < NI; a++) X[a] = Dual(x[a], a == k ? 1.0 : 0.0);
        f(Pz, X, Y);
        double s = 0.0;
        for (int o = 0; o < NO; o++) s = s + ly[o] * Y[o].d;
        lx[k] = s;
    }
#pragma unroll 1
    for (int p = 0; p < A.np; p++) {
        const TP Pk = dclaw_tl::make_params(A.P, A.dP[p]);
        Dual X[NI], Y[NO];
        for (int a = 0; a < NI; a++) X[a] = Dual(x[a], 0.0);
        f(Pk, X, Y);
        double s = 0.0;
        for (int o = 0; o < NO; o++) s = s + ly[o] * Y[o].d;
        gp[p] += s;
    }
}

// The same products as vjp, K seeds per evaluation (dclaw_tlk): the NI input
// seeds, then one seed per parameter direction, dealt K at a time.  Each
// tangent component is bit for bit the tangent of vjp's evaluation with that
// seed, so lx and gp equal vjp's exactly.
// Only the first nis inputs are seeded (the caller writes the adjoint of an
// input that passes through the stage untouched); lx of the others is 0.
template <int NI, int NO, class F>
DC_HD void vjpk(const AdjArgs& A, const F& f, const double* x, const double* ly, double* lx,
                double* gp, int nis = NI)
{
    using dclaw_tlk::DualK;
    constexpr int KK = dclaw_tlk::K;
    for (int k = 0; k < NI; k++) lx[k] = 0.0;
    if (all_zero<NO>(ly)) return;
    const int ns = nis + A.np;
#pragma unroll 1
    for (int s0 = 0; s0 < ns; s0 += KK) {
        const dclaw_params* dp[KK];
        for (int k = 0; k < KK; k++) {
            const int sd = s0 + k;
            dp[k] = (sd >= nis && sd < ns) ? &A.dP[sd - nis] : nullptr;
        }
        const dclaw_tlk::Params Pk = dclaw_tlk::make_params(A.P, dp);
        DualK X[NI], Y[NO];
#pragma unroll
        for (int a = 0; a < NI; a++) {
            X[a] = DualK(x[a]);
#pragma unroll
            for (int k = 0; k < KK; k++) X[a].d[k] = (a == s0 + k && a < nis) ? 1.0 : 0.0;
        }
        f(Pk, X, Y);
#pragma unroll
        for (int k = 0; k < KK; k++) {
            const int sd = s0 + k;
            double t = 0.0;
#pragma unroll
            for (int o = 0; o < NO; o++) t = t + ly[o] * Y[o].d[k];
            // compile-time indices (predicated), so lx and gp stay in registers
#pragma unroll
            for (int a = 0; a < NI; a++)
                if (a == sd && sd < nis) lx[a] = t;
#pragma unroll
            for (int p = 0; p < MAXP; p++)
                if (p == sd - nis && sd < ns) gp[p] += t;
        }
    }
}

// ------------------------------------------------------- input gathering

// q of cell (i, j) as the step loads it (boundary map for a ghost cell when
// bc[0] != 0) and the momentum signs of the map.
DC_HD void load_q(const AdjArgs& A, int i, int j, double* q7, double& fu, double& fv)
{
    int si = i, sj = j;
    fu = 1.0;
    fv = 1.0;
    if (!interior(A, i, j) && A.bc[0] != 0) bc_source(A.bc, A.mx, A.my, i, j, si, sj, fu, fv);
    const size_t o = qidx(0, si, sj, A.mx);
    for (int m = 0; m < 7; m++) q7[m] = A.q_in[o + m];
    q7[1] = fu * q7[1];
    q7[2] = fv * q7[2];
}

DC_HD void taudir_inputs(const AdjArgs& A, int i, int j, double* x, TaudirF& f)
{
    const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
    for (int s = 0; s < 5; s++) {
        const size_t n = cix(A, i + di[s], j + dj[s]);
        x[7 * s + 0] = wv(A, W_PC + PC_H, n);
        x[7 * s + 1] = wv(A, W_PC + PC_HU, n);
        x[7 * s + 2] = wv(A, W_PC + PC_HV, n);
        x[7 * s + 3] = wv(A, W_PC + PC_HM0, n);
        x[7 * s + 4] = wv(A, W_PC + PC_P0, n);
        x[7 * s + 5] = wv(A, W_PC + PC_HCHI0, n);
        x[7 * s + 6] = wv(A, W_PC + PC_BDIF, n);
        f.b[s] = bed(A, i + di[s], j + dj[s]);
    }
    const size_t n = cix(A, i, j);
    x[35] = wv(A, W_PC + PC_RHO1, n);
    x[36] = wv(A, W_PC + PC_TAU1, n);
    f.dx = A.dx;
    f.dy = A.dy;
}

// The inputs taudir_vjp reads (depths, bdifs, the centre's hu, hv, rho1,
// tau1; the others 0).
DC_HD void taudir_inputs_live(const AdjArgs& A, int i, int j, double* x, TaudirF& f)
{
    const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
    for (int k = 0; k < NTI; k++) x[k] = 0.0;
    for (int s = 0; s < 5; s++) {
        const size_t n = cix(A, i + di[s], j + dj[s]);
        x[7 * s + 0] = wv(A, W_PC + PC_H, n);
        x[7 * s + 6] = wv(A, W_PC + PC_BDIF, n);
        f.b[s] = bed(A, i + di[s], j + dj[s]);
    }
    const size_t n = cix(A, i, j);
    x[1] = wv(A, W_PC + PC_HU, n);
    x[2] = wv(A, W_PC + PC_HV, n);
    x[35] = wv(A, W_PC + PC_RHO1, n);
    x[36] = wv(A, W_PC + PC_TAU1, n);
    f.dx = A.dx;
    f.dy = A.dy;
}

// rpn2 side of cell n: F2 state ("second") or F1; ixy = 2 swaps u and v.
DC_HD void side_inputs(const AdjArgs& A, size_t n, bool second, int ixy, double* s)
{
    s[SD_H] = wv(A, W_PC + PC_H, n);
    s[SD_HU] = wv(A, W_PC + PC_HU, n);
    s[SD_HV] = wv(A, W_PC + PC_HV, n);
    s[SD_HM] = wv(A, W_PC + (second ? PC_HM2 : PC_HM1), n);
    s[SD_P] = wv(A, W_PC + (second ? PC_P2 : PC_P1), n);
    s[SD_HCHI] = wv(A, W_PC + (second ? PC_HCHI2 : PC_HCHI1), n);
    s[SD_U] = wv(A, W_PC + (ixy == 1 ? PC_U : PC_V), n);
    s[SD_V] = wv(A, W_PC + (ixy == 1 ? PC_V : PC_U), n);
    s[SD_M] = wv(A, W_PC + (second ? PC_M2 : PC_M1), n);
    s[SD_CHI] = wv(A, W_PC + (second ? PC_CHI2 : PC_CHI1), n);
    s[SD_BDIF] = wv(A, W_PC + PC_BDIF, n);
    s[SD_TAUDIR] = wv(A, W_TD + (ixy == 1 ? 1 : 2), n);
    s[SD_FS] = wv(A, W_TD + 0, n);
    s[SD_RHO] = wv(A, W_PC + (second ? PC_RHO2 : PC_RHO1), n);
    s[SD_TAU] = wv(A, W_PC + (second ? PC_TAU2 : PC_TAU1), n);
}

// Edge ranges: x-edge (i, j) exists for i in [0, mx+2], j in [0, my+1]; its
// flux is used for i in [1, mx+1].  y-edge (i, k): k in [0, my+2], i in
// [0, mx+1]; flux for k in [1, my+1].
DC_HD bool xedge(const AdjArgs& A, int i, int j)
{
    return i >= 0 && i <= A.mx + 2 && j >= 0 && j <= A.my + 1;
}
DC_HD bool yedge(const AdjArgs& A, int i, int k)
{
    return k >= 0 && k <= A.my + 2 && i >= 0 && i <= A.mx + 1;
}
DC_HD bool xflux(const AdjArgs& A, int i, int j)
{
    return i >= 1 && i <= A.mx + 1 && j >= 0 && j <= A.my + 1;
}
DC_HD bool yflux(const AdjArgs& A, int i, int k)
{
    return k >= 1 && k <= A.my + 1 && i >= 0 && i <= A.mx + 1;
}

DC_HD void rpn2_inputs(const AdjArgs& A, int ixy, int i, int j, double* x, Rpn2F& f)
{
    // left/bottom cell, right/top cell
    const int li = ixy == 1 ? i - 1 : i, lj = ixy == 1 ? j : j - 1;
    const bool second = (ixy == 1 ? li : lj) >= 0;
    side_inputs(A, cix(A, li, lj), second, ixy, x);
    side_inputs(A, cix(A, i, j), false, ixy, x + NSD);
    f.ixy = ixy;
    f.bedL = bed(A, li, lj);
    f.bedR = bed(A, i, j);
}

DC_HD void wave_dots3(const double* a, const double* b, double* d3)
{
    double d = 0.0;
    for (int k = 0; k < 6; k++) d = d + a[3 + k] * b[3 + k];
    d3[0] = d;
    d = 0.0;
    d = d + a[15] * b[15];
    d = d + a[16] * b[16];
    d3[1] = d;
    d = 0.0;
    for (int k = 0; k < 6; k++) d = d + a[9 + k] * b[9 + k];
    d3[2] = d;
}

DC_HD void load_waves(const AdjArgs& A, int f0, size_t n, double* w)
{
    for (int k = 0; k < NW; k++) w[k] = wv(A, f0 + k, n);
}

DC_HD void flux_inputs(const AdjArgs& A, int ixy, int i, int j, double* x, FluxF& f)
{
    const int f0 = ixy == 1 ? W_XW : W_YW;
    const int di = ixy == 1 ? 1 : 0, dj = ixy == 1 ? 0 : 1;
    double wl[NW], wr[NW];
    load_waves(A, f0, cix(A, i - di, j - dj), wl);
    load_waves(A, f0, cix(A, i, j), x);
    load_waves(A, f0, cix(A, i + di, j + dj), wr);
    wave_dots3(wl, x, x + NW);
    wave_dots3(x, wr, x + NW + 3);
    // wet-dry gate h(-1) h(0) h(-2) h(+1) in flux2's product order
    const double hm1 = wv(A, W_PC + PC_H, cix(A, i - di, j - dj));
    const double h0 = wv(A, W_PC + PC_H, cix(A, i, j));
    const double hm2 = wv(A, W_PC + PC_H, cix(A, i - 2 * di, j - 2 * dj));
    const double hp1 = wv(A, W_PC + PC_H, cix(A, i + di, j + dj));
    f.ixy = ixy;
    f.gate = (hm1 * h0 * hm2 * hp1 > 0.0);
    f.dtdx = ixy == 1 ? A.dt / A.dx : A.dt / A.dy;
}

DC_HD void split_state(const AdjArgs& A, size_t n, double* x)
{
    x[0] = wv(A, W_PC + PC_H, n);
    x[1] = wv(A, W_PC + PC_HU, n);
    x[2] = wv(A, W_PC + PC_HV, n);
    x[3] = wv(A, W_PC + PC_HM2, n);
    x[4] = wv(A, W_PC + PC_HCHI2, n);
    x[5] = wv(A, W_PC + PC_BDIF, n);
}

// The four splits of cell (i, j): s = 0 x imp1 (amdq of x-edge i+1), 1 x imp2
// (apdq of x-edge i), 2 y imp1 (amdq of y-edge j+1), 3 y imp2 (apdq of y-edge
// j).  Returns false if the split does not take place.
DC_HD bool split_inputs(const AdjArgs& A, int s, int i, int j, double* x, SplitF& f)
{
    if (A.P.transverse == 0) return false;
    const size_t n = cix(A, i, j);
    split_state(A, n, x);
    f.bedc = bed(A, i, j);
    if (s < 2) {
        if (!(i >= 1 && i <= A.mx && j >= 0 && j <= A.my + 1)) return false;
        f.ixy = 1;
        f.topo1 = bed(A, i, j - 1);
        f.topo3 = bed(A, i, j + 1);
        f.active = true;
        const size_t e = s == 0 ? cix(A, i + 1, j) : n;
        const int o = s == 0 ? 12 : 18;
        for (int k = 0; k < 6; k++) x[6 + k] = wv(A, W_XE + o + k, e);
    } else {
        if (!(i >= 0 && i <= A.mx + 1 && j >= -1 && j <= A.my + 1)) return false;
        f.ixy = 2;
        f.topo1 = bed(A, i - 1, j);
        f.topo3 = bed(A, i + 1, j);
        const int k = s == 2 ? j + 1 : j;
        f.active = (k >= 1 && k <= A.my + 1);
        const size_t e = cix(A, i, k);
        const int o = s == 2 ? 12 : 18;
        for (int z = 0; z < 6; z++) x[6 + z] = wv(A, W_YE + o + z, e);
    }
    return true;
}

// ---------------------------------------------------------- record pass

DC_HD void rec_prepare(const AdjArgs& A, const TP& Pz, int i, int j)
{
    double q7[7], fu, fv, y[NPC];
    load_q(A, i, j, q7, fu, fv);
    PrepF f{bed(A, i, j)};
    evald<7, NPC>(f, A.P, q7, y);
    const size_t n = cix(A, i, j);
    for (int k = 0; k < NPC; k++) ws(A, W_PC + k, n, y[k]);
}

DC_HD void rec_taudir(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double y[3] = {0.0, 0.0, 0.0};
    if (i >= 0 && i <= A.mx + 1 && j >= 0 && j <= A.my + 1) {
        double x[NTI];
        TaudirF f;
        taudir_inputs(A, i, j, x, f);
        evald<NTI, 3>(f, A.P, x, y);
    }
    for (int k = 0; k < 3; k++) ws(A, W_TD + k, n, y[k]);
}

DC_HD void rec_waves(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    // both edges' inputs are loaded before the first store: every plane goes
    // through one pointer, so a load after a store stays behind it
    double x[2][2 * NSD];
    Rpn2F f[2];
    bool e[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const int ixy = d + 1;
        e[d] = ixy == 1 ? xedge(A, i, j) : yedge(A, i, j);
        if (e[d]) rpn2_inputs(A, ixy, i, j, x[d], f[d]);
    }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        double y[NW];
        for (int k = 0; k < NW; k++) y[k] = 0.0;
        if (e[d]) evald<2 * NSD, NW>(f[d], A.P, x[d], y);
        const int f0 = d == 0 ? W_XW : W_YW;
        for (int k = 0; k < NW; k++) ws(A, f0 + k, n, y[k]);
    }
}

DC_HD void rec_flux(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    // both edges' inputs before the first store (as rec_waves)
    double x[2][NEI];
    FluxF f[2];
    bool e[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const int ixy = d + 1;
        e[d] = ixy == 1 ? xflux(A, i, j) : yflux(A, i, j);
        if (e[d]) flux_inputs(A, ixy, i, j, x[d], f[d]);
    }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        double y[NEF];
        for (int k = 0; k < NEF; k++) y[k] = 0.0;
        if (e[d]) evald<NEI, NEF>(f[d], A.P, x[d], y);
        const int f0 = d == 0 ? W_XE : W_YE;
        for (int k = 0; k < NEF; k++) ws(A, f0 + k, n, y[k]);
    }
}

DC_HD void rec_split(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double b[4][12];
    for (int s = 0; s < 4; s++) {
        for (int k = 0; k < 12; k++) b[s][k] = 0.0;
        double x[NSI];
        SplitF f;
        if (split_inputs(A, s, i, j, x, f)) evald<NSI, 12>(f, A.P, x, b[s]);
    }
    // G = (0 - (dt/2dx) b_imp1) - (dt/2dx) b_imp2 (step2's transverse terms)
    const double dtdx = A.dt / A.dx, dtdy = A.dt / A.dy;
    for (int k = 0; k < 6; k++) {
        ws(A, W_G + k, n, (0.0 - 0.5 * dtdx * b[0][k]) - 0.5 * dtdx * b[1][k]);
        ws(A, W_G + 6 + k, n, (0.0 - 0.5 * dtdx * b[0][6 + k]) - 0.5 * dtdx * b[1][6 + k]);
        ws(A, W_G + 12 + k, n, (0.0 - 0.5 * dtdy * b[2][k]) - 0.5 * dtdy * b[3][k]);
        ws(A, W_G + 18 + k, n, (0.0 - 0.5 * dtdy * b[2][6 + k]) - 0.5 * dtdy * b[3][6 + k]);
    }
}

// The conservative update of cell (i, j) into q (the state src2 receives);
// false for a ghost cell, whose q_out (when wanted) is its prepared state.
DC_HD bool update_state(const AdjArgs& A, int i, int j, double* q)
{
    const size_t n = cix(A, i, j);
    q[0] = wv(A, W_PC + PC_H, n); q[1] = wv(A, W_PC + PC_HU, n); q[2] = wv(A, W_PC + PC_HV, n);
    q[3] = wv(A, W_PC + PC_HM0, n); q[4] = wv(A, W_PC + PC_P0, n);
    q[5] = wv(A, W_PC + PC_HCHI0, n); q[6] = wv(A, W_PC + PC_BDIF, n);
    if (!interior(A, i, j)) {
        if (A.q_out) {
            double* o = A.q_out + qidx(0, i, j, A.mx);
            for (int m = 0; m < 7; m++) o[m] = q[m];
        }
        return false;
    }
    const double dtdx = A.dt / A.dx, dtdy = A.dt / A.dy;
    const size_t nl = cix(A, i - 1, j), nr = cix(A, i + 1, j);
    const size_t nb = cix(A, i, j - 1), nt = cix(A, i, j + 1);
    for (int z = 0; z < 6; z++) {
        const double fm_next = ((0.0 + wv(A, W_XE + z, nr)) + wv(A, W_G + 18 + z, n)) +
                               wv(A, W_G + 12 + z, nr);
        const double fpc = ((0.0 + wv(A, W_XE + 6 + z, n)) + wv(A, W_G + 18 + z, nl)) +
                           wv(A, W_G + 12 + z, n);
        const double gm_next = ((0.0 + wv(A, W_G + 6 + z, n)) + wv(A, W_G + z, nt)) +
                               wv(A, W_YE + z, nt);
        const double gp_prev = ((0.0 + wv(A, W_G + 6 + z, nb)) + wv(A, W_G + z, n)) +
                               wv(A, W_YE + 6 + z, n);
        q[z] = q[z] - dtdx * (fm_next - fpc) - dtdy * (gm_next - gp_prev);
    }
    return true;
}

// q_out of an interior cell: src2 of its updated state q (only when the
// caller asked for q_out).
DC_HD void update_out(const AdjArgs& A, const TP& Pz, int i, int j, const double* q)
{
    if (!A.q_out) return;
    const double h_e = A.P.entrainment ? A.aux[aidx(i, j, A.mx, A.maux) + 6] : 0.0;
    Src2F f{A.dt, h_e};
    double y[7];
    // src2method 2 is not differentiated: the tangent type's src2 writes NaN,
    // and the record keeps that (the step itself integrates it)
    if (A.P.src2method == 2) eval<7, 7>(f, Pz, q, y);
    else evald<7, 7>(f, A.P, q, y);
    double* o = A.q_out + qidx(0, i, j, A.mx);
    for (int m = 0; m < 7; m++) o[m] = y[m];
}

DC_HD void rec_update(const AdjArgs& A, const TP& Pz, int i, int j)
{
    double q[7];
    if (!update_state(A, i, j, q)) return;
    const size_t n = cix(A, i, j);
    for (int m = 0; m < 7; m++) ws(A, W_QN + m, n, q[m]);
    update_out(A, Pz, i, j, q);
}

// --------------------------------------------------------- reverse pass

// adjoint of fm(i, j) + fp(i, j) and of gm(i, j) + gp(i, j) (each flux
// entry feeds one update: fm(i, j) cell i-1, fp(i, j) cell i, gm(i, j) cell
// j-1, gp(i, j) cell j)
DC_HD double lam_qn(const AdjArgs& A, int i, int j, int z)
{
    return interior(A, i, j) ? wv(A, W_LQN + z, cix(A, i, j)) : 0.0;
}
DC_HD double lam_fm(const AdjArgs& A, int i, int j, int z)
{
    return -(A.dt / A.dx) * lam_qn(A, i - 1, j, z);
}
DC_HD double lam_fp(const AdjArgs& A, int i, int j, int z)
{
    return (A.dt / A.dx) * lam_qn(A, i, j, z);
}
DC_HD double lam_gm(const AdjArgs& A, int i, int j, int z)
{
    return -(A.dt / A.dy) * lam_qn(A, i, j - 1, z);
}
DC_HD double lam_gp(const AdjArgs& A, int i, int j, int z)
{
    return (A.dt / A.dy) * lam_qn(A, i, j, z);
}

DC_HD void gp_add(const AdjArgs& A, size_t n, const double* g)
{
#pragma unroll
    for (int p = 0; p < MAXP; p++)
        if (p < A.np) ws(A, W_GP + p, n, wv(A, W_GP + p, n) + g[p]);
}

// B1: src2 (and the parameter gradient is started here).  WRITTEN selects
// the configurations src2_vjp covers (the written-out reverse) or the others
// (seeded products); a call for the other kind does nothing, so the two run
// as separate passes and each keeps only its own registers.
// The reverse of src2 at cell (i, j) given its updated state x (interior
// cells; x unused for ghost cells).
template <bool WRITTEN>
DC_HD void rev_src2_cell(const AdjArgs& A, int i, int j, const double* x)
{
    const size_t n = cix(A, i, j);
    double lx[7] = {0, 0, 0, 0, 0, 0, 0}, g[MAXP] = {0, 0, 0, 0};
    if (interior(A, i, j)) {
        double ly[7];
        for (int m = 0; m < 7; m++) ly[m] = A.lam_out[qidx(m, i, j, A.mx)];
        if (x[0] <= A.P.dry_tolerance) {
            // src2 zeroes a dry cell and leaves bdif (src2.f90:103-111)
            lx[6] = ly[6];
        } else {
            const double h_e = A.P.entrainment ? A.aux[aidx(i, j, A.mx, A.maux) + 6] : 0.0;
            if constexpr (WRITTEN) {
                // the written-out reverse (bdif passes through)
                ParAdj pa;
                pa_zero(pa);
                src2_vjp(A.P, A.dt, x, ly, lx, pa);
                lx[6] = ly[6];
#pragma unroll
                for (int p = 0; p < MAXP; p++)
                    if (p < A.np) g[p] = pa_dot(pa, A.dP[p]);
            } else {
                Src2F f{A.dt, h_e};
                // without ent_dclaw4's erosion src2 passes bdif through unchanged
                const bool erodes = A.P.entrainment == 1 && A.P.entrainment_method == 0;
                vjpk<7, 7>(A, f, x, ly, lx, g, erodes ? 7 : 6);
                if (!erodes) lx[6] = ly[6];
            }
        }
    }
    // the parameter gradient of the cell starts here (0 + g: the first term
    // of its sum)
#pragma unroll
    for (int p = 0; p < MAXP; p++) ws(A, W_GP + p, n, 0.0 + g[p]);
    for (int m = 0; m < 7; m++) ws(A, W_LQN + m, n, lx[m]);
}

template <bool WRITTEN>
DC_HD void rev_src2_t(const AdjArgs& A, const TP& Pz, int i, int j)
{
    if (src2_vjp_covers(A.P) != WRITTEN) return;
    const size_t n = cix(A, i, j);
    double x[7] = {0, 0, 0, 0, 0, 0, 0};
    if (interior(A, i, j))
        for (int m = 0; m < 7; m++) x[m] = wv(A, W_QN + m, n);
    rev_src2_cell<WRITTEN>(A, i, j, x);
}

DC_HD void rev_src2(const AdjArgs& A, const TP& Pz, int i, int j)
{
    rev_src2_t<true>(A, Pz, i, j);
    rev_src2_t<false>(A, Pz, i, j);
}

// F6 and B1 in one pass: the update (kept in registers instead of going
// through the QN planes), q_out when wanted, and src2's written-out reverse.
// A member whose src2 needs the seeded products gets its QN planes instead,
// and rev_src2_t<false> runs after.  Equals rec_update then rev_src2 bit for
// bit.
DC_HD void upd_src2(const AdjArgs& A, const TP& Pz, int i, int j)
{
    double q[7] = {0, 0, 0, 0, 0, 0, 0};
    const bool in = update_state(A, i, j, q);
    if (in) update_out(A, Pz, i, j, q);
    if (src2_vjp_covers(A.P)) {
        rev_src2_cell<true>(A, i, j, q);
        return;
    }
    if (in) {
        const size_t n = cix(A, i, j);
        for (int m = 0; m < 7; m++) ws(A, W_QN + m, n, q[m]);
    }
}

// Vector-Jacobian product of one rpt2 split (SplitF, rpt2_cell_split)
// written out: given the adjoints ly of bm[6], bp[6], the adjoints lx of
// (h, hu, hv, hm, hchi, bdif, asdq[6]) and, in gp[k], the parameter
// directions' contributions (the split reads grav, rho_s, rho_f).  The same
// branches as the split (dry cell, eta below both neighbours' bed, the sign
// of each speed and the eta test of its side); equals vjp<NSI, 12> of SplitF
// up to rounding.
template <int MU>
DC_HD void split_vjp_t(const AdjArgs& A, const SplitF& f, const double* x, const double* ly,
                       double* lx, double* gp)
{
    const dclaw_params& P = A.P;
    for (int k = 0; k < NSI; k++) lx[k] = 0.0;
    const double h = x[0];
    if (!f.active || !(h > P.dry_tolerance)) return;
    constexpr bool xd = (MU == 1);
    constexpr int mu = MU, mv = 3 - MU;
    const double nu = xd ? x[1] : x[2], nv = xd ? x[2] : x[1];
    const double u = nu / h, v = nv / h, m = x[3] / h, chi = x[4] / h;
    const double eta = h + f.bedc - x[5];
    const double g = P.grav;
    const double rho = m * P.rho_s + (1.0 - m) * P.rho_f;
    // the reverse divides by D, h, rho and 2 sq through their reciprocals
    // (one division each; an FP64 division is a chain of 11 instructions)
    const double rrho = 1.0 / rho;
    double gamma;
    if (P.src2method == -1) gamma = P.rho_f * rrho;
    else gamma = 0.25 * (P.rho_f + 3.0 * rho) * rrho;
    const double tmin = (f.topo1 < f.topo3) ? f.topo1 : f.topo3;
    if (eta < tmin) return;
    const double sq = sqrt(g * h);
    const double s[3] = {v - sq, v, v + sq};
    const double* asdq = x + 6;
    const double d1 = asdq[0], d2 = asdq[mu], d3 = asdq[mv];
    const double D = s[2] - s[0];
    const double rD = 1.0 / D;
    const double be[3] = {(s[2] * d1 - d3) * rD, -(u * d1) + d2, (d3 - s[0] * d1) * rD};
    const double r5 = gamma * rho * g;
    double r[3][6];
    r[0][0] = 1.0; r[0][mu] = u; r[0][mv] = s[0]; r[0][3] = m; r[0][4] = r5; r[0][5] = chi;
    r[2][0] = 1.0; r[2][mu] = u; r[2][mv] = s[2]; r[2][3] = m; r[2][4] = r5; r[2][5] = chi;
    for (int k = 0; k < 6; k++) r[1][k] = 0.0;
    r[1][mu] = 1.0;
    double ls[3] = {0, 0, 0}, lbe[3] = {0, 0, 0}, lr[3][6];
    for (int w = 0; w < 3; w++)
        for (int k = 0; k < 6; k++) lr[w][k] = 0.0;
#pragma unroll
    for (int w = 0; w < 3; w++) {
        // the wave goes to bm (ly[0..5]) or bp (ly[6..11]), or nowhere
        const bool to_m = s[w] < 0.0 && eta >= f.topo1;
        const bool to_p = !to_m && s[w] > 0.0 && eta >= f.topo3;
        if (!to_m && !to_p) continue;
        // out[k] += ((1 * s) * be) * r[k]
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const double lt = to_m ? ly[k] : ly[6 + k];
            ls[w] += lt * be[w] * r[w][k];
            lbe[w] += lt * s[w] * r[w][k];
            lr[w][k] += lt * s[w] * be[w];
        }
    }
    double lu = lr[0][mu] + lr[2][mu], lm = lr[0][3] + lr[2][3];
    double lr5 = lr[0][4] + lr[2][4], lchi = lr[0][5] + lr[2][5];
    ls[0] += lr[0][mv];
    ls[2] += lr[2][mv];
    // betas
    double ld1 = 0.0, ld2 = 0.0, ld3 = 0.0;
    ls[2] += lbe[0] * (d1 - be[0]) * rD;
    ls[0] += lbe[0] * be[0] * rD;
    ld1 += lbe[0] * s[2] * rD;
    ld3 += -lbe[0] * rD;
    lu += -(lbe[1] * d1);
    ld1 += -(lbe[1] * u);
    ld2 += lbe[1];
    ls[0] += lbe[2] * (be[2] - d1) * rD;
    ls[2] += -(lbe[2] * be[2]) * rD;
    ld3 += lbe[2] * rD;
    ld1 += -(lbe[2] * s[0]) * rD;
    lx[6 + 0] += ld1;
    lx[6 + mu] += ld2;
    lx[6 + mv] += ld3;
    // speeds: s1 = v - sq, s2 = v, s3 = v + sq; sq = sqrt(g h)
    const double lv = ls[0] + ls[1] + ls[2];
    const double lsq = ls[2] - ls[0];
    double lh = 0.0, lg = 0.0;
    if (sq > 0.0) {
        const double r2sq = 0.5 / sq;
        lh += lsq * g * r2sq;
        lg += lsq * h * r2sq;
    }
    // r5 = gamma rho g; gamma(rho, rho_f); rho = m rho_s + (1 - m) rho_f
    double lrho = lr5 * gamma * g, lrhof = 0.0;
    lg += lr5 * gamma * rho;
    const double lgam = lr5 * rho * g;
    if (P.src2method == -1) {
        lrho += -(lgam * gamma) * rrho;
        lrhof += lgam * rrho;
    } else {
        lrho += -(lgam * 0.25 * P.rho_f) * (rrho * rrho);
        lrhof += lgam * 0.25 * rrho;
    }
    lm += lrho * (P.rho_s - P.rho_f);
    const double lrhos = lrho * m;
    lrhof += lrho * (1.0 - m);
    // u = nu / h, v = nv / h, m = hm / h, chi = hchi / h
    const double rh = 1.0 / h;
    lx[xd ? 1 : 2] += lu * rh;
    lx[xd ? 2 : 1] += lv * rh;
    lx[3] += lm * rh;
    lx[4] += lchi * rh;
    lh += -(lu * u + lv * v + lm * m + lchi * chi) * rh;
    lx[0] += lh;
#pragma unroll
    for (int p = 0; p < MAXP; p++)
        if (p < A.np) gp[p] += lg * A.dP[p].grav + lrhos * A.dP[p].rho_s + lrhof * A.dP[p].rho_f;
}

// The normal momentum's slot (1 for an x-sweep split, 2 for y) as a
// compile-time constant, so no array is indexed at run time.
DC_HD void split_vjp(const AdjArgs& A, const SplitF& f, const double* x, const double* ly,
                     double* lx, double* gp)
{
    if (f.ixy == 1) split_vjp_t<1>(A, f, x, ly, lx, gp);
    else split_vjp_t<2>(A, f, x, ly, lx, gp);
}

// B2: the four transverse splits of the cell.  Every value the splits read
// is loaded before the first store: all planes share one pointer, so the
// compiler keeps a load that follows a store behind it, and each split's
// loads would wait for the previous split's stores.
DC_HD void rev_split(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    const bool tr = A.P.transverse != 0;
    // the x-splits (s = 0, 1) and y-splits (s = 2, 3) take place (split_inputs)
    const bool sx = tr && i >= 1 && i <= A.mx && j >= 0 && j <= A.my + 1;
    const bool sy = tr && i >= 0 && i <= A.mx + 1 && j >= -1 && j <= A.my + 1;
    double x[4][NSI];
    split_state(A, n, x[0]);
    double ax[2][6], ay[2][6];
    for (int z = 0; z < 6; z++) {
        ax[0][z] = sx ? wv(A, W_XE + 12 + z, cix(A, i + 1, j)) : 0.0;   // amdq of x-edge i+1
        ax[1][z] = sx ? wv(A, W_XE + 18 + z, n) : 0.0;                  // apdq of x-edge i
        ay[0][z] = sy ? wv(A, W_YE + 12 + z, cix(A, i, j + 1)) : 0.0;   // amdq of y-edge j+1
        ay[1][z] = sy ? wv(A, W_YE + 18 + z, n) : 0.0;                  // apdq of y-edge j
    }
    const double bc = bed(A, i, j);
    const double bjm = sx ? bed(A, i, j - 1) : 0.0, bjp = sx ? bed(A, i, j + 1) : 0.0;
    const double bim = sy ? bed(A, i - 1, j) : 0.0, bip = sy ? bed(A, i + 1, j) : 0.0;
    double lc[6], ljm[6], ljp[6], lim[6], lip[6], gp0[MAXP];
    for (int z = 0; z < 6; z++) {
        lc[z] = lam_qn(A, i, j, z);
        ljm[z] = sx ? lam_qn(A, i, j - 1, z) : 0.0;
        ljp[z] = sx ? lam_qn(A, i, j + 1, z) : 0.0;
        lim[z] = sy ? lam_qn(A, i - 1, j, z) : 0.0;
        lip[z] = sy ? lam_qn(A, i + 1, j, z) : 0.0;
    }
#pragma unroll
    for (int p = 0; p < MAXP; p++) gp0[p] = p < A.np ? wv(A, W_GP + p, n) : 0.0;

    for (int s = 1; s < 4; s++)
        for (int k = 0; k < 6; k++) x[s][k] = x[0][k];
    for (int k = 0; k < 6; k++) {
        x[0][6 + k] = ax[0][k];
        x[1][6 + k] = ax[1][k];
        x[2][6 + k] = ay[0][k];
        x[3][6 + k] = ay[1][k];
    }
    // adjoints of bm, bp of each split: Gx1(i, j) feeds gm(i, j) and gp(i, j),
    // Gx2(i, j) feeds gm, gp(i, j+1); Gy1(i, j) feeds fm(i, j) and fp(i, j),
    // Gy2(i, j) feeds fm, fp(i+1, j) (lam_gm, lam_gp, lam_fm, lam_fp)
    const double hx = 0.5 * (A.dt / A.dx), hy = 0.5 * (A.dt / A.dy);
    const double ry = A.dt / A.dy, rx = A.dt / A.dx;
    double lyx[12], lyy[12];
    for (int z = 0; z < 6; z++) {
        lyx[z] = -hx * (-ry * ljm[z] + ry * lc[z]);
        lyx[6 + z] = -hx * (-ry * lc[z] + ry * ljp[z]);
        lyy[z] = -hy * (-rx * lim[z] + rx * lc[z]);
        lyy[6 + z] = -hy * (-rx * lc[z] + rx * lip[z]);
    }
    double ls[6] = {0, 0, 0, 0, 0, 0};
    double g[MAXP] = {0, 0, 0, 0};
    double la[4][6];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        for (int k = 0; k < 6; k++) la[s][k] = 0.0;
        if (!(s < 2 ? sx : sy)) continue;
        SplitF f;
        f.bedc = bc;
        if (s < 2) {
            f.ixy = 1;
            f.topo1 = bjm;
            f.topo3 = bjp;
            f.active = true;
        } else {
            f.ixy = 2;
            f.topo1 = bim;
            f.topo3 = bip;
            const int k = s == 2 ? j + 1 : j;
            f.active = (k >= 1 && k <= A.my + 1);
        }
        double lx[NSI];
        split_vjp(A, f, x[s], s < 2 ? lyx : lyy, lx, g);
        for (int k = 0; k < 6; k++) ls[k] += lx[k];
        for (int k = 0; k < 6; k++) la[s][k] = lx[6 + k];
    }
    // adjoint of the split's asdq: amdq of x-edge i+1 / apdq of x-edge i
    // (s = 0, 1), the same for y-edges (s = 2, 3)
    for (int k = 0; k < 6; k++) {
        ws(A, W_LXA + k, n, la[0][k]);
        ws(A, W_LXA + 6 + k, n, la[1][k]);
        ws(A, W_LYA + k, n, la[2][k]);
        ws(A, W_LYA + 6 + k, n, la[3][k]);
    }
    for (int k = 0; k < 6; k++) ws(A, W_LS + k, n, ls[k]);
#pragma unroll
    for (int p = 0; p < MAXP; p++)
        if (p < A.np) ws(A, W_GP + p, n, gp0[p] + g[p]);
}

// d limiter_phi / dr on the branch the value takes (max/min ties as dmax and
// dmin resolve them: the second argument).
DC_HD double limiter_dphi(int mthlim, double r)
{
    switch (mthlim) {
    case 1: {
        const double m = (1.0 < r) ? 1.0 : r;
        const double dm = (1.0 < r) ? 0.0 : 1.0;
        return (0.0 > m) ? 0.0 : dm;
    }
    case 2: {
        const double i = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r;
        const double di = (1.0 < 2.0 * r) ? 0.0 : 2.0;
        const double a = (0.0 > i) ? 0.0 : i;
        const double da = (0.0 > i) ? 0.0 : di;
        const double b = (2.0 < r) ? 2.0 : r;
        const double db = (2.0 < r) ? 0.0 : 1.0;
        return (a > b) ? da : db;
    }
    case 3: {
        const double a = fabs(r), da = signbit(r) ? -1.0 : 1.0;
        const double den = 1.0 + a;
        const double q = (r + a) / den;
        return ((1.0 + da) - q * da) / den;
    }
    case 4: {
        const double c = (1.0 + r) / 2.0;
        const double m1 = (c < 2.0) ? c : 2.0;
        const double dm1 = (c < 2.0) ? 0.5 : 0.0;
        const double m2 = (m1 < 2.0 * r) ? m1 : 2.0 * r;
        const double dm2 = (m1 < 2.0 * r) ? dm1 : 2.0;
        return (0.0 > m2) ? 0.0 : dm2;
    }
    default: return 1.0;
    }
}

// Vector-Jacobian product of the flux2 edge body (FluxF, dclaw_strip_body.inc
// edge_flux) written out: the function is piecewise linear in the waves given
// the signs of the speeds, the limiter ratios and the wet-dry gate, so its
// transpose is exact and cheap.  x: the 23 inputs (waves, dotl, dotr); ly:
// adjoints of faddm, faddp, amdq, apdq; lx: adjoints of the inputs.  It
// equals vjp<NEI, NEF> of FluxF up to rounding (the derivatives are those of
// the branch the values take, as the tangent-linear evaluation defines them).
DC_HD void flux_vjp(const dclaw_params& P, const FluxF& f, const double* x, const double* ly,
                    double* lx)
{
    const int nrm = (f.ixy == 1) ? 1 : 2;
    const double* s = x;
    const double* w1 = x + 3;
    const double* w5 = x + 9;
    const double w2h = x[15], w2n = x[16];
    const double* dotl = x + NW;
    const double* dotr = x + NW + 3;
    for (int k = 0; k < NEI; k++) lx[k] = 0.0;
    double* ls = lx;
    double* lw1 = lx + 3;
    double* lw5 = lx + 9;
    double* ldotl = lx + NW;
    double* ldotr = lx + NW + 3;
    const bool fold = (P.order > 1 && P.transverse == 2);
    double lam[6], lap[6], lcq[6];
    for (int k = 0; k < 6; k++) {
        const double lfm = ly[k], lfp = ly[6 + k], lam_o = ly[12 + k], lap_o = ly[18 + k];
        lam[k] = lfm + lam_o;
        lap[k] = -lfp + lap_o;
        lcq[k] = 0.5 * lfm + 0.5 * lfp + (fold ? lam_o - lap_o : 0.0);
    }
    // fluctuations: each wave joins amdq or apdq by the sign of its speed
    if (s[0] < 0.0) { for (int k = 0; k < 6; k++) lw1[k] += lam[k]; }
    else if (s[0] > 0.0) { for (int k = 0; k < 6; k++) lw1[k] += lap[k]; }
    if (s[1] < 0.0) { lx[15] += lam[0]; lx[16] += lam[nrm]; }
    else if (s[1] > 0.0) { lx[15] += lap[0]; lx[16] += lap[nrm]; }
    if (s[2] < 0.0) { for (int k = 0; k < 6; k++) lw5[k] += lam[k]; }
    else if (s[2] > 0.0) { for (int k = 0; k < 6; k++) lw5[k] += lap[k]; }
    if (P.order == 1 || !f.gate) return;   // no correction: cq = 0
    // limiter
    double wn2[3];
    {
        double d = 0.0;
        for (int k = 0; k < 6; k++) d = d + w1[k] * w1[k];
        wn2[0] = d;
        d = 0.0;
        d = d + w2h * w2h;
        d = d + w2n * w2n;
        wn2[1] = d;
        d = 0.0;
        for (int k = 0; k < 6; k++) d = d + w5[k] * w5[k];
        wn2[2] = d;
    }
    const int fam_mw[3] = {0, 1, 4};
    double phi[3] = {1.0, 1.0, 1.0}, rr[3] = {0.0, 0.0, 0.0};
    bool lim_on[3];
    for (int q = 0; q < 3; q++) {
        const int lim = P.mthlim[fam_mw[q]];
        lim_on[q] = (lim != 0 && wn2[q] != 0.0);
        if (!lim_on[q]) continue;
        rr[q] = (s[q] > 0.0 ? dotl[q] : dotr[q]) / wn2[q];
        // limiter_phi's value, as dclaw::limiter_phi computes it
        phi[q] = dclaw::limiter_phi(lim, rr[q]);
    }
    // cq[k] = c0 l1[k] + [k = 0] c1 l2h + [k = nrm] c1 l2n + c2 l5[k]
    double c[3], dcds[3];
    for (int q = 0; q < 3; q++) {
        const double sg = copysign(1.0, s[q]);
        c[q] = sg * (1.0 - fabs(s[q]) * f.dtdx);
        dcds[q] = sg * (-((signbit(s[q]) ? -1.0 : 1.0) * f.dtdx));
    }
    double lc[3] = {0.0, 0.0, 0.0}, lphi[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 6; k++) {
        const double l1 = lim_on[0] ? phi[0] * w1[k] : w1[k];
        const double l5 = lim_on[2] ? phi[2] * w5[k] : w5[k];
        lc[0] += l1 * lcq[k];
        lc[2] += l5 * lcq[k];
        const double ll1 = c[0] * lcq[k], ll5 = c[2] * lcq[k];
        lw1[k] += lim_on[0] ? phi[0] * ll1 : ll1;
        lw5[k] += lim_on[2] ? phi[2] * ll5 : ll5;
        if (lim_on[0]) lphi[0] += w1[k] * ll1;
        if (lim_on[2]) lphi[2] += w5[k] * ll5;
    }
    {
        const double l2h = lim_on[1] ? phi[1] * w2h : w2h;
        const double l2n = lim_on[1] ? phi[1] * w2n : w2n;
        lc[1] += l2h * lcq[0] + l2n * lcq[nrm];
        const double ll2h = c[1] * lcq[0], ll2n = c[1] * lcq[nrm];
        lx[15] += lim_on[1] ? phi[1] * ll2h : ll2h;
        lx[16] += lim_on[1] ? phi[1] * ll2n : ll2n;
        if (lim_on[1]) lphi[1] += w2h * ll2h + w2n * ll2n;
    }
    for (int q = 0; q < 3; q++) ls[q] += lc[q] * dcds[q];
    // back through phi(r), r = dot / wn2, wn2 = sum of squares of the family
    for (int q = 0; q < 3; q++) {
        if (!lim_on[q] || lphi[q] == 0.0) continue;
        const double lr = lphi[q] * limiter_dphi(P.mthlim[fam_mw[q]], rr[q]);
        const double ldot = lr / wn2[q];
        const double lwn2 = -(lr * rr[q]) / wn2[q];
        if (s[q] > 0.0) ldotl[q] += ldot;
        else ldotr[q] += ldot;
        if (q == 0) {
            for (int k = 0; k < 6; k++) lw1[k] += 2.0 * w1[k] * lwn2;
        } else if (q == 1) {
            lx[15] += 2.0 * w2h * lwn2;
            lx[16] += 2.0 * w2n * lwn2;
        } else {
            for (int k = 0; k < 6; k++) lw5[k] += 2.0 * w5[k] * lwn2;
        }
    }
}

// B3: the flux body of the x- and y-edge at (i, j).  The edge body has no
// parameter dependence.  Both edges' inputs are loaded before the first
// store (as rec_waves).
DC_HD void rev_flux(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double x[2][NEI], ly[2][NEF];
    FluxF f[2];
    bool e[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const int ixy = d + 1;
        e[d] = ixy == 1 ? xflux(A, i, j) : yflux(A, i, j);
        if (!e[d]) continue;
        flux_inputs(A, ixy, i, j, x[d], f[d]);
        for (int z = 0; z < 6; z++) {
            if (ixy == 1) {
                ly[d][z] = lam_fm(A, i, j, z);
                ly[d][6 + z] = lam_fp(A, i, j, z);
                ly[d][12 + z] = wv(A, W_LXA + z, cix(A, i - 1, j));   // amdq: cell i-1's split
                ly[d][18 + z] = wv(A, W_LXA + 6 + z, n);              // apdq: cell i's split
            } else {
                ly[d][z] = lam_gm(A, i, j, z);
                ly[d][6 + z] = lam_gp(A, i, j, z);
                ly[d][12 + z] = wv(A, W_LYA + z, cix(A, i, j - 1));
                ly[d][18 + z] = wv(A, W_LYA + 6 + z, n);
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        double lx[NEI];
        for (int k = 0; k < NEI; k++) lx[k] = 0.0;
        if (e[d]) flux_vjp(A.P, f[d], x[d], ly[d], lx);
        const int f0 = d == 0 ? W_LXE : W_LYE;
        for (int k = 0; k < NEI; k++) ws(A, f0 + k, n, lx[k]);
    }
}

// Adjoint of the waves of an edge: its own flux body's, plus the limiter dot
// products of its flux body and of the two neighbouring edges'.
DC_HD void lam_waves(const AdjArgs& A, int ixy, int i, int j, double* lw)
{
    const int fw = ixy == 1 ? W_XW : W_YW, fl = ixy == 1 ? W_LXE : W_LYE;
    const int di = ixy == 1 ? 1 : 0, dj = ixy == 1 ? 0 : 1;
    const size_t n = cix(A, i, j);
    for (int k = 0; k < NW; k++) lw[k] = wv(A, fl + k, n);
    // dotl(e) = dots(w(e-1), w(e)) and dotr(e) = dots(w(e), w(e+1)); w(e-1)
    // multiplies the own dotl's and the left neighbour's dotr's adjoints,
    // w(e+1) the own dotr's and the right neighbour's dotl's
    double wl[NW], wr[NW];
    const bool has_l = in_grid(A, i - di, j - dj), has_r = in_grid(A, i + di, j + dj);
    for (int k = 0; k < NW; k++) wl[k] = wr[k] = 0.0;
    if (has_l) load_waves(A, fw, cix(A, i - di, j - dj), wl);
    if (has_r) load_waves(A, fw, cix(A, i + di, j + dj), wr);
    double cl[3], cr[3];
    for (int f = 0; f < 3; f++) {
        cl[f] = wv(A, fl + NW + f, n) +
                (has_l ? wv(A, fl + NW + 3 + f, cix(A, i - di, j - dj)) : 0.0);
        cr[f] = wv(A, fl + NW + 3 + f, n) +
                (has_r ? wv(A, fl + NW + f, cix(A, i + di, j + dj)) : 0.0);
    }
    for (int k = 0; k < 6; k++) lw[3 + k] += cl[0] * wl[3 + k] + cr[0] * wr[3 + k];
    lw[15] += cl[1] * wl[15] + cr[1] * wr[15];
    lw[16] += cl[1] * wl[16] + cr[1] * wr[16];
    for (int k = 0; k < 6; k++) lw[9 + k] += cl[2] * wl[9 + k] + cr[2] * wr[9 + k];
}

// B4: rpn2 of the x- and y-edge at (i, j).
// The inputs and wave adjoints of one edge's rpn2 product; false when the
// product is zero (no such edge, both sides dry: rpn2 returns constant zero
// waves, or a zero wave adjoint).
DC_HD bool wave_edge(const AdjArgs& A, int ixy, int i, int j, double* x, Rpn2F& f, double* ly)
{
    if (!(ixy == 1 ? xedge(A, i, j) : yedge(A, i, j))) return false;
    rpn2_inputs(A, ixy, i, j, x, f);
    const double drytol = A.P.dry_tolerance;
    if (!(x[SD_H] > drytol || x[NSD + SD_H] > drytol)) return false;
    lam_waves(A, ixy, i, j, ly);
    return !all_zero<NW>(ly);
}

// d sqrt(a) = da / (2 sqrt(a)), 0 where sqrt(a) = 0 (the tangent-linear
// convention): the adjoint of a from that of s = sqrt(a).
DC_HD double sqrt_adj(double s, double sb) { return s > 0.0 ? sb / (2.0 * s) : 0.0; }

// Adjoint of riemanntype's speeds s1m, s2m (riemannsolvers_dclaw.f:526-637)
// with respect to hL, hR, uL, uR and g, added to hLb ... gb.  hm enters rpn2
// only through these speeds and comparisons.  Same branches as riemanntype;
// in the mixed branch the term (h0 - h_min) sqrt(...) is 0 with derivative 0.
DC_HD void riemanntype_vjp(double hL, double hR, double uL, double uR, double drytol, double g,
                           double s1b, double s2b, double& hLb, double& hRb, double& uLb,
                           double& uRb, double& gb)
{
    if (s1b == 0.0 && s2b == 0.0) return;
    // s = sqrt(g a): adjoints of g and a from that of s
    auto sg = [&](double a, double sb, double& ab) {
        const double t = sqrt_adj(sqrt(g * a), sb);
        gb += t * a;
        ab += t * g;
    };
    const bool minR = hR < hL, maxR = hR > hL;
    const double h_min = minR ? hR : hL, h_max = maxR ? hR : hL;
    const double delu = uR - uL;
    double delub = 0.0, hminb = 0.0, hmaxb = 0.0;
    if (h_min <= drytol) {
        const double S = s1b + s2b;
        uRb += S;
        uLb += S;
        sg(hR, -2.0 * S, hRb);
        sg(hL, 2.0 * S, hLb);
        return;
    }
    const double F_min = delu + 2.0 * (sqrt(g * h_min) - sqrt(g * h_max));
    const double FQ = sqrt(.5 * g * (h_max + h_min) / (h_max * h_min));
    const double F_max = delu + (h_max - h_min) * FQ;
    if (F_min > 0.0) {
        const double arg = -delu + 2.0 * (sqrt(g * hL) + sqrt(g * hR));
        const double t = 0.0 > arg ? 0.0 : arg;
        const double c16 = 1.0 / (16.0 * g);
        const double hm = c16 * (t * t);
        uLb += s1b;
        uRb += s2b;
        double sLgb = 2.0 * s1b, sRgb = -2.0 * s2b, hmb = 0.0;
        sg(hm, -3.0 * s1b + 3.0 * s2b, hmb);
        const double c16b = hmb * (t * t);
        const double tb = hmb * c16 * 2.0 * t;
        gb += -(c16b * c16) / g;
        if (!(0.0 > arg)) {
            delub -= tb;
            sLgb += 2.0 * tb;
            sRgb += 2.0 * tb;
        }
        sg(hL, sLgb, hLb);
        sg(hR, sRgb, hRb);
    } else if (F_max <= 0.0) {
        const double h0 = h_max;
        const double iA = 1.0 / h0 + 1.0 / hL, iB = 1.0 / h0 + 1.0 / hR;
        const double gL = sqrt(.5 * g * iA), gR = sqrt(.5 * g * iB);
        const double dL = h0 - hL, dR = h0 - hR;
        const double F0 = delu + dL * gL + dR * gR;
        const double h02 = h0 * h0;
        const double DL = 4.0 * h02 * gL, DR = 4.0 * h02 * gR;
        const double dfdh = gL - g * dL / DL + gR - g * dR / DR;
        const double sq0 = sqrt(h0);
        const double slope = 2.0 * sq0 * dfdh;
        const double q = F0 / slope;
        const double t = sq0 - q;
        const double hm = t * t;
        const double iC = 1.0 / hm + 1.0 / hL, iD = 1.0 / hm + 1.0 / hR;
        const double qC = sqrt((.5 * g) * iC), qD = sqrt((.5 * g) * iD);
        // s1m = uL - (hm - hL) qC - sqrt(g hm), s2m = uR + (hm - hR) qD + sqrt(g hm)
        // the reverse divides through reciprocals (an FP64 division is a
        // chain of 11 instructions)
        const double rhm = 1.0 / hm, rhL = 1.0 / hL, rhR = 1.0 / hR, rh0 = 1.0 / h0;
        const double rslope = 1.0 / slope, rDL = 1.0 / DL, rDR = 1.0 / DR;
        double hmb = 0.0;
        sg(hm, -s1b + s2b, hmb);
        uRb += s2b;
        hmb += s2b * qD;
        hRb -= s2b * qD;
        const double qDb = s2b * (hm - hR);
        uLb += s1b;
        hmb -= s1b * qC;
        hLb += s1b * qC;
        const double qCb = -(s1b * (hm - hL));
        {
            const double tc = sqrt_adj(qC, qCb), td = sqrt_adj(qD, qDb);
            gb += tc * .5 * iC + td * .5 * iD;
            const double iCb = tc * .5 * g, iDb = td * .5 * g;
            hmb += -(iCb * rhm) * rhm - (iDb * rhm) * rhm;
            hLb += -(iCb * rhL) * rhL;
            hRb += -(iDb * rhR) * rhR;
        }
        // hm = t^2, t = sq0 - F0 / slope, slope = 2 sq0 dfdh
        const double tb = hmb * 2.0 * t;
        double sq0b = tb;
        const double F0b = -tb * rslope;
        const double slopeb = tb * q * rslope;
        sq0b += slopeb * 2.0 * dfdh;
        const double dfdhb = slopeb * 2.0 * sq0;
        double h0b = 0.0, gLb = dfdhb, gRb = dfdhb, dLb = 0.0, dRb = 0.0, h02b = 0.0;
        {
            // dfdh -= g dL / DL (and the same with R)
            const double TL = g * dL * rDL, TR = g * dR * rDR;
            const double nLb = -dfdhb * rDL, nRb = -dfdhb * rDR;
            const double DLb = dfdhb * TL * rDL, DRb = dfdhb * TR * rDR;
            gb += nLb * dL + nRb * dR;
            dLb += nLb * g;
            dRb += nRb * g;
            h02b += DLb * 4.0 * gL + DRb * 4.0 * gR;
            gLb += DLb * 4.0 * h02;
            gRb += DRb * 4.0 * h02;
        }
        h0b += h02b * 2.0 * h0;
        // F0 = delu + dL gL + dR gR
        delub += F0b;
        dLb += F0b * gL;
        gLb += F0b * dL;
        dRb += F0b * gR;
        gRb += F0b * dR;
        h0b += dLb + dRb;
        hLb -= dLb;
        hRb -= dRb;
        {
            const double ta = sqrt_adj(gL, gLb), tb2 = sqrt_adj(gR, gRb);
            gb += ta * .5 * iA + tb2 * .5 * iB;
            const double iAb = ta * .5 * g, iBb = tb2 * .5 * g;
            h0b += -(iAb * rh0) * rh0 - (iBb * rh0) * rh0;
            hLb += -(iAb * rhL) * rhL;
            hRb += -(iBb * rhR) * rhR;
        }
        h0b += sqrt_adj(sq0, sq0b);
        hmaxb += h0b;
    } else {
        const double h0 = h_min;
        const double sg0 = sqrt(g * h0), sgM = sqrt(g * h_max);
        const double F0 = delu + 2.0 * (sg0 - sgM)
                          + (h0 - h_min) * sqrt(.5 * g * (1.0 / h0 + 1.0 / h_min));
        const double dH = h_max - h_min;
        const double slope = (F_max - F0) / dH;
        const double q = F0 / slope;
        const double hm = h0 - q;
        double hmb = 0.0;
        if (hL > hR) {
            // s1m = uL + 2 sqrt(g hL) - 3 sqrt(g hm), s2m = uL + 2 sqrt(g hL) - sqrt(g hm)
            uLb += s1b + s2b;
            sg(hL, 2.0 * (s1b + s2b), hLb);
            sg(hm, -3.0 * s1b - s2b, hmb);
        } else {
            // s2m = uR - 2 sqrt(g hR) + 3 sqrt(g hm), s1m = uR - 2 sqrt(g hR) + sqrt(g hm)
            uRb += s1b + s2b;
            sg(hR, -2.0 * (s1b + s2b), hRb);
            sg(hm, 3.0 * s2b + s1b, hmb);
        }
        // hm = h0 - F0 / slope, slope = (F_max - F0) / dH
        double h0b = hmb;
        const double qb = -hmb;
        const double rslope = 1.0 / slope, rdH = 1.0 / dH;
        double F0b = qb * rslope;
        const double slopeb = -(qb * q) * rslope;
        const double nb = slopeb * rdH;
        const double F_maxb = nb;
        F0b -= nb;
        double dHb = -(slopeb * slope) * rdH;
        // F_max = delu + dH FQ, FQ = sqrt(.5 g (h_max + h_min) / (h_max h_min))
        delub += F_maxb;
        dHb += F_maxb * FQ;
        {
            const double t = sqrt_adj(FQ, F_maxb * dH);
            const double NQ = .5 * g * (h_max + h_min), DQ = h_max * h_min;
            const double rDQ = 1.0 / DQ;
            const double NQb = t * rDQ, DQb = -(t * (NQ * rDQ)) * rDQ;
            gb += NQb * .5 * (h_max + h_min);
            hmaxb += NQb * .5 * g + DQb * h_min;
            hminb += NQb * .5 * g + DQb * h_max;
        }
        hmaxb += dHb;
        hminb -= dHb;
        // F0 = delu + 2 (sqrt(g h0) - sqrt(g h_max))
        delub += F0b;
        sg(h0, 2.0 * F0b, h0b);
        sg(h_max, -2.0 * F0b, hmaxb);
        hminb += h0b;
    }
    if (minR) hRb += hminb;
    else hLb += hminb;
    if (maxR) hRb += hmaxb;
    else hLb += hmaxb;
    uRb += delub;
    uLb -= delub;
}

// Vector-Jacobian product of rpn2 (Rpn2F, rpn2_edge: rpn2_dclaw.f:125-315 and
// riemann_dig2_aug_sswave_ez, riemannsolvers_dclaw.f:3-521) written out for
// dry_tolerance > 0: given the adjoints ly of the 17 wave values, the
// adjoints lx of the 30 side inputs and, in gp[k], the parameter directions'
// contributions (rpn2 reads grav, rho_f, kappa and beta_seg).  The forward
// part replays rpn2_edge's arithmetic exactly, so every branch (dry resets,
// wall mirror, the rho/tau side choice, riemanntype's structure, the rare
// corrector, sonic, the source2dx and deldelh clamps, the static test,
// delta_huvm_revert's fold) is the one the step took; the tangent-linear
// conventions at kinks (max/min ties, sqrt at 0, copysign) are followed.
// Equals vjp<2 NSD, NW> of Rpn2F up to rounding.
DC_HD void rpn2_vjp(const AdjArgs& A, const Rpn2F& f, const double* x, const double* ly,
                    double* lx, double* gp)
{
    const dclaw_params& P = A.P;
    for (int k = 0; k < 2 * NSD; k++) lx[k] = 0.0;
    const double drytol = P.dry_tolerance, grav = P.grav;
    const double* XL = x;
    const double* XR = x + NSD;
    if (XL[SD_H] <= drytol && XR[SD_H] <= drytol) return;
    const bool xd = (f.ixy == 1);
    // effective side values after the dry resets / wall mirror; `sides`
    // records which inputs they come from (indices stay compile-time
    // constants, so the side arrays live in registers)
    double hL = XL[SD_H], hR = XR[SD_H];
    double huL = xd ? XL[SD_HU] : XL[SD_HV], huR = xd ? XR[SD_HU] : XR[SD_HV];  // normal
    double hvL = xd ? XL[SD_HV] : XL[SD_HU], hvR = xd ? XR[SD_HV] : XR[SD_HU];  // transverse
    double hmL = XL[SD_HM], hmR = XR[SD_HM], pL = XL[SD_P], pR = XR[SD_P];
    double bL = f.bedL - XL[SD_BDIF], bR = f.bedR - XR[SD_BDIF];
    double uL = XL[SD_U], uR = XR[SD_U], vL = XL[SD_V], vR = XR[SD_V];
    double mL = XL[SD_M], mR = XR[SD_M], chiL = XL[SD_CHI], chiR = XR[SD_CHI];
    double svrL = XL[SD_RHO], svtL = XL[SD_TAU], svrR = XR[SD_RHO], svtR = XR[SD_TAU];
    const double taudirR = XR[SD_TAUDIR], fsL = XL[SD_FS], fsR = XR[SD_FS];
    enum { OWN, RESET_R, WALL_R, RESET_L, WALL_L } sides = OWN;
    double wall0 = 1.0, wall1 = 1.0, wall2 = 1.0;
    bool wallprob = false;
    if (hR <= drytol) {
        hR = 0.0; pR = 0.0; mR = mL; chiR = chiL;
        sides = RESET_R;
        const dclaw::RType rt = dclaw::riemanntype(hL, hL, uL, -uL, drytol, grav);
        const double hstartest = dclaw::dmax(hL, rt.hm);
        if (hstartest + bL < bR) {
            wall1 = 0.0; wall2 = 0.0; wallprob = true;
            hR = hL; huR = -huL; hvR = hvL; hmR = hmL; bR = bL;
            uR = -uL; vR = vL; mR = mL; pR = pL; chiR = chiL; svrR = svrL; svtR = svtL;
            sides = WALL_R;
        }
    } else if (hL <= drytol) {
        hL = 0.0; pL = 0.0; mL = mR; chiL = chiR;
        sides = RESET_L;
        const dclaw::RType rt = dclaw::riemanntype(hR, hR, -uR, uR, drytol, grav);
        const double hstartest = dclaw::dmax(hR, rt.hm);
        if (hstartest + bR < bL) {
            wall0 = 0.0; wall1 = 0.0; wallprob = true;
            hL = hR; huL = -huR; hvL = hvR; hmL = hmR; mL = mR; bL = bR;
            uL = -uR; vL = vR; pL = pR; chiL = chiR; svrL = svrR; svtL = svtR;
            sides = WALL_L;
        }
    }

    // ---- forward replay of the solver
    const double rho_f = P.rho_f, kappa = P.kappa;
    const double criticaltol = dclaw::dmax(drytol * grav, 1e-6);
    const double criticaltol_2 = sqrt(criticaltol);
    const double gz = grav * cos(0.0);
    const double alpha_seg = 1.0 - P.beta_seg;
    int rtcase;
    double rhoL, rhoR, tauLs, tauRs, h;
    if (hL >= drytol && hR >= drytol) {
        rtcase = 0; rhoL = svrL; tauLs = svtL; rhoR = svrR; tauRs = svtR; h = 0.5 * (hL + hR);
    } else if (hL >= drytol) {
        rtcase = 1; rhoL = svrL; tauLs = svtL; rhoR = svrL; tauRs = 0.5 * tauLs; h = 0.5 * hL;
    } else {
        rtcase = 2; rhoL = svrR; rhoR = svrR; tauRs = svtR; tauLs = 0.5 * tauRs; h = 0.5 * hR;
    }
    const double tauL = fsL * tauLs, tauR = fsR * tauRs;
    const double rho = 0.5 * (rhoL + rhoR);
    const bool m1 = P.src2method == -1;
    double gamma, gammaL, gammaR;
    if (m1) {
        gamma = rho_f / rho; gammaL = rho_f / rhoL; gammaR = rho_f / rhoR;
    } else {
        gamma = 0.25 * (rho_f + 3.0 * rho) / rho;
        gammaL = 0.25 * (rho_f + 3.0 * rhoL) / rhoL;
        gammaR = 0.25 * (rho_f + 3.0 * rhoR) / rhoR;
    }
    const double eps = kappa + (1.0 - kappa) * gamma;
    const double geps = gz * eps;
    const double sqhL = sqrt(hL), sqhR = sqrt(hR);
    const double gsL = sqrt(geps * hL), gsR = sqrt(geps * hR);
    const double sL = uL - gsL, sR = uR + gsR;
    const double uhat = (sqhL * uL + sqhR * uR) / (sqhR + sqhL);
    const double chat = sqrt(geps * 0.5 * (hR + hL));
    const double sRoe1 = uhat - chat, sRoe2 = uhat + chat;
    const double sE1 = dclaw::dmin(sL, sRoe1), sE2 = dclaw::dmax(sR, sRoe2);
    const double u = uhat;
    const dclaw::RType rt = dclaw::riemanntype(hL, hR, uL, uR, drytol, geps);
    const double sw0 = dclaw::dmin(sE1, rt.s2m), sw2 = dclaw::dmax(sE2, rt.s1m);
    double sw1 = 0.5 * (sw2 + sw0);
    int sw1src = 0;
    const double hsQ = (huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1);
    const double hstarHLL = dclaw::dmax(hsQ, 0.0);
    bool rarecorrector = false;
    if (P.rarecorrectortest) {
        const double sdelta = sw2 - sw0;
        double raremin = 0.5;
        const double raremax = 0.9;
        if (rt.rare1 && sE1 * rt.s1m < 0.0) raremin = 0.2;
        if (rt.rare2 && sE2 * rt.s2m < 0.0) raremin = 0.2;
        if (rt.rare1 || rt.rare2) {
            const double rare1st = 3.0 * (sqrt(geps * hL) - sqrt(geps * rt.hm));
            const double rare2st = 3.0 * (sqrt(geps * hR) - sqrt(geps * rt.hm));
            if (dclaw::dmax(rare1st, rare2st) > raremin * sdelta &&
                dclaw::dmax(rare1st, rare2st) < raremax * sdelta) {
                rarecorrector = true;
                if (rare1st > rare2st) { sw1 = rt.s1m; sw1src = 1; }
                else if (rare2st > rare1st) { sw1 = rt.s2m; sw1src = 2; }
                else { sw1 = 0.5 * (rt.s1m + rt.s2m); sw1src = 3; }
            }
        }
        if (hstarHLL < dclaw::dmin(hL, hR) / 5.0) rarecorrector = false;
    }
    const double delb = bR - bL;
    const double hbar = 0.5 * (hL + hR);
    const double tuu = uL + uR;
    const double s1s2bar = 0.25 * (tuu * tuu) - gz * hbar;
    const double uLuR = uL * uR;
    const double s1s2tilde = dclaw::dmax(0.0, uLuR) - gz * hbar;
    bool sonic = false;
    if (fabs(s1s2bar) <= criticaltol) sonic = true;
    else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) sonic = true;
    else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) sonic = true;
    else if (dclaw::dmin(fabs(sE1), fabs(sE2)) < criticaltol_2) sonic = true;
    else if (sE1 < criticaltol_2 && rt.s1m > -criticaltol_2) sonic = true;
    else if (sE2 > -criticaltol_2 && rt.s2m < criticaltol_2) sonic = true;
    else if ((uL + sqrt(geps * hL)) * (uR + sqrt(geps * hR)) < 0.0) sonic = true;
    else if ((uL - sqrt(geps * hL)) * (uR - sqrt(geps * hR)) < 0.0) sonic = true;

    // source2dx: which expression it ends as (0/4: -(gz hbar delb), 1: the
    // non-sonic ratio, 2: gz max(-hL delb, -hR delb), 3: gz min(...))
    int s2src;
    double source2dx;
    if (sonic) { source2dx = -(gz * hbar * delb); s2src = 0; }
    else { source2dx = -(gz * hbar * delb * s1s2tilde / s1s2bar); s2src = 1; }
    const double mLd = -(hL * delb), mRd = -(hR * delb);
    const double cl1 = gz * dclaw::dmax(mLd, mRd);
    if (!(source2dx < cl1)) { source2dx = cl1; s2src = 2; }
    const double cl2 = gz * dclaw::dmin(mLd, mRd);
    if (!(source2dx > cl2)) { source2dx = cl2; s2src = 3; }
    if (fabs(u) <= 0.0) { source2dx = -(hbar * gz * delb); s2src = 4; }

    // deldelh: 0 -delb, 1 the non-sonic ratio, 2 hstarHLL dE / sE2,
    // 3 hstarHLL dE / sE1, 4 -hL, 5 hR
    int ddsrc;
    double deldelh;
    if (sonic) { deldelh = -delb; ddsrc = 0; }
    else { deldelh = delb * gz * hbar / s1s2bar; ddsrc = 1; }
    const double dE = sE2 - sE1;
    if (sE1 < -criticaltol && sE2 > criticaltol) {
        const double a1 = hstarHLL * dE / sE2;
        if (!(deldelh < a1)) { deldelh = a1; ddsrc = 2; }
        const double a2 = hstarHLL * dE / sE1;
        if (!(deldelh > a2)) { deldelh = a2; ddsrc = 3; }
    } else if (sE1 >= criticaltol) {
        const double a1 = hstarHLL * dE / sE1;
        if (!(deldelh < a1)) { deldelh = a1; ddsrc = 3; }
        if (!(deldelh > -hL)) { deldelh = -hL; ddsrc = 4; }
    } else if (sE2 <= -criticaltol) {
        if (!(deldelh < hR)) { deldelh = hR; ddsrc = 5; }
        const double a2 = hstarHLL * dE / sE2;
        if (!(deldelh > a2)) { deldelh = a2; ddsrc = 2; }
    }

    double del0 = hR - hL - deldelh;
    double del1 = huR - huL;
    double del2 = hR * (uR * uR) + 0.5 * kappa * gz * (hR * hR)
                  - (hL * (uL * uL) + 0.5 * kappa * gz * (hL * hL));
    del2 = del2 + (1.0 - kappa) * h * (pR - pL) / rho;
    const double del2pre = del2;
    const double vnorm = sqrt(uR * uR + uL * uL + vR * vR + vL * vL);
    const double T1 = taudirR * tauR / rhoR, T2 = tauL * taudirR / rhoL;
    bool is_static = false;
    int tscase;  // 0: tausource constant, 1: del2 - source2dx, 2: dsign(...)
    double ts0 = 0.0;
    if (vnorm > 0.0) {
        tscase = 0;
    } else if (0.5 * fabs(T1 + T2) > fabs(del2 - source2dx)) {
        tscase = 1;
        del1 = 0.0;
        del0 = 0.0;
        is_static = true;
    } else {
        tscase = 2;
        ts0 = 0.5 * (T1 + T2);
    }
    if (wallprob) tscase = 0;

    const double a = sw0, b = sw1, c = sw2;
    double beta1, beta2, beta3, R12, R22;
    if (!rarecorrector) {
        beta1 = (c * del0 - del1) / (c - a);
        beta2 = 0.0;  // its value is not needed
        beta3 = (del1 - a * del0) / (c - a);
        R12 = 0.0;
        R22 = 1.0;
    } else {
        beta1 = 0.0;
        beta2 = 0.0;
        beta3 = 0.0;
        R12 = sw1;
        R22 = sw1 * sw1;
    }
    double D1 = 0.0, D2 = 0.0, D3 = 0.0;
    if (rarecorrector) {
        // del2 after the tausource subtraction
        double tausource = 0.0;
        if (tscase == 1) tausource = del2pre - source2dx;
        else if (tscase == 2) tausource = dclaw::dsign(ts0, del2pre - source2dx);
        const double d2 = del2pre - source2dx - tausource;
        D1 = a * a - a * b - a * c + b * c;
        D2 = a * b - a * c - b * b + b * c;
        D3 = a * b - a * c - b * c + c * c;
        beta1 = (b * c * del0 - (b + c) * del1 + d2) / D1;
        beta2 = (-(a * c * del0) + (a + c) * del1 - d2) / D2;
        beta3 = (a * b * del0 - (a + b) * del1 + d2) / D3;
    } else {
        double tausource = 0.0;
        if (tscase == 1) tausource = del2pre - source2dx;
        else if (tscase == 2) tausource = dclaw::dsign(ts0, del2pre - source2dx);
        const double d2 = del2pre - source2dx - tausource;
        beta2 = a * c * del0 - (a + c) * del1 + d2;
    }
    const double f11 = beta1 * a, f13 = beta3 * c;
    const double AL = 1.0 + (1.0 - alpha_seg) * (1.0 - chiL);
    const double AR = 1.0 + (1.0 - alpha_seg) * (1.0 - chiR);
    const double hustar = huL + f11;
    const double u_i = is_static ? 0.0 : hustar / hstarHLL;
    const double K = gamma * rho * gz;

    // ---- reverse sweep (it divides through reciprocals: one division per
    // divisor, an FP64 division being a chain of 11 instructions; the forward
    // replay above keeps rpn2's own divisions, since they select branches)
    double hLb = 0, hRb = 0, huLb = 0, huRb = 0, hvLb = 0, hvRb = 0, hmLb = 0, hmRb = 0;
    double pLb = 0, pRb = 0, bLb = 0, bRb = 0, uLb = 0, uRb = 0, vLb = 0, vRb = 0;
    double mLb = 0, mRb = 0, chiLb = 0, chiRb = 0, svrLb = 0, svrRb = 0, svtLb = 0, svtRb = 0;
    double fsLb = 0, fsRb = 0, taudirRb = 0;
    double gravb = 0, rhofb = 0, kappab = 0, betasegb = 0;
    double gzb = 0, gammab = 0, gammaLb = 0, gammaRb = 0, rhob = 0, rhoLb = 0, rhoRb = 0;
    double tauLb = 0, tauRb = 0, hb = 0, hstarb = 0, sE1b = 0, sE2b = 0;
    // outputs (the wall factors are constants)
    const double* lw1 = ly + 3;
    const double* lw5 = ly + 9;
    double sw0b = ly[0] * wall0, sw1b = ly[1] * wall1, sw2b = ly[2] * wall2;
    double f11b = lw1[0] * wall0;
    const double f21b = (xd ? lw1[1] : lw1[2]) * wall0;
    const double f31Fb = (xd ? lw1[2] : lw1[1]) * wall0;
    const double f41Fb = lw1[3] * wall0, f51Fb = lw1[4] * wall0, f61Fb = lw1[5] * wall0;
    double f13b = lw5[0] * wall2;
    const double f23b = (xd ? lw5[1] : lw5[2]) * wall2;
    const double f33Fb = (xd ? lw5[2] : lw5[1]) * wall2;
    const double f43Fb = lw5[3] * wall2, f53Fb = lw5[4] * wall2, f63Fb = lw5[5] * wall2;
    const double f12b = ly[15] * wall1, f22b = ly[16] * wall1;
    // delta_huvm_revert's fold of the middle wave into wave 1 or 3
    double f31b = f31Fb, f33b = f33Fb, f41b = f41Fb, f43b = f43Fb;
    double f51b = f51Fb, f53b = f53Fb, f61b = f61Fb, f63b = f63Fb;
    const bool left = hustar <= 0.0;
    const double f32b = left ? f31Fb : f33Fb, f42b = left ? f41Fb : f43Fb;
    const double f52b = left ? f51Fb : f53Fb, f62b = left ? f61Fb : f63Fb;
    double del0b = 0, del1b = 0, del2b = 0;
    // f52 = del4 - f53 - f51, del4 = 0 (static) or recomputed with u_interface
    f53b -= f52b;
    f51b -= f52b;
    if (!is_static) {
        const double dhRL = hR - hL, dp = pR - pL;
        const double Kb = f52b * (-(u_i * dhRL) + del1);
        const double uib = f52b * (-(K * dhRL) + dp);
        hRb -= f52b * K * u_i;
        hLb += f52b * K * u_i;
        del1b += f52b * K;
        pRb += f52b * u_i;
        pLb -= f52b * u_i;
        gammab += Kb * rho * gz;
        rhob += Kb * gamma * gz;
        gzb += Kb * gamma * rho;
        const double rhs = 1.0 / hstarHLL;
        const double hsb = uib * rhs;
        hstarb -= uib * u_i * rhs;
        huLb += hsb;
        f11b += hsb;
    }
    // f62 = seg_R - seg_L - f61 - f63; f42, f32 likewise
    const double segRb = f62b, segLb = -f62b;
    f61b -= f62b;
    f63b -= f62b;
    hmRb += f42b * uR; uRb += f42b * hmR; hmLb -= f42b * uL; uLb -= f42b * hmL;
    f41b -= f42b;
    f43b -= f42b;
    hvRb += f32b * uR; uRb += f32b * hvR; hvLb -= f32b * uL; uLb -= f32b * hvL;
    f31b -= f32b;
    f33b -= f32b;
    double ALb = 0.0, ARb = 0.0;
    f11b += f61b * chiL * AL; chiLb += f61b * f11 * AL; ALb += f61b * f11 * chiL;
    f13b += f63b * chiR * AR; chiRb += f63b * f13 * AR; ARb += f63b * f13 * chiR;
    chiLb += segLb * hL * uL * AL; hLb += segLb * chiL * uL * AL;
    uLb += segLb * chiL * hL * AL; ALb += segLb * chiL * hL * uL;
    chiRb += segRb * hR * uR * AR; hRb += segRb * chiR * uR * AR;
    uRb += segRb * chiR * hR * AR; ARb += segRb * chiR * hR * uR;
    betasegb += ALb * (1.0 - chiL) + ARb * (1.0 - chiR);  // alpha_seg = 1 - beta_seg
    chiLb -= ALb * (1.0 - alpha_seg);
    chiRb -= ARb * (1.0 - alpha_seg);
    // f51 = f11 gammaL rhoL grav, f53 = f13 gammaR rhoR grav
    f11b += f51b * gammaL * rhoL * grav; gammaLb += f51b * f11 * rhoL * grav;
    rhoLb += f51b * f11 * gammaL * grav; gravb += f51b * f11 * gammaL * rhoL;
    f13b += f53b * gammaR * rhoR * grav; gammaRb += f53b * f13 * rhoR * grav;
    rhoRb += f53b * f13 * gammaR * grav; gravb += f53b * f13 * gammaR * rhoR;
    // f41 = f11 mL, f31 = f11 vL, f43 = f13 mR, f33 = f13 vR
    f11b += f41b * mL + f31b * vL; mLb += f41b * f11; vLb += f31b * f11;
    f13b += f43b * mR + f33b * vR; mRb += f43b * f13; vRb += f33b * f13;
    // f11 = beta1 a, f21 = beta1 a^2, f12 = beta2 R12, f22 = beta2 R22,
    // f13 = beta3 c, f23 = beta3 c^2
    const double beta1b = f11b * a + f21b * (a * a);
    double ab = f11b * beta1 + f21b * beta1 * 2.0 * a;
    const double beta2b = f12b * R12 + f22b * R22;
    const double R12b = f12b * beta2, R22b = f22b * beta2;
    const double beta3b = f13b * c + f23b * (c * c);
    double cb = f13b * beta3 + f23b * beta3 * 2.0 * c;
    double bb = 0.0;
    if (!rarecorrector) {
        const double rD = 1.0 / (c - a);
        const double N1b = beta1b * rD, N3b = beta3b * rD;
        const double Db = -(beta1b * beta1) * rD - (beta3b * beta3) * rD;
        cb += N1b * del0; del0b += N1b * c; del1b -= N1b;
        ab += beta2b * (c * del0 - del1); cb += beta2b * (a * del0 - del1);
        del0b += beta2b * a * c; del1b -= beta2b * (a + c); del2b += beta2b;
        del1b += N3b; ab -= N3b * del0; del0b -= N3b * a;
        cb += Db;
        ab -= Db;
    } else {
        const double rD1 = 1.0 / D1;
        const double N1b = beta1b * rD1, D1b = -(beta1b * beta1) * rD1;
        bb += N1b * (c * del0 - del1); cb += N1b * (b * del0 - del1);
        del0b += N1b * b * c; del1b -= N1b * (b + c); del2b += N1b;
        ab += D1b * (2.0 * a - b - c); bb += D1b * (c - a); cb += D1b * (b - a);
        const double rD2 = 1.0 / D2;
        const double N2b = beta2b * rD2, D2b = -(beta2b * beta2) * rD2;
        ab += N2b * (del1 - c * del0); cb += N2b * (del1 - a * del0);
        del0b -= N2b * a * c; del1b += N2b * (a + c); del2b -= N2b;
        ab += D2b * (b - c); bb += D2b * (a - 2.0 * b + c); cb += D2b * (b - a);
        const double rD3 = 1.0 / D3;
        const double N3b = beta3b * rD3, D3b = -(beta3b * beta3) * rD3;
        ab += N3b * (b * del0 - del1); bb += N3b * (a * del0 - del1);
        del0b += N3b * a * b; del1b -= N3b * (a + b); del2b += N3b;
        ab += D3b * (b - c); bb += D3b * (a - c); cb += D3b * (2.0 * c - a - b);
        bb += R12b + R22b * 2.0 * b;
    }
    if (is_static) {
        // del0 and del1 were replaced by constants
        del0b = 0.0;
        del1b = 0.0;
    }
    sw0b += ab;
    sw1b += bb;
    sw2b += cb;
    // del2 = del2pre - source2dx - tausource
    double del2preb = del2b, s2dxb = -del2b;
    const double tsb = -del2b;
    if (tscase == 1) {
        del2preb += tsb;
        s2dxb -= tsb;
    } else if (tscase == 2) {
        const double sgn = (signbit(ts0) ? -1.0 : 1.0) * (signbit(del2pre - source2dx) ? -1.0 : 1.0);
        const double T12b = 0.5 * tsb * sgn;
        const double rR = 1.0 / rhoR, rL = 1.0 / rhoL;
        taudirRb += T12b * tauR * rR + T12b * tauL * rL;
        tauRb += T12b * taudirR * rR;
        rhoRb -= T12b * T1 * rR;
        tauLb += T12b * taudirR * rL;
        rhoLb -= T12b * T2 * rL;
    }
    // del2pre
    {
        const double KG = 0.5 * kappa * gz;
        hRb += del2preb * (uR * uR + KG * 2.0 * hR);
        uRb += del2preb * hR * 2.0 * uR;
        hLb -= del2preb * (uL * uL + KG * 2.0 * hL);
        uLb -= del2preb * hL * 2.0 * uL;
        const double KGb = del2preb * (hR * hR - hL * hL);
        kappab += KGb * 0.5 * gz;
        gzb += KGb * 0.5 * kappa;
        const double dp = pR - pL, rr = 1.0 / rho;
        const double E = (1.0 - kappa) * h * dp * rr;
        kappab -= del2preb * h * dp * rr;
        hb += del2preb * (1.0 - kappa) * dp * rr;
        pRb += del2preb * (1.0 - kappa) * h * rr;
        pLb -= del2preb * (1.0 - kappa) * h * rr;
        rhob -= del2preb * E * rr;
    }
    huRb += del1b;
    huLb -= del1b;
    hRb += del0b;
    hLb -= del0b;
    const double deldelhb = -del0b;
    double delbb = 0.0, hbarb = 0.0, s1s2barb = 0.0, s1s2tildeb = 0.0;
    switch (ddsrc) {
    case 0: delbb -= deldelhb; break;
    case 1: {
        const double r = 1.0 / s1s2bar;
        const double v = delb * gz * hbar * r;
        delbb += deldelhb * gz * hbar * r;
        gzb += deldelhb * delb * hbar * r;
        hbarb += deldelhb * delb * gz * r;
        s1s2barb -= deldelhb * v * r;
        break;
    }
    case 2: {
        const double r = 1.0 / sE2;
        const double v = hstarHLL * dE * r, dEb = deldelhb * hstarHLL * r;
        hstarb += deldelhb * dE * r;
        sE2b += dEb - deldelhb * v * r;
        sE1b -= dEb;
        break;
    }
    case 3: {
        const double r = 1.0 / sE1;
        const double v = hstarHLL * dE * r, dEb = deldelhb * hstarHLL * r;
        hstarb += deldelhb * dE * r;
        sE2b += dEb;
        sE1b += -dEb - deldelhb * v * r;
        break;
    }
    case 4: hLb -= deldelhb; break;
    default: hRb += deldelhb; break;
    }
    switch (s2src) {
    case 1: {
        const double rb = 1.0 / s1s2bar;
        const double r = s1s2tilde * rb;
        const double w = gz * hbar * delb * s1s2tilde * rb;
        gzb -= s2dxb * hbar * delb * r;
        hbarb -= s2dxb * gz * delb * r;
        delbb -= s2dxb * gz * hbar * r;
        s1s2tildeb -= s2dxb * gz * hbar * delb * rb;
        s1s2barb += s2dxb * w * rb;
        break;
    }
    case 2:
    case 3: {
        const bool lsel = s2src == 2 ? (mLd > mRd) : (mLd < mRd);
        const double mv = lsel ? mLd : mRd;
        gzb += s2dxb * mv;
        const double mb = s2dxb * gz;
        if (lsel) { hLb -= mb * delb; delbb -= mb * hL; }
        else { hRb -= mb * delb; delbb -= mb * hR; }
        break;
    }
    default:
        gzb -= s2dxb * hbar * delb;
        hbarb -= s2dxb * gz * delb;
        delbb -= s2dxb * gz * hbar;
        break;
    }
    // s1s2bar = tuu^2 / 4 - gz hbar, s1s2tilde = max(0, uL uR) - gz hbar
    const double tuub = s1s2barb * 0.5 * tuu;
    gzb -= (s1s2barb + s1s2tildeb) * hbar;
    hbarb -= (s1s2barb + s1s2tildeb) * gz;
    if (!(0.0 > uLuR)) {
        uLb += s1s2tildeb * uR;
        uRb += s1s2tildeb * uL;
    }
    uLb += tuub;
    uRb += tuub;
    hLb += 0.5 * hbarb;
    hRb += 0.5 * hbarb;
    bRb += delbb;
    bLb -= delbb;
    // sw1 and the speeds
    double s1mb = 0.0, s2mb = 0.0;
    switch (sw1src) {
    case 0: sw2b += 0.5 * sw1b; sw0b += 0.5 * sw1b; break;
    case 1: s1mb += sw1b; break;
    case 2: s2mb += sw1b; break;
    default: s1mb += 0.5 * sw1b; s2mb += 0.5 * sw1b; break;
    }
    if (sE1 < rt.s2m) sE1b += sw0b;
    else s2mb += sw0b;
    if (sE2 > rt.s1m) sE2b += sw2b;
    else s1mb += sw2b;
    if (hsQ > 0.0) {
        const double rDq = 1.0 / (sE2 - sE1);
        const double Nb = hstarb * rDq, Dqb = -(hstarb * hsQ) * rDq;
        huLb += Nb;
        huRb -= Nb;
        sE2b += Nb * hR + Dqb;
        hRb += Nb * sE2;
        sE1b -= Nb * hL + Dqb;
        hLb -= Nb * sE1;
    }
    double gepsb = 0.0;
    riemanntype_vjp(hL, hR, uL, uR, drytol, geps, s1mb, s2mb, hLb, hRb, uLb, uRb, gepsb);
    double sLb = 0.0, sRb = 0.0, uhatb = 0.0, chatb = 0.0;
    if (sL < sRoe1) sLb += sE1b;
    else { uhatb += sE1b; chatb -= sE1b; }
    if (sR > sRoe2) sRb += sE2b;
    else { uhatb += sE2b; chatb += sE2b; }
    uLb += sLb;
    uRb += sRb;
    {
        const double tl = sqrt_adj(gsL, -sLb), tr = sqrt_adj(gsR, sRb);
        gepsb += tl * hL + tr * hR;
        hLb += tl * geps;
        hRb += tr * geps;
        const double tc = sqrt_adj(chat, chatb);
        gepsb += tc * 0.5 * (hR + hL);
        hRb += tc * geps * 0.5;
        hLb += tc * geps * 0.5;
    }
    {
        const double rDu = 1.0 / (sqhR + sqhL);
        const double Nb = uhatb * rDu, Dub = -(uhatb * uhat) * rDu;
        uLb += Nb * sqhL;
        uRb += Nb * sqhR;
        hLb += sqrt_adj(sqhL, Nb * uL + Dub);
        hRb += sqrt_adj(sqhR, Nb * uR + Dub);
    }
    // geps = gz eps, eps = kappa + (1 - kappa) gamma
    gzb += gepsb * eps;
    const double epsb = gepsb * gz;
    kappab += epsb * (1.0 - gamma);
    gammab += epsb * (1.0 - kappa);
    {
        const double rr = 1.0 / rho, rL = 1.0 / rhoL, rR = 1.0 / rhoR;
        if (m1) {
            rhofb += gammab * rr + gammaLb * rL + gammaRb * rR;
            rhob -= gammab * gamma * rr;
            rhoLb -= gammaLb * gammaL * rL;
            rhoRb -= gammaRb * gammaR * rR;
        } else {
            rhofb += 0.25 * (gammab * rr + gammaLb * rL + gammaRb * rR);
            rhob += gammab * (0.75 - gamma) * rr;
            rhoLb += gammaLb * (0.75 - gammaL) * rL;
            rhoRb += gammaRb * (0.75 - gammaR) * rR;
        }
    }
    rhoLb += 0.5 * rhob;
    rhoRb += 0.5 * rhob;
    fsLb += tauLb * tauLs;
    const double tauLsb = tauLb * fsL;
    fsRb += tauRb * tauRs;
    const double tauRsb = tauRb * fsR;
    gravb += gzb;
    switch (rtcase) {
    case 0:
        svrLb += rhoLb; svtLb += tauLsb; svrRb += rhoRb; svtRb += tauRsb;
        hLb += 0.5 * hb; hRb += 0.5 * hb;
        break;
    case 1:
        svrLb += rhoLb + rhoRb; svtLb += tauLsb + 0.5 * tauRsb; hLb += 0.5 * hb;
        break;
    default:
        svrRb += rhoLb + rhoRb; svtRb += tauRsb + 0.5 * tauLsb; hRb += 0.5 * hb;
        break;
    }
    // back to the side inputs: b = bed - bdif; a reset side's h and p are
    // constants and its m, chi the other side's; a mirrored side is the other
    // side with the normal momentum and velocity negated
    auto side = [&](double* l, double h_, double hu_, double hv_, double hm_, double p_,
                    double b_, double u_, double v_, double m_, double chi_, double svr_,
                    double svt_) {
        l[SD_H] += h_;
        if (xd) {
            l[SD_HU] += hu_;
            l[SD_HV] += hv_;
        } else {
            l[SD_HV] += hu_;
            l[SD_HU] += hv_;
        }
        l[SD_HM] += hm_;
        l[SD_P] += p_;
        l[SD_BDIF] -= b_;
        l[SD_U] += u_;
        l[SD_V] += v_;
        l[SD_M] += m_;
        l[SD_CHI] += chi_;
        l[SD_RHO] += svr_;
        l[SD_TAU] += svt_;
    };
    switch (sides) {
    case OWN:
        side(lx, hLb, huLb, hvLb, hmLb, pLb, bLb, uLb, vLb, mLb, chiLb, svrLb, svtLb);
        side(lx + NSD, hRb, huRb, hvRb, hmRb, pRb, bRb, uRb, vRb, mRb, chiRb, svrRb, svtRb);
        break;
    case RESET_R:
        side(lx, hLb, huLb, hvLb, hmLb, pLb, bLb, uLb, vLb, mLb + mRb, chiLb + chiRb, svrLb, svtLb);
        side(lx + NSD, 0.0, huRb, hvRb, hmRb, 0.0, bRb, uRb, vRb, 0.0, 0.0, svrRb, svtRb);
        break;
    case WALL_R:
        side(lx, hLb + hRb, huLb - huRb, hvLb + hvRb, hmLb + hmRb, pLb + pRb, bLb + bRb,
             uLb - uRb, vLb + vRb, mLb + mRb, chiLb + chiRb, svrLb + svrRb, svtLb + svtRb);
        break;
    case RESET_L:
        side(lx, 0.0, huLb, hvLb, hmLb, 0.0, bLb, uLb, vLb, 0.0, 0.0, svrLb, svtLb);
        side(lx + NSD, hRb, huRb, hvRb, hmRb, pRb, bRb, uRb, vRb, mRb + mLb, chiRb + chiLb, svrRb,
             svtRb);
        break;
    default:
        side(lx + NSD, hRb + hLb, huRb - huLb, hvRb + hvLb, hmRb + hmLb, pRb + pLb, bRb + bLb,
             uRb - uLb, vRb + vLb, mRb + mLb, chiRb + chiLb, svrRb + svrLb, svtRb + svtLb);
        break;
    }
    lx[SD_FS] += fsLb;
    lx[NSD + SD_FS] += fsRb;
    lx[NSD + SD_TAUDIR] += taudirRb;
#pragma unroll
    for (int p = 0; p < MAXP; p++)
        if (p < A.np)
            gp[p] += gravb * A.dP[p].grav + rhofb * A.dP[p].rho_f + kappab * A.dP[p].kappa +
                     betasegb * A.dP[p].beta_seg;
}

// B4: rpn2 of the x- and y-edge at (i, j): the hand-written product for
// dry_tolerance > 0, the seeded one otherwise (a reset dry side then reaches
// setvars).
DC_HD void rev_waves(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double g[MAXP] = {0, 0, 0, 0};
    double lr[NLR];
    for (int ixy = 1; ixy <= 2; ixy++) {
        double lx[2 * NSD];
        for (int k = 0; k < 2 * NSD; k++) lx[k] = 0.0;
        double x[2 * NSD], ly[NW];
        Rpn2F f;
        if (wave_edge(A, ixy, i, j, x, f, ly)) {
            if (A.P.dry_tolerance > 0.0) {
                rpn2_vjp(A, f, x, ly, lx, g);
            } else {
                // (its own arrays: the seeded product indexes them at run time)
                double ls[2 * NSD], gs[MAXP] = {0, 0, 0, 0};
                vjp<2 * NSD, NW>(A, f, Pz, x, ly, ls, gs);
                for (int k = 0; k < 2 * NSD; k++) lx[k] = ls[k];
                for (int p = 0; p < MAXP; p++) g[p] += gs[p];
            }
        }
        // the left (bottom) side belongs to the neighbour; the right (top)
        // side is this cell's F1 state for both edges
        const int f0 = ixy == 1 ? W_LXL : W_LYL;
        for (int k = 0; k < NSD; k++) ws(A, f0 + k, n, lx[k]);
        const double* r = lx + NSD;
        if (ixy == 1) {
            lr[LR_H] = r[SD_H]; lr[LR_HU] = r[SD_HU]; lr[LR_HV] = r[SD_HV];
            lr[LR_HM1] = r[SD_HM]; lr[LR_P1] = r[SD_P]; lr[LR_HCHI1] = r[SD_HCHI];
            lr[LR_U] = r[SD_U]; lr[LR_V] = r[SD_V]; lr[LR_M1] = r[SD_M];
            lr[LR_CHI1] = r[SD_CHI]; lr[LR_BDIF] = r[SD_BDIF]; lr[LR_RHO1] = r[SD_RHO];
            lr[LR_TAU1] = r[SD_TAU]; lr[LR_FS] = r[SD_FS]; lr[LR_TDX] = r[SD_TAUDIR];
        } else {
            // a y-edge's side has u and v swapped
            lr[LR_H] += r[SD_H]; lr[LR_HU] += r[SD_HU]; lr[LR_HV] += r[SD_HV];
            lr[LR_HM1] += r[SD_HM]; lr[LR_P1] += r[SD_P]; lr[LR_HCHI1] += r[SD_HCHI];
            lr[LR_U] += r[SD_V]; lr[LR_V] += r[SD_U]; lr[LR_M1] += r[SD_M];
            lr[LR_CHI1] += r[SD_CHI]; lr[LR_BDIF] += r[SD_BDIF]; lr[LR_RHO1] += r[SD_RHO];
            lr[LR_TAU1] += r[SD_TAU]; lr[LR_FS] += r[SD_FS]; lr[LR_TDY] = r[SD_TAUDIR];
        }
    }
    for (int k = 0; k < NLR; k++) ws(A, W_LR + k, n, lr[k]);
    gp_add(A, n, g);
}

// Field k of a neighbour's left (bottom) side adjoint, 0 outside the grid.
DC_HD double lside(const AdjArgs& A, int f0, int i, int j, int k)
{
    return in_grid(A, i, j) ? wv(A, f0 + k, cix(A, i, j)) : 0.0;
}

// calc_taudir's reverse written out (TaudirF: calc_taudir_cell on the five
// Q0 states, digclaw_module.f90:546-713).  Of its 37 inputs only the depths
// and bdif of the five cells, the centre's hu, hv and its rho1, tau1 enter
// the outputs (the qfix passes leave h, hu, hv as they are on a wet cell and
// zero them on a dry one); the rest get 0.  lx: adjoints of the 37 inputs;
// b_grav: the adjoint of grav (added).
DC_HD void taudir_vjp(const dclaw_params& P, double dx, double dy, const double* bed5,
                      const double* x, const double* ly, double* lx, double& b_grav)
{
    for (int k = 0; k < NTI; k++) lx[k] = 0.0;
    const double gz = P.grav, drytol = P.dry_tolerance;
    double hs[5], e0[5];
    bool wet[5];
    for (int s = 0; s < 5; s++) {
        wet[s] = x[7 * s] > drytol;
        hs[s] = wet[s] ? x[7 * s] : 0.0;
        e0[s] = hs[s] + (bed5[s] - x[7 * s + 6]);
    }
    const double h = hs[0];
    const double hu = wet[0] ? x[1] : 0.0, hv = wet[0] ? x[2] : 0.0;
    // neighbours' surfaces, limited to the centre's where they are dry
    double en[5];
    bool nlim[5];
    for (int s = 1; s < 5; s++) {
        en[s] = e0[s];
        nlim[s] = false;
        if (hs[s] < drytol) {
            nlim[s] = !(en[s] < e0[0]);   // dmin picks the centre's surface
            en[s] = nlim[s] ? e0[0] : en[s];
        }
    }
    // the centre's surface, limited by the left then the bottom one when dry
    double eta = e0[0];
    bool c1 = false, c2 = false;
    if (h < drytol) {
        c1 = en[1] < eta;
        eta = c1 ? en[1] : eta;
        c2 = en[3] < eta;
        eta = c2 ? en[3] : eta;
    }
    if ((h + hs[1] + hs[3] + hs[2] + hs[4]) < drytol) return;
    const double A = gz * 0.5;
    const double FxL = -(A * (h + hs[1]) * (eta - en[1]) / dx);
    const double FyL = -(A * (h + hs[3]) * (eta - en[3]) / dy);
    const double FxR = -(A * (h + hs[2]) * (en[2] - eta) / dx);
    const double FyR = -(A * (h + hs[4]) * (en[4] - eta) / dy);
    const bool fx_on = FxL * FxR > 0.0, fy_on = FyL * FyR > 0.0;
    const bool fx_l = fabs(FxL) < fabs(FxR), fy_l = fabs(FyL) < fabs(FyR);
    const double Fx = fx_on ? (fx_l ? FxL : FxR) : 0.0;
    const double Fy = fy_on ? (fy_l ? FyL : FyR) : 0.0;
    double bFx = 0.0, bFy = 0.0, bFxL = 0.0, bFyL = 0.0, bhu = 0.0, bhv = 0.0;
    const double hvn = sqrt(hu * hu + hv * hv);
    // The reverse divides through reciprocals (one division each; an FP64
    // division is a chain of 11 instructions).  The forward values that
    // select a branch (the F's, Fproj, X) keep calc_taudir's own divisions.
    if (hvn > 0.0) {
        double bhvn = 0.0;
        const double rhv = 1.0 / hvn;
        const double tx = -(dx * hu * rhv), ty = -(dy * hv * rhv);
        bhu -= ly[1] * dx * rhv;
        bhvn -= ly[1] * tx * rhv;
        bhv -= ly[2] * dy * rhv;
        bhvn -= ly[2] * ty * rhv;
        const double a1p = Fx * hu, a2p = Fy * hv;
        const double a1 = 0.0 > a1p ? 0.0 : a1p, a2 = 0.0 > a2p ? 0.0 : a2p;
        const double dot = a1 < a2 ? a1 : a2;
        if (dot > 0.0) {
            const double Fproj = dot / hvn;
            const double tau = x[36], rho = x[35];
            const double tmax = tau > 1e-16 ? tau : 1e-16;
            const double X = Fproj * rho / tmax;
            if (!(1.0 < X)) {
                const double bX = ly[0];
                const double rtm = 1.0 / tmax;
                const double bFproj = bX * rho * rtm;
                lx[35] = bX * Fproj * rtm;
                if (tau > 1e-16) lx[36] = -bX * X * rtm;
                const double bdot = bFproj * rhv;
                bhvn -= bFproj * Fproj * rhv;
                double ba1 = 0.0, ba2 = 0.0;
                if (a1 < a2) ba1 = bdot;
                else ba2 = bdot;
                if (!(0.0 > a1p)) {
                    bFx += ba1 * hu;
                    bhu += ba1 * Fx;
                }
                if (!(0.0 > a2p)) {
                    bFy += ba2 * hv;
                    bhv += ba2 * Fy;
                }
            }
        }
        bhu += bhvn * hu * rhv;
        bhv += bhvn * hv * rhv;
    } else {
        const double S1 = FxL * FxL + Fy * Fy, S2 = Fx * Fx + FyL * FyL;
        if (S1 > 0.0) {
            const double r1 = sqrt(S1), rr1 = 1.0 / r1;
            bFxL -= ly[1] * dx * rr1;
            const double br1 = ly[1] * dx * FxL * (rr1 * rr1);
            bFxL += br1 * FxL * rr1;
            bFy += br1 * Fy * rr1;
        }
        if (S2 > 0.0) {
            const double r2 = sqrt(S2), rr2 = 1.0 / r2;
            bFyL -= ly[2] * dy * rr2;
            const double br2 = ly[2] * dy * FyL * (rr2 * rr2);
            bFx += br2 * Fx * rr2;
            bFyL += br2 * FyL * rr2;
        }
    }
    double bFxR = 0.0, bFyR = 0.0;
    if (fx_on) {
        if (fx_l) bFxL += bFx;
        else bFxR += bFx;
    }
    if (fy_on) {
        if (fy_l) bFyL += bFy;
        else bFyR += bFy;
    }
    // F = -(A (h + h_n) d_eta / d); A = grav / 2
    double bh[5] = {0, 0, 0, 0, 0}, ben[5] = {0, 0, 0, 0, 0}, beta = 0.0;
    {
        const double rdx = 1.0 / dx, rdy = 1.0 / dy;
        const double sL = A * (h + hs[1]) * rdx, dL = eta - en[1];
        bh[0] -= bFxL * A * dL * rdx; bh[1] -= bFxL * A * dL * rdx;
        beta -= bFxL * sL; ben[1] += bFxL * sL;
        b_grav -= bFxL * 0.5 * (h + hs[1]) * dL * rdx;
        const double sB = A * (h + hs[3]) * rdy, dB = eta - en[3];
        bh[0] -= bFyL * A * dB * rdy; bh[3] -= bFyL * A * dB * rdy;
        beta -= bFyL * sB; ben[3] += bFyL * sB;
        b_grav -= bFyL * 0.5 * (h + hs[3]) * dB * rdy;
        const double sR = A * (h + hs[2]) * rdx, dR = en[2] - eta;
        bh[0] -= bFxR * A * dR * rdx; bh[2] -= bFxR * A * dR * rdx;
        ben[2] -= bFxR * sR; beta += bFxR * sR;
        b_grav -= bFxR * 0.5 * (h + hs[2]) * dR * rdx;
        const double sT = A * (h + hs[4]) * rdy, dT = en[4] - eta;
        bh[0] -= bFyR * A * dT * rdy; bh[4] -= bFyR * A * dT * rdy;
        ben[4] -= bFyR * sT; beta += bFyR * sT;
        b_grav -= bFyR * 0.5 * (h + hs[4]) * dT * rdy;
    }
    // the centre's limits (bottom, then left), then the neighbours'
    double be0[5] = {0, 0, 0, 0, 0};
    if (h < drytol) {
        if (c2) ben[3] += beta;
        else if (c1) ben[1] += beta;
        else be0[0] += beta;
    } else {
        be0[0] += beta;
    }
    for (int s = 1; s < 5; s++) {
        if (nlim[s]) be0[0] += ben[s];
        else be0[s] += ben[s];
    }
    for (int s = 0; s < 5; s++) {
        // e0 = h + (bed - bdif); the depth's derivative passes on a wet cell
        if (wet[s]) lx[7 * s] = bh[s] + be0[s];
        lx[7 * s + 6] = -be0[s];
    }
    if (wet[0]) {
        lx[1] = bhu;
        lx[2] = bhv;
    }
}

// B5: calc_taudir of the cell.
DC_HD void rev_taudir(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double lx[NTI];
    for (int k = 0; k < NTI; k++) lx[k] = 0.0;
    if (i >= 0 && i <= A.mx + 1 && j >= 0 && j <= A.my + 1) {
        double ly[3];
        // the cell is the right side of x-edge i, the left of x-edge i+1, the
        // right (top) side of y-edge j, the left of y-edge j+1
        ly[0] = (wv(A, W_LR + LR_FS, n) + lside(A, W_LXL, i + 1, j, SD_FS)) +
                lside(A, W_LYL, i, j + 1, SD_FS);
        ly[1] = wv(A, W_LR + LR_TDX, n) + lside(A, W_LXL, i + 1, j, SD_TAUDIR);
        ly[2] = wv(A, W_LR + LR_TDY, n) + lside(A, W_LYL, i, j + 1, SD_TAUDIR);
        double x[NTI], g[MAXP] = {0, 0, 0, 0};
        TaudirF f;
        taudir_inputs_live(A, i, j, x, f);
        // five dry states: calc_taudir returns zeros (its h-sum test, with
        // dry_tolerance > 0)
        const double dtol = A.P.dry_tolerance;
        const bool dry5 = dtol > 0.0 && x[0] <= dtol && x[7] <= dtol && x[14] <= dtol &&
                          x[21] <= dtol && x[28] <= dtol;
        if (!dry5) {
            double b_grav = 0.0;
            taudir_vjp(A.P, A.dx, A.dy, f.b, x, ly, lx, b_grav);
#pragma unroll
            for (int p = 0; p < MAXP; p++)
                if (p < A.np) g[p] = b_grav * A.dP[p].grav;
            gp_add(A, n, g);
        }
    }
    for (int s5 = 0; s5 < 5; s5++) {
        ws(A, W_LTD + lt_h(s5), n, lx[7 * s5]);
        ws(A, W_LTD + lt_bdif(s5), n, lx[7 * s5 + 6]);
    }
    ws(A, W_LTD + LT_HU, n, lx[1]);
    ws(A, W_LTD + LT_HV, n, lx[2]);
    ws(A, W_LTD + LT_RHO, n, lx[35]);
    ws(A, W_LTD + LT_TAU, n, lx[36]);
}

// Add a side's adjoint to the prepared fields it was read from.
DC_HD void side_to_pc(const double* ls, bool second, int ixy, double* lp)
{
    lp[PC_H] += ls[SD_H];
    lp[PC_HU] += ls[SD_HU];
    lp[PC_HV] += ls[SD_HV];
    lp[second ? PC_HM2 : PC_HM1] += ls[SD_HM];
    lp[second ? PC_P2 : PC_P1] += ls[SD_P];
    lp[second ? PC_HCHI2 : PC_HCHI1] += ls[SD_HCHI];
    lp[ixy == 1 ? PC_U : PC_V] += ls[SD_U];
    lp[ixy == 1 ? PC_V : PC_U] += ls[SD_V];
    lp[second ? PC_M2 : PC_M1] += ls[SD_M];
    lp[second ? PC_CHI2 : PC_CHI1] += ls[SD_CHI];
    lp[PC_BDIF] += ls[SD_BDIF];
    lp[second ? PC_RHO2 : PC_RHO1] += ls[SD_RHO];
    lp[second ? PC_TAU2 : PC_TAU1] += ls[SD_TAU];
}

// ------------------------------------------------------------------ prepare
// The prepare stage (PrepF: prepare_cell, dclaw_prepare_body.inc, then
// rho_tau of F1 and F2) on a wet loaded cell (h > dry_tolerance): b4step2's
// qfix (digclaw_module.f90:285-345) and speed limit (b4step2.f90:74-81),
// rpn2's two qfix passes F1 and F2, setvars of both.  lp: adjoints of the
// NPC prepared fields; lx: adjoints of h, hu, hv, hm, pb, hchi (bdif is the
// caller's: it enters as max(bdif, 0) only); pa: parameter adjoints (added).
DC_HD void prepare_vjp(const dclaw_params& P, const double* q7, const double* lp, double* lx,
                       ParAdj& pa)
{
    const double gz = P.grav;
    QfFwd q0;
    qfix_fwd(P, q7[0], q7[1], q7[2], q7[3], q7[4], q7[5], gz, q0);
    const double h = q0.h;
    const double hm0 = h * q0.m, p0 = q0.p, hchi0 = h * q0.chi;
    // speed limit
    const double n0 = sqrt((q7[1] * q7[1] + q7[2] * q7[2]));
    const double sp = n0 / h;
    const bool lim = h > 0.0 && sp > P.speed_limit;
    const double sratio = lim ? P.speed_limit / sp : 1.0;
    const double hu = lim ? q7[1] * sratio : q7[1], hv = lim ? q7[2] * sratio : q7[2];
    QfFwd f1, f2;
    qfix_fwd(P, h, hu, hv, hm0, p0, hchi0, gz, f1);
    const double hm1 = h * f1.m, p1 = f1.p, hchi1 = h * f1.chi;
    qfix_fwd(P, h, hu, hv, hm1, p1, hchi1, gz, f2);
    SvFwd s1, s2;
    setvars_fwd(P, h, f1.u, f1.v, f1.m, f1.p, f1.chi, gz, s1);
    setvars_fwd(P, h, f1.u, f1.v, f2.m, f2.p, f2.chi, gz, s2);
    double b_gz = 0.0;
    // setvars of F2 and F1 (rho, tau)
    double bh = lp[PC_H], bu = lp[PC_U], bv = lp[PC_V];
    double bm2 = lp[PC_M2], bp2 = lp[PC_P2], bchi2 = lp[PC_CHI2];
    setvars_rev(P, h, f1.u, f1.v, f2.m, f2.chi, gz, s2, lp[PC_RHO2], 0.0, 0.0, 0.0, 0.0,
                lp[PC_TAU2], bh, bu, bv, bm2, bp2, bchi2, b_gz, pa);
    double bm1 = lp[PC_M1], bp1 = lp[PC_P1], bchi1 = lp[PC_CHI1];
    setvars_rev(P, h, f1.u, f1.v, f1.m, f1.chi, gz, s1, lp[PC_RHO1], 0.0, 0.0, 0.0, 0.0,
                lp[PC_TAU1], bh, bu, bv, bm1, bp1, bchi1, b_gz, pa);
    // F2 (its h, hu, hv, u, v, rho are not read)
    double t_h = 0.0, b_hu = 0.0, b_hv = 0.0, t_hm = lp[PC_HM2], t_hchi = lp[PC_HCHI2];
    qfix_rev(P, f2, gz, t_h, b_hu, b_hv, t_hm, bp2, t_hchi, 0.0, 0.0, bm2, bchi2, 0.0, b_gz,
             pa);
    bh += t_h;
    // F1: hm1, p1, hchi1 feed F2 and are outputs
    double u_h = 0.0, b_hm0 = lp[PC_HM1] + t_hm, b_hchi0 = lp[PC_HCHI1] + t_hchi;
    bp1 += bp2;
    qfix_rev(P, f1, gz, u_h, b_hu, b_hv, b_hm0, bp1, b_hchi0, bu, bv, bm1, bchi1, 0.0, b_gz, pa);
    bh += u_h;
    // Q0's hm, p, hchi are outputs and feed F1; hu, hv (after the limit) too
    b_hm0 += lp[PC_HM0];
    double b_p0 = bp1 + lp[PC_P0];
    b_hchi0 += lp[PC_HCHI0];
    b_hu += lp[PC_HU];
    b_hv += lp[PC_HV];
    double b_hu0 = b_hu, b_hv0 = b_hv;
    if (lim) {
        // hu = hu0 L / s, s = |(hu0, hv0)| / h
        b_hu0 = b_hu * sratio;
        b_hv0 = b_hv * sratio;
        const double b_sr = b_hu * q7[1] + b_hv * q7[2];
        pa.speed_limit += b_sr / sp;
        const double b_s = -b_sr * sratio / sp;
        const double b_n = b_s / h;
        bh -= b_s * sp / h;
        if (n0 > 0.0) {
            b_hu0 += b_n * q7[1] / n0;
            b_hv0 += b_n * q7[2] / n0;
        }
    }
    // Q0: the loaded cell's qfix (its u, v, m, chi, rho are not read)
    double b_hm = b_hm0, b_hchi = b_hchi0;
    qfix_rev(P, q0, gz, bh, b_hu0, b_hv0, b_hm, b_p0, b_hchi, 0.0, 0.0, 0.0, 0.0, 0.0, b_gz,
             pa);
    lx[0] = bh;
    lx[1] = b_hu0;
    lx[2] = b_hv0;
    lx[3] = b_hm;
    lx[4] = b_p0;
    lx[5] = b_hchi;
    pa.grav += b_gz;
}


// Whether lambda_n of cell (i, j) comes from the boundary map's pass (rev_bc):
// with the map on load, the ghost cells and the interior cells within 2 of
// the border; the prepare pass writes every other cell's lambda_n itself.
DC_HD bool bc_gathers(const AdjArgs& A, int i, int j)
{
    if (A.bc[0] == 0) return false;
    return !interior(A, i, j) || i <= 2 || i >= A.mx - 1 || j <= 2 || j >= A.my - 1;
}

// B6: gather the adjoint of the prepared fields, then b4step2's pointwise part.
DC_HD void rev_prepare(const AdjArgs& A, const TP& Pz, int i, int j)
{
    const size_t n = cix(A, i, j);
    double lp[NPC];
    for (int k = 0; k < NPC; k++) lp[k] = 0.0;
    // Q0 feeds the update (interior) or is the output (ghost cells)
    {
        double l7[7];
        if (interior(A, i, j)) {
            for (int m = 0; m < 7; m++) l7[m] = wv(A, W_LQN + m, n);
        } else {
            for (int m = 0; m < 7; m++) l7[m] = A.lam_out[qidx(m, i, j, A.mx)];
        }
        lp[PC_H] += l7[0]; lp[PC_HU] += l7[1]; lp[PC_HV] += l7[2]; lp[PC_HM0] += l7[3];
        lp[PC_P0] += l7[4]; lp[PC_HCHI0] += l7[5]; lp[PC_BDIF] += l7[6];
    }
    // transverse splits in this cell
    lp[PC_H] += wv(A, W_LS + 0, n); lp[PC_HU] += wv(A, W_LS + 1, n);
    lp[PC_HV] += wv(A, W_LS + 2, n); lp[PC_HM2] += wv(A, W_LS + 3, n);
    lp[PC_HCHI2] += wv(A, W_LS + 4, n); lp[PC_BDIF] += wv(A, W_LS + 5, n);
    // rpn2 sides: right side of x-edge i and y-edge j (F1, summed by the
    // waves' reverse pass; 0 where the edge does not exist), left side of
    // x-edge i+1 and y-edge j+1 (F2 unless the cell is the slice's first)
    {
        lp[PC_H] += wv(A, W_LR + LR_H, n); lp[PC_HU] += wv(A, W_LR + LR_HU, n);
        lp[PC_HV] += wv(A, W_LR + LR_HV, n); lp[PC_HM1] += wv(A, W_LR + LR_HM1, n);
        lp[PC_P1] += wv(A, W_LR + LR_P1, n); lp[PC_HCHI1] += wv(A, W_LR + LR_HCHI1, n);
        lp[PC_U] += wv(A, W_LR + LR_U, n); lp[PC_V] += wv(A, W_LR + LR_V, n);
        lp[PC_M1] += wv(A, W_LR + LR_M1, n); lp[PC_CHI1] += wv(A, W_LR + LR_CHI1, n);
        lp[PC_BDIF] += wv(A, W_LR + LR_BDIF, n); lp[PC_RHO1] += wv(A, W_LR + LR_RHO1, n);
        lp[PC_TAU1] += wv(A, W_LR + LR_TAU1, n);
        double ls[NSD];
        if (xedge(A, i + 1, j)) {
            for (int k = 0; k < NSD; k++) ls[k] = wv(A, W_LXL + k, cix(A, i + 1, j));
            side_to_pc(ls, i >= 0, 1, lp);
        }
        if (yedge(A, i, j + 1)) {
            for (int k = 0; k < NSD; k++) ls[k] = wv(A, W_LYL + k, cix(A, i, j + 1));
            side_to_pc(ls, j >= 0, 2, lp);
        }
    }
    // calc_taudir of this cell (slot 0) and of its neighbours: the cell is the
    // left neighbour (slot 1) of (i+1, j), the right (2) of (i-1, j), the
    // bottom (3) of (i, j+1), the top (4) of (i, j-1)
    {
        const int ci[5] = {i, i + 1, i - 1, i, i}, cj[5] = {j, j, j, j + 1, j - 1};
        for (int s = 0; s < 5; s++) {
            if (!in_grid(A, ci[s], cj[s])) continue;
            const size_t c = cix(A, ci[s], cj[s]);
            lp[PC_H] += wv(A, W_LTD + lt_h(s), c);
            lp[PC_BDIF] += wv(A, W_LTD + lt_bdif(s), c);
        }
        // the other inputs of calc_taudir have no derivative (taudir_vjp)
        lp[PC_HU] += wv(A, W_LTD + LT_HU, n);
        lp[PC_HV] += wv(A, W_LTD + LT_HV, n);
        lp[PC_RHO1] += wv(A, W_LTD + LT_RHO, n);
        lp[PC_TAU1] += wv(A, W_LTD + LT_TAU, n);
    }
    double q7[7], fu, fv, lx[7], g[MAXP] = {0, 0, 0, 0};
    load_q(A, i, j, q7, fu, fv);
    if (q7[0] <= A.P.dry_tolerance) {
        // qfix zeroes a dry cell; only bdif = max(bdif, 0) remains
        for (int m = 0; m < 6; m++) lx[m] = 0.0;
        lx[6] = q7[6] > 0.0 ? lp[PC_BDIF] : 0.0;
    } else {
        // bdif enters the prepared fields only as max(bdif, 0)
        ParAdj pa;
        pa_zero(pa);
        prepare_vjp(A.P, q7, lp, lx, pa);
        lx[6] = q7[6] > 0.0 ? lp[PC_BDIF] : 0.0;
#pragma unroll
        for (int p = 0; p < MAXP; p++)
            if (p < A.np) g[p] = pa_dot(pa, A.dP[p]);
        gp_add(A, n, g);
    }
    if (bc_gathers(A, i, j)) {
        for (int m = 0; m < 7; m++) ws(A, W_LQ7 + m, n, lx[m]);
    } else {
        // lambda_n of a cell the boundary map does not touch
        double* o = A.lam_in + qidx(0, i, j, A.mx);
        for (int m = 0; m < 7; m++) o[m] = lx[m];
    }
}

// B7: lambda_n.  With the boundary map on load a ghost cell's q is its source
// cell's (normal momentum negated at a wall), so its adjoint goes to the
// source; the stored ghosts of q_n are not read and get 0.
// The ghost cells whose source under the boundary map is interior cell
// (i, j) -- columns -1, 0 / mx+1, mx+2 and rows -1, 0 / my+1, my+2 around
// the cell's own -- in a fixed order: f(ci, cj, fu, fv).
template <class F>
DC_HD void for_ghost_images(const AdjArgs& A, int i, int j, F f)
{
    int gi[5], gj[5], ni = 0, nj = 0;
    gi[ni++] = i;
    if (i <= 2) { gi[ni++] = -1; gi[ni++] = 0; }
    if (i >= A.mx - 1) { gi[ni++] = A.mx + 1; gi[ni++] = A.mx + 2; }
    gj[nj++] = j;
    if (j <= 2) { gj[nj++] = -1; gj[nj++] = 0; }
    if (j >= A.my - 1) { gj[nj++] = A.my + 1; gj[nj++] = A.my + 2; }
    for (int b = 0; b < nj; b++)
        for (int a = 0; a < ni; a++) {
            const int ci = gi[a], cj = gj[b];
            if (interior(A, ci, cj)) continue;
            int si, sj;
            double fu, fv;
            bc_source(A.bc, A.mx, A.my, ci, cj, si, sj, fu, fv);
            if (si != i || sj != j) continue;
            f(ci, cj, fu, fv);
        }
}

// Whether the tile passes ran on cell (i, j) this call.
DC_HD bool cell_live(const AdjArgs& A, int i, int j)
{
    return !A.tflags || A.tflags[((j + 1) / TILE_H) * A.tntx + (i + 1) / TILE_W] != 0;
}

// The bdif adjoint a dry ghost cell sends to its source (i, j) (bdif raw
// value b there): the prepare pass's value far from wet cells, where the
// ghost's Q0 is the dry state and only its output's bdif adjoint reaches
// max(bdif, 0).
DC_HD double dry_ghost_bdif(const AdjArgs& A, int ci, int cj, double b)
{
    return b > 0.0 ? A.lam_out[qidx(6, ci, cj, A.mx)] : 0.0;
}

DC_HD void rev_bc(const AdjArgs& A, int i, int j)
{
    if (!bc_gathers(A, i, j)) return;   // written by the prepare pass
    const size_t n = cix(A, i, j);
    double l[7];
    for (int m = 0; m < 7; m++) l[m] = wv(A, W_LQ7 + m, n);
    if (!interior(A, i, j)) {
        for (int m = 0; m < 7; m++) l[m] = 0.0;
    } else {
        for_ghost_images(A, i, j, [&](int ci, int cj, double fu, double fv) {
            if (cell_live(A, ci, cj)) {
                const size_t c = cix(A, ci, cj);
                for (int m = 0; m < 7; m++) {
                    const double s = m == 1 ? fu : (m == 2 ? fv : 1.0);
                    l[m] += s * wv(A, W_LQ7 + m, c);
                }
            } else {
                // a ghost in a tile the passes skipped (far from wet cells)
                l[6] += dry_ghost_bdif(A, ci, cj, A.q_in[qidx(6, i, j, A.mx)]);
            }
        });
    }
    double* o = A.lam_in + qidx(0, i, j, A.mx);
    for (int m = 0; m < 7; m++) o[m] = l[m];
}

}  // namespace dclaw_adj
