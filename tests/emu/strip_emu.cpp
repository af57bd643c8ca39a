// strip_emu.cpp -- runs the product's strip kernel body (dclaw_strip.h) on
// the host: 64 coroutines play the 64 lanes of one wavefront and meet at
// every lane shuffle, so control flow and data movement are exactly the
// GPU's.  With glibc libm this reproduces the CPU oracle's
// arithmetic, so tests compare the two bit for bit.  Test tool only.
#include <ucontext.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dclaw_adjoint.h"
#include "dclaw_amr.h"
#include "dclaw_strip.h"
#include "dclaw_tangent.h"

using namespace dclaw;

namespace {

// Cooperative scheduler: the 64 lanes are coroutines run round-robin in one
// thread.  Every lane yields once per shuffle, after publishing its value in
// one of two buffers (alternating per shuffle), so when a lane resumes all
// lanes have published and nobody has overwritten the buffer it reads.
struct Sched {
    ucontext_t main;
    ucontext_t ctx[STRIP_LANES];
    double buf[2][STRIP_LANES];
    double cfl[STRIP_LANES];
    bool done[STRIP_LANES];
    const dclaw_params* P;
    const StripArgs* A;
    int mode;
    int seg_rows;                 // MODE_CLASSIFY
    unsigned mask[STRIP_LANES];   // MODE_CLASSIFY result per lane
    const dclaw_tl::Params* Ptl;  // MODE_TANGENT
    const dclaw_tl::StripArgs* Atl;
    std::vector<char> stacks;
};

constexpr int MODE_CLASSIFY = 100;   // strip_classify instead of the strip body
constexpr int MODE_TANGENT = 101;    // the tangent-linear strip body (fused step)

struct CoLane {
    int lid;
    int n;
    Sched* s;
    int id() const { return lid; }
    double shfl(double v, int src)
    {
        const int b = n & 1;
        n++;
        s->buf[b][lid] = v;
        swapcontext(&s->ctx[lid], &s->main);
        return s->buf[b][src];
    }
    double prev(double v) { return shfl(v, lid > 0 ? lid - 1 : lid); }
    double next(double v) { return shfl(v, lid < STRIP_LANES - 1 ? lid + 1 : lid); }
    dclaw_tl::Dual prev(const dclaw_tl::Dual& v)
    {
        const double a = prev(v.v);
        return dclaw_tl::Dual(a, prev(v.d));
    }
    dclaw_tl::Dual next(const dclaw_tl::Dual& v)
    {
        const double a = next(v.v);
        return dclaw_tl::Dual(a, next(v.d));
    }
    bool any(bool p)
    {
        const int b = n & 1;
        n++;
        s->buf[b][lid] = p ? 1.0 : 0.0;
        swapcontext(&s->ctx[lid], &s->main);
        for (int l = 0; l < STRIP_LANES; l++)
            if (s->buf[b][l] != 0.0) return true;
        return false;
    }
};

Sched* g_sched = nullptr;

void lane_entry(int lid)
{
    Sched* s = g_sched;
    CoLane L{lid, 0, s};
    if (s->mode == MODE_TANGENT) {
        s->cfl[lid] = dclaw_tl::strip_step<MODE_STEP>(L, *s->Ptl, *s->Atl).v;
        s->done[lid] = true;
        return;
    }
    if (s->mode == MODE_CLASSIFY) {
        s->mask[lid] = strip_classify(L, s->P->dry_tolerance, *s->A, s->seg_rows);
        s->cfl[lid] = 0.0;
        s->done[lid] = true;
        return;
    }
    s->cfl[lid] = s->mode == MODE_FLUX ? strip_step_dispatch<MODE_FLUX>(L, *s->P, *s->A)
                  : s->mode == MODE_STEP_NOSRC2
                      ? strip_step_dispatch<MODE_STEP_NOSRC2>(L, *s->P, *s->A)
                      : strip_step_dispatch<MODE_STEP>(L, *s->P, *s->A);
    s->done[lid] = true;
}

// Runs one wavefront; returns the classify mask (MODE_CLASSIFY; -1 if the
// lanes disagree), else 0.
long run_strip(const dclaw_params& P, const StripArgs& A, double* cfl_out, int mode = MODE_STEP,
               int seg_rows = 0, const dclaw_tl::Params* Ptl = nullptr,
               const dclaw_tl::StripArgs* Atl = nullptr)
{
    constexpr size_t STACK = 1 << 20;
    Sched* s = new Sched();
    s->P = &P;
    s->A = &A;
    s->Ptl = Ptl;
    s->Atl = Atl;
    s->mode = mode;
    s->seg_rows = seg_rows;
    s->stacks.resize(STACK * STRIP_LANES);
    g_sched = s;
    for (int l = 0; l < STRIP_LANES; l++) {
        s->done[l] = false;
        getcontext(&s->ctx[l]);
        s->ctx[l].uc_stack.ss_sp = s->stacks.data() + STACK * l;
        s->ctx[l].uc_stack.ss_size = STACK;
        s->ctx[l].uc_link = &s->main;
        makecontext(&s->ctx[l], (void (*)())lane_entry, 1, l);
    }
    bool all_done = false;
    while (!all_done) {
        all_done = true;
        for (int l = 0; l < STRIP_LANES; l++) {
            if (s->done[l]) continue;
            swapcontext(&s->main, &s->ctx[l]);
            if (!s->done[l]) all_done = false;
        }
    }
    for (int l = 0; l < STRIP_LANES; l++) *cfl_out = dmax(*cfl_out, s->cfl[l]);
    long mask = 0;
    if (mode == MODE_CLASSIFY) {
        mask = (long)s->mask[0];
        for (int l = 1; l < STRIP_LANES; l++)
            if (s->mask[l] != s->mask[0]) mask = -1;
    }
    delete s;
    return mask;
}

}  // namespace

extern "C" int emu_stepgrid(const dclaw_params* P, const double* q_in, double* q_out,
                            double* aux, int maux, int write_aux, int mx, int my, double dx,
                            double dy, double dt, int seg_rows, const int* bc, double* cflgrid)
{
    *cflgrid = 0.0;
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            StripArgs A;
            A.q_in = q_in;
            A.q_out = q_out;
            A.aux = aux;
            A.aux_out = write_aux ? aux : nullptr;
            A.maux = maux;
            A.mx = mx;
            A.my = my;
            A.dx = dx;
            A.dy = dy;
            A.dt = dt;
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g * seg_rows;
            A.jB = std::min(my, A.jA + seg_rows - 1);
            for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
            A.fm = A.fp = A.gm = A.gp = nullptr;
            A.tile_flag = nullptr;
            run_strip(*P, A, cflgrid);
        }
    return 0;
}

// The strip pass of the split step alone: q_out receives the state before
// src2 (the cells src2 would zero already zeroed) and flags[segment *
// strips + strip] each tile's flag (TF_* bits).
extern "C" int emu_strip_flags(const dclaw_params* P, const double* q_in, double* q_out,
                               const double* aux, int maux, int mx, int my, double dx, double dy,
                               double dt, int seg_rows, const int* bc, int* flags)
{
    double cflgrid = 0.0;
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    for (int t = 0; t < nstrips * nseg; t++) flags[t] = -1;
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            StripArgs A;
            A.q_in = q_in;
            A.q_out = q_out;
            A.aux = aux;
            A.aux_out = nullptr;
            A.maux = maux;
            A.mx = mx;
            A.my = my;
            A.dx = dx;
            A.dy = dy;
            A.dt = dt;
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g * seg_rows;
            A.jB = std::min(my, A.jA + seg_rows - 1);
            for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
            A.fm = A.fp = A.gm = A.gp = nullptr;
            A.tile_flag = &flags[(size_t)g * nstrips + s];
            run_strip(*P, A, &cflgrid, MODE_STEP_NOSRC2);
        }
    for (int t = 0; t < nstrips * nseg; t++)
        if (flags[t] < 0) return -1;       // a tile did not report
    return 0;
}

// The step with src2 split out: the strip body without src2 flags its tiles,
// then src2 runs over the interior cells of the flagged tiles.
extern "C" int emu_stepgrid_split(const dclaw_params* P, const double* q_in, double* q_out,
                                  const double* aux, int maux, int mx, int my, double dx,
                                  double dy, double dt, int seg_rows, const int* bc,
                                  double* cflgrid, int* nflagged)
{
    *cflgrid = 0.0;
    *nflagged = 0;
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    std::vector<int> flags((size_t)nstrips * nseg, -1);
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            StripArgs A;
            A.q_in = q_in;
            A.q_out = q_out;
            A.aux = aux;
            A.aux_out = nullptr;
            A.maux = maux;
            A.mx = mx;
            A.my = my;
            A.dx = dx;
            A.dy = dy;
            A.dt = dt;
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g * seg_rows;
            A.jB = std::min(my, A.jA + seg_rows - 1);
            for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
            A.fm = A.fp = A.gm = A.gp = nullptr;
            A.tile_flag = &flags[(size_t)g * nstrips + s];
            run_strip(*P, A, cflgrid, MODE_STEP_NOSRC2);
        }
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            const int f = flags[(size_t)g * nstrips + s];
            if (f < 0) return -1;          // a tile did not report
            if (!f) continue;
            (*nflagged)++;
            const int i0 = 1 + s * STRIP_W, jA = 1 + g * seg_rows;
            const int jB = std::min(my, jA + seg_rows - 1);
            for (int k = 0; k < STRIP_W * (jB - jA + 1); k++)
                src2_tile_cell(*P, dt, q_out, aux, maux, mx, i0, jA, jB, k);
        }
    return 0;
}

// The split step with its dry pass, as the GPU runs it: strip_classify over
// groups of `group` segments writes the dry state and lists the other
// segments; the strip body (known_wet, no src2) runs on the listed segments
// in reverse list order (the GPU's order is arbitrary), then src2 on the
// flagged ones.  nwet: listed segments.
extern "C" int emu_stepgrid_classified(const dclaw_params* P, const double* q_in, double* q_out,
                                       const double* aux, int maux, int mx, int my, double dx,
                                       double dy, double dt, int seg_rows, int group,
                                       const int* bc, double* cflgrid, int* nwet)
{
    *cflgrid = 0.0;
    *nwet = 0;
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    std::vector<int> flags((size_t)nstrips * nseg, -1);
    std::vector<int> list;
    StripArgs A;
    A.q_in = q_in;
    A.q_out = q_out;
    A.aux = aux;
    A.aux_out = nullptr;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = dt;
    for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = nullptr;
    for (int s = 0; s < nstrips; s++)
        for (int g0 = 0; g0 < nseg; g0 += group) {
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g0 * seg_rows;
            A.jB = std::min(my, A.jA + group * seg_rows - 1);
            A.known_wet = false;
            const long m = run_strip(*P, A, cflgrid, MODE_CLASSIFY, seg_rows);
            if (m < 0) return -2;          // lanes disagree on the mask
            const int ng = std::min(group, nseg - g0);
            if (m >> ng) return -3;        // a bit beyond the group
            for (int k = 0; k < ng; k++) {
                const int t = (g0 + k) * nstrips + s;
                if ((m >> k) & 1) list.push_back(t);
                else flags[t] = 0;
            }
        }
    *nwet = (int)list.size();
    for (auto it = list.rbegin(); it != list.rend(); ++it) {
        const int t = *it;
        A.i0 = 1 + (t % nstrips) * STRIP_W;
        A.jA = 1 + (t / nstrips) * seg_rows;
        A.jB = std::min(my, A.jA + seg_rows - 1);
        A.known_wet = true;
        A.tile_flag = &flags[t];
        run_strip(*P, A, cflgrid, MODE_STEP_NOSRC2);
    }
    for (int t = 0; t < nstrips * nseg; t++) {
        if (flags[t] < 0) return -1;       // a tile did not report
        if (!flags[t]) continue;
        const int i0 = 1 + (t % nstrips) * STRIP_W, jA = 1 + (t / nstrips) * seg_rows;
        const int jB = std::min(my, jA + seg_rows - 1);
        for (int k = 0; k < STRIP_W * (jB - jA + 1); k++)
            src2_tile_cell(*P, dt, q_out, aux, maux, mx, i0, jA, jB, k);
    }
    return 0;
}

// step2 (flux mode of the strip body) on the b4step2 output: fm, fp, gm, gp
// are zero-filled here, then the strips store their entries.
extern "C" int emu_step2(const dclaw_params* P, const double* q, const double* aux, int maux,
                         int mx, int my, double dx, double dy, double dt, int seg_rows,
                         double* fm, double* fp, double* gm, double* gp, double* cflgrid)
{
    *cflgrid = 0.0;
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    for (size_t k = 0; k < nq; k++) fm[k] = fp[k] = gm[k] = gp[k] = 0.0;
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            StripArgs A;
            A.q_in = q;
            A.q_out = nullptr;
            A.aux = aux;
            A.aux_out = nullptr;
            A.maux = maux;
            A.mx = mx;
            A.my = my;
            A.dx = dx;
            A.dy = dy;
            A.dt = dt;
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g * seg_rows;
            A.jB = std::min(my, A.jA + seg_rows - 1);
            for (int k = 0; k < 4; k++) A.bc[k] = 0;
            A.fm = fm;
            A.fp = fp;
            A.gm = gm;
            A.gp = gp;
            A.tile_flag = nullptr;
            run_strip(*P, A, cflgrid, MODE_FLUX);
        }
    return 0;
}

// rpn2 over one slice with the product's per-edge routine, in the
// reference's array layout: q1d(7, 1-mbc:mx+mbc) (not modified: the qfix
// chain is applied to copies), aux2(maux, ...), fwave(7, 5, ...), s(5, ...).
extern "C" int emu_rpn2_slice(const dclaw_params* P, int ixy, int mx, int maux,
                              const double* q1d, const double* aux2, double* fwave, double* s)
{
    const int n = mx + 4;
    const int itx = (ixy == 1) ? 4 : 5;   // 0-based aux index of taudir_x / taudir_y
    std::vector<RowCell> cells(n);
    for (int k = 0; k < n; k++) {
        double q7[7];
        for (int m = 0; m < 7; m++) q7[m] = q1d[7 * k + m];
        // the slice is already b4step2-processed: prepare_cell's qfix is then
        // the identity on wet cells only up to rounding, so bypass it by
        // building F1 / F2 exactly as rpn2 applies them.
        RowCell r;
        r.h = q7[0]; r.hu = q7[1]; r.hv = q7[2]; r.bdif = q7[6]; r.b = aux2[maux * k];
        r.hm0 = q7[3]; r.p0 = q7[4]; r.hchi0 = q7[5];
        double h = r.h, hu = r.hu, hv = r.hv, hm = r.hm0, p = r.p0, hchi = r.hchi0;
        double u, v, m_, chi, rho;
        qfix(*P, h, hu, hv, hm, p, hchi, u, v, m_, chi, rho, P->grav);
        r.h = h; r.hu = hu; r.hv = hv;
        r.hm1 = hm; r.p1 = p; r.hchi1 = hchi; r.m1 = m_; r.chi1 = chi; r.u = u; r.v = v;
        qfix(*P, h, hu, hv, hm, p, hchi, u, v, m_, chi, rho, P->grav);
        r.hm2 = hm; r.p2 = p; r.hchi2 = hchi; r.m2 = m_; r.chi2 = chi;
        cells[k] = r;
    }
    for (int k = 0; k < 35 * n; k++) fwave[k] = 0.0;
    for (int k = 0; k < 5 * n; k++) s[k] = 0.0;
    for (int e = 1; e < n; e++) {          // edge between array cells e-1 and e
        const RowCell& cl = cells[e - 1];
        const RowCell& cr = cells[e];
        const bool second = e - 1 >= 1;
        RhoTau vl = {0.0, 0.0}, vr = {0.0, 0.0};
        if (cl.h > P->dry_tolerance)
            vl = second ? rho_tau(*P, cl.h, cl.u, cl.v, cl.m2, cl.p2, cl.chi2)
                        : rho_tau(*P, cl.h, cl.u, cl.v, cl.m1, cl.p1, cl.chi1);
        if (cr.h > P->dry_tolerance) vr = rho_tau(*P, cr.h, cr.u, cr.v, cr.m1, cr.p1, cr.chi1);
        Side sL = side_of(cl, second, aux2[maux * (e - 1) + itx], aux2[maux * (e - 1) + 3], vl);
        Side sR = side_of(cr, false, aux2[maux * e + itx], aux2[maux * e + 3], vr);
        if (ixy == 2) {
            sL.u = cells[e - 1].v; sL.v = cells[e - 1].u;
            sR.u = cells[e].v; sR.v = cells[e].u;
        }
        const EdgeWaves w = rpn2_edge(*P, ixy, sL, sR);
        if (getenv("EMU_DBG") && e == atoi(getenv("EMU_DBG")))
            fprintf(stderr, "L h%.17g hu%.17g hv%.17g hm%.17g p%.17g hc%.17g u%.17g v%.17g m%.17g chi%.17g b%.17g t%.17g fs%.17g\nR h%.17g hu%.17g hv%.17g hm%.17g p%.17g hc%.17g u%.17g v%.17g m%.17g chi%.17g b%.17g t%.17g fs%.17g\n",
                sL.h, sL.hu, sL.hv, sL.hm, sL.p, sL.hchi, sL.u, sL.v, sL.m, sL.chi, sL.b, sL.taudir, sL.fs,
                sR.h, sR.hu, sR.hv, sR.hm, sR.p, sR.hchi, sR.u, sR.v, sR.m, sR.chi, sR.b, sR.taudir, sR.fs);
        double* sv = s + 5 * e;
        sv[0] = w.s[0]; sv[1] = w.s[1]; sv[2] = w.s[1]; sv[3] = w.s[1]; sv[4] = w.s[2];
        double* f = fwave + 35 * e;
        const int nrm = (ixy == 1) ? 1 : 2;
        for (int m = 0; m < 6; m++) { f[0 * 7 + m] = w.w1[m]; f[4 * 7 + m] = w.w5[m]; }
        f[1 * 7 + 0] = w.w2h;
        f[1 * 7 + nrm] = w.w2n;
    }
    return 0;
}

// One rpn2 edge from explicit side states (in[0..12] left, in[13..25] right:
// h hu hv hm p hchi u v m chi b taudir fs), outputs s[3] and w1, w5 (q-index).
extern "C" int emu_rpn2_edge(const dclaw_params* P, int ixy, const double* in, double* out)
{
    Side S[2];
    for (int k = 0; k < 2; k++) {
        const double* x = in + 13 * k;
        S[k].h = x[0]; S[k].hu = x[1]; S[k].hv = x[2]; S[k].hm = x[3]; S[k].p = x[4];
        S[k].hchi = x[5]; S[k].u = x[6]; S[k].v = x[7]; S[k].m = x[8]; S[k].chi = x[9];
        S[k].b = x[10]; S[k].taudir = x[11]; S[k].fs = x[12];
        S[k].rho = S[k].tau = 0.0;
        if (S[k].h > P->dry_tolerance) {
            const RhoTau t = rho_tau(*P, S[k].h, S[k].u, S[k].v, S[k].m, S[k].p, S[k].chi);
            S[k].rho = t.rho;
            S[k].tau = t.tau;
        }
    }
    const EdgeWaves w = rpn2_edge(*P, ixy, S[0], S[1]);
    for (int k = 0; k < 3; k++) out[k] = w.s[k];
    for (int k = 0; k < 6; k++) { out[3 + k] = w.w1[k]; out[9 + k] = w.w5[k]; }
    out[15] = w.w2h; out[16] = w.w2n;
    return 0;
}

// Elementary functions of dcmath.h over an array (fn: 0 exp, 1 atan, 2 tan,
// 3 tanh, 4 log, 5 pow(x, y2[k])) -- for the accuracy tests.
extern "C" int emu_dcm(int fn, const double* x, const double* y2, double* out, int n)
{
    for (int k = 0; k < n; k++) {
        switch (fn) {
        case 0: out[k] = dcm_exp(x[k]); break;
        case 1: out[k] = dcm_atan(x[k]); break;
        case 2: out[k] = dcm_tan(x[k]); break;
        case 3: out[k] = dcm_tanh(x[k]); break;
        case 4: out[k] = dcm_log(x[k]); break;
        default: out[k] = dcm_pow(x[k], y2[k]); break;
        }
    }
    return 0;
}

// The AMR transfer bodies (dclaw_amr.h) run cell by cell on the host.
extern "C" void emu_amr_average(double* qc, const double* auxc, int mauxc, int mitc,
                                const double* qf, const double* auxf, int mauxf, int mitf, int ci0,
                                int cj0, int ni, int nj, int fi0, int fj0, int rx, int ry,
                                double dry_tolerance)
{
    for (int b = 0; b < nj; b++)
        for (int a = 0; a < ni; a++)
            amr_average_cell(qc, auxc, mauxc, mitc, qf, auxf, mauxf, mitf, ci0 + a, cj0 + b,
                             fi0 + a * rx, fj0 + b * ry, rx, ry, dry_tolerance);
}

extern "C" void emu_amr_interp(const double* valc, const double* auxc, int mauxc, int mic,
                               int mjc, double* val, const double* auxf, int mauxf, int mitot,
                               const unsigned char* needs_set, int rx, int ry,
                               double dry_tolerance, double sea_level, double rho_f, double grav)
{
    for (int j = 1; j <= mjc - 2; j++)
        for (int i = 1; i <= mic - 2; i++)
            amr_interp_cell(valc, auxc, mauxc, mic, val, auxf, mauxf, mitot, needs_set, rx, ry, i,
                            j, dry_tolerance, sea_level, rho_f, grav);
}

// filpatch's interpolation (amr_filpatch_cell) over every interior coarse
// cell of the rectangle, in reverse order (cells are independent).
extern "C" void emu_amr_filpatch(const double* valc, const double* auxc, int mauxc, int mxc,
                                 int myc, int iplo, int jplo, double* q, const double* aux,
                                 int maux, int mx, int r0, int c0, int ilo, int ihi, int jlo,
                                 int jhi, const unsigned char* flag, int rx, int ry,
                                 double xlower, double ylower, double dxf, double dyf,
                                 double dxc, double dyc, double sea_level, double drytol,
                                 double rho_f, double grav)
{
    AmrFill G;
    G.ilo = ilo; G.ihi = ihi; G.jlo = jlo; G.jhi = jhi; G.r0 = r0; G.c0 = c0; G.mx = mx;
    G.maux = maux; G.iplo = iplo; G.jplo = jplo; G.mxc = mxc; G.myc = myc; G.mauxc = mauxc;
    G.rx = rx; G.ry = ry; G.xlower = xlower; G.ylower = ylower; G.dxf = dxf; G.dyf = dyf;
    G.dxc = dxc; G.dyc = dyc; G.sea_level = sea_level; G.drytol = drytol; G.rho_f = rho_f;
    G.grav = grav;
    for (int jc = myc - 1; jc >= 2; jc--)
        for (int ic = mxc - 1; ic >= 2; ic--) amr_filpatch_cell(valc, auxc, q, aux, flag, G, ic, jc);
}

// The tangent-linear fused step: q_out, cflgrid as emu_stepgrid (aux not
// written); dq_out = J_q dq_in + J_P dP at dt.
extern "C" int emu_stepgrid_tl(const dclaw_params* P, const dclaw_params* dP, const double* q_in,
                               const double* dq_in, double* q_out, double* dq_out,
                               const double* aux, int maux, int mx, int my, double dx, double dy,
                               double dt, int seg_rows, const int* bc, double* cflgrid)
{
    *cflgrid = 0.0;
    const dclaw_tl::Params Ptl = dclaw_tl::make_params(*P, *dP);
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    StripArgs dummy{};
    for (int s = 0; s < nstrips; s++)
        for (int g = 0; g < nseg; g++) {
            dclaw_tl::StripArgs A;
            A.q_in.v = q_in;
            A.q_in.d = dq_in;
            A.q_out.v = q_out;
            A.q_out.d = dq_out;
            A.aux = aux;
            A.aux_out = nullptr;
            A.maux = maux;
            A.mx = mx;
            A.my = my;
            A.dx = dx;
            A.dy = dy;
            A.dt = dt;
            A.i0 = 1 + s * STRIP_W;
            A.jA = 1 + g * seg_rows;
            A.jB = std::min(my, A.jA + seg_rows - 1);
            for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
            A.fm = A.fp = A.gm = A.gp = nullptr;
            A.tile_flag = nullptr;
            run_strip(*P, dummy, cflgrid, MODE_TANGENT, 0, &Ptl, &A);
        }
    return 0;
}
