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

// The adjoint step (dclaw_adjoint.h) with its stages run over the grid in
// order on the host: q_out is the record pass's step, lam_in = J_q^T lam_out
// and gp[k] = lam_out . J_P dP[k] summed over the cells in index order.
// emu_set_adjoint_fused(1) runs the update and src2's reverse as the GPU
// does (one fused pass); 0 (default) as two passes through the QN planes.
static int g_emu_fused_update = 0;
extern "C" void emu_set_adjoint_fused(int on) { g_emu_fused_update = on; }

extern "C" int emu_adjoint_step(const dclaw_params* P, const dclaw_params* dP, int np,
                                const double* q_in, const double* aux, int maux, int mx, int my,
                                double dx, double dy, double dt, const int* bc,
                                const double* lam_out, double* lam_in, double* q_out, double* gp)
{
    namespace D = dclaw_adj;
    if (np < 0 || np > D::MAXP) return 1;
    D::AdjArgs A;
    A.q_in = q_in;
    A.aux = aux;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = dt;
    for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
    A.P = *P;
    A.dP = dP;
    A.np = np;
    A.lam_out = lam_out;
    A.lam_in = lam_in;
    A.q_out = q_out;
    A.N = D::plane_size(mx, my);
    A.ld = D::plane_ld(mx);
    A.Pd = nullptr;
    A.dt_dev = nullptr;
    std::vector<double> W((size_t)D::W_FIELDS * A.N, 0.0);
    A.W = W.data();
    const D::TP Pz = D::zero_params(*P);
    auto all = [&](auto fn) {
        for (int j = -1; j <= my + 2; j++)
            for (int i = -1; i <= mx + 2; i++) fn(A, Pz, i, j);
    };
    all(D::rec_prepare);
    all(D::rec_taudir);
    all(D::rec_waves);
    all(D::rec_flux);
    all(D::rec_split);
    if (g_emu_fused_update) {
        // the GPU passes: update and src2's reverse in one pass, then the
        // seeded src2 reverse for configurations the written one does not cover
        all(D::upd_src2);
        all(D::rev_src2_t<false>);
    } else {
        all(D::rec_update);
        all(D::rev_src2);
    }
    all(D::rev_split);
    all(D::rev_flux);
    all(D::rev_waves);
    all(D::rev_taudir);
    all(D::rev_prepare);
    for (int j = -1; j <= my + 2; j++)
        for (int i = -1; i <= mx + 2; i++) D::rev_bc(A, i, j);
    for (int k = 0; k < np; k++) {
        double s = 0.0;
        for (size_t n = 0; n < A.N; n++) s += W[(size_t)(D::W_GP + k) * A.N + n];
        gp[k] = s;
    }
    return 0;
}

// The hand-written flux2 edge-body adjoint (flux_vjp) against the seeded
// tangent-linear vector-Jacobian product of the same function: n random
// cases (waves, limiter dot products, adjoints) for the given scheme switches;
// returns the largest |difference| relative to the largest |entry|.
extern "C" double emu_flux_vjp_check(const dclaw_params* P, int n, unsigned seed)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    D::AdjArgs A{};
    A.P = *P;
    A.np = 0;
    const D::TP Pz = D::zero_params(*P);
    double worst = 0.0;
    for (int c = 0; c < n; c++) {
        double x[D::NEI], ly[D::NEF], la[D::NEI], lb[D::NEI];
        for (int k = 0; k < D::NEI; k++) x[k] = rnd();
        // some speeds exactly 0 and both signs; dot products of either sign
        if (c % 7 == 0) x[c % 3] = 0.0;
        for (int k = 0; k < D::NEF; k++) ly[k] = rnd();
        D::FluxF f;
        f.ixy = 1 + (c & 1);
        f.gate = (c % 5) != 0;
        f.dtdx = 0.3 + 0.2 * rnd();
        D::vjp<D::NEI, D::NEF>(A, f, Pz, x, ly, la, nullptr);
        D::flux_vjp(*P, f, x, ly, lb);
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < D::NEI; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        if (scale > 0.0) worst = std::max(worst, diff / scale);
        else worst = std::max(worst, diff);
    }
    return worst;
}

// The hand-written rpt2 split adjoint (split_vjp) against the seeded
// tangent-linear vector-Jacobian product (inputs and the grav, rho_s, rho_f
// directions): n random wet cells with random fluctuations, both sweep
// directions and random neighbour beds; returns the largest difference
// relative to the largest input adjoint (parameters: to the largest
// sensitivity to a relative parameter change).
extern "C" double emu_split_vjp_check(const dclaw_params* P, int n, unsigned seed)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    dclaw_params dP[3] = {};
    dP[0].grav = 1.0;
    dP[1].rho_s = 1.0;
    dP[2].rho_f = 1.0;
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    A.np = 3;
    const D::TP Pz = D::zero_params(*P);
    double worst = 0.0;
    for (int c = 0; c < n; c++) {
        double x[D::NSI], ly[12], la[D::NSI], lb[D::NSI];
        x[0] = 0.05 + 2.0 * (rnd() + 1.0);                    // h
        x[1] = 3.0 * rnd() * x[0];
        x[2] = 3.0 * rnd() * x[0];
        x[3] = (0.5 + 0.2 * rnd()) * x[0];
        x[4] = 0.3 * (rnd() + 1.0) * x[0];
        x[5] = 0.1 * rnd();
        for (int k = 6; k < D::NSI; k++) x[k] = rnd();
        for (int k = 0; k < 12; k++) ly[k] = rnd();
        D::SplitF f;
        f.ixy = 1 + (c & 1);
        f.bedc = rnd();
        f.topo1 = f.bedc + 2.0 * rnd();
        f.topo3 = f.bedc + 2.0 * rnd();
        f.active = true;
        double ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
        D::vjp<D::NSI, 12>(A, f, Pz, x, ly, la, ga);
        D::split_vjp(A, f, x, ly, lb, gb);
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < D::NSI; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        // parameter derivatives compared as sensitivities to a relative change
        // of each parameter (|dL/dp| |p|), against the largest of the three: a
        // derivative that is 0 in exact arithmetic (d/drho_s with src2method
        // -1, where gamma rho = rho_f) is rounding residue in both forms
        const double pm[3] = {P->grav, P->rho_s, P->rho_f};
        double gs = 0.0;
        for (int k = 0; k < 3; k++) gs = std::max(gs, std::fabs(ga[k] * pm[k]));
        for (int k = 0; k < 3; k++) {
            const double d = std::fabs((ga[k] - gb[k]) * pm[k]);
            if (gs > 0.0) worst = std::max(worst, d / gs);
        }
        if (scale > 0.0) worst = std::max(worst, diff / scale);
        else worst = std::max(worst, diff);
    }
    return worst;
}

// The written-out src2 reverse (src2_vjp) against the seeded vector-Jacobian
// product vjp<7, 7>(Src2F) on n random wet cells (velocities of both signs
// and at rest, m and chi inside and outside [0, 1], pb below and above
// rho g h, a few with hm < 0), for the six state inputs and the 13
// parameter fields it covers (one unit direction each).  Returns the
// largest difference relative to the largest input adjoint (parameters:
// as sensitivities |dL/dp| |p| against the largest of them); worst2[0..1]
// get the input and the parameter parts when not null.
extern "C" double emu_src2_vjp_check(const dclaw_params* P, int n, unsigned seed, double dt,
                                     double* worst2)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    constexpr int NP = 13;
    dclaw_params dP[4];
    const double pv[NP] = {P->grav, P->rho_s, P->rho_f, P->mu, P->phi, P->kref, P->mref,
                           P->delta_kr_order, P->alpha_c, P->sigma_0, P->delta, P->m_crit,
                           P->c1};
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    const D::TP Pz = D::zero_params(*P);
    double wi = 0.0, wp = 0.0;
    for (int c = 0; c < n; c++) {
        double x[7], ly[7], la[7], lb[7];
        const double h = 0.01 + 1.5 * (rnd() + 1.0);
        x[0] = h;
        const bool still = c % 9 == 0;
        x[1] = still ? 0.0 : 3.0 * rnd() * h;
        x[2] = still ? 0.0 : 3.0 * rnd() * h;
        double m = 0.5 + 0.25 * rnd();
        if (c % 11 == 3) m = 1.0 + 0.05 * (rnd() + 1.0);
        if (c % 13 == 5) m = -0.05 * (rnd() + 1.0);
        x[3] = m * h;
        const double rho = P->rho_s * m + (1.0 - m) * P->rho_f;
        x[4] = (0.6 + 0.5 * rnd()) * rho * P->grav * h;
        if (c % 7 == 2) x[4] = -0.1 * rho * P->grav * h;
        x[5] = (0.5 + 0.6 * rnd()) * h;
        x[6] = 0.1 * rnd();
        for (int k = 0; k < 7; k++) ly[k] = rnd();
        D::Src2F f{dt, 0.0};
        // inputs
        A.np = 0;
        D::vjp<7, 7>(A, f, Pz, x, ly, la, nullptr);
        D::ParAdj pa;
        D::pa_zero(pa);
        D::src2_vjp(*P, dt, x, ly, lb, pa);
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < 6; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        wi = std::max(wi, scale > 0.0 ? diff / scale : diff);
        // parameters, four directions per seeded call
        double ga[NP], gb[NP];
        for (int p0 = 0; p0 < NP; p0 += 4) {
            const int np = std::min(4, NP - p0);
            for (int k = 0; k < np; k++) {
                std::memset(&dP[k], 0, sizeof(dclaw_params));
                double* fld[NP] = {&dP[k].grav, &dP[k].rho_s, &dP[k].rho_f, &dP[k].mu,
                                   &dP[k].phi, &dP[k].kref, &dP[k].mref, &dP[k].delta_kr_order,
                                   &dP[k].alpha_c, &dP[k].sigma_0, &dP[k].delta,
                                   &dP[k].m_crit, &dP[k].c1};
                *fld[p0 + k] = 1.0;
            }
            A.np = np;
            double g4[4] = {0, 0, 0, 0}, tmp[7];
            D::vjp<7, 7>(A, f, Pz, x, ly, tmp, g4);
            for (int k = 0; k < np; k++) {
                ga[p0 + k] = g4[k];
                gb[p0 + k] = D::pa_dot(pa, dP[k]);
            }
        }
        double gs = 0.0;
        for (int k = 0; k < NP; k++) gs = std::max(gs, std::fabs(ga[k] * pv[k]));
        for (int k = 0; k < NP; k++)
            if (gs > 0.0) wp = std::max(wp, std::fabs((ga[k] - gb[k]) * pv[k]) / gs);
    }
    if (worst2) {
        worst2[0] = wi;
        worst2[1] = wp;
    }
    return std::max(wi, wp);
}

// The written-out prepare reverse (prepare_vjp) against the seeded product
// vjp<7, NPC>(PrepF) on n random wet cells (speeds above and below the
// speed limit, m, chi and pb inside and outside their bounds), for the six
// state inputs and 14 parameter fields; worst2 as emu_src2_vjp_check.
extern "C" double emu_prepare_vjp_check(const dclaw_params* P, int n, unsigned seed,
                                        double* worst2)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    constexpr int NP = 14;
    dclaw_params dP[4];
    const double pv[NP] = {P->grav, P->rho_s, P->rho_f, P->mu, P->phi, P->kref, P->mref,
                           P->delta_kr_order, P->alpha_c, P->sigma_0, P->delta, P->m_crit,
                           P->c1, P->speed_limit};
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    const D::TP Pz = D::zero_params(*P);
    double wi = 0.0, wp = 0.0;
    for (int c = 0; c < n; c++) {
        double x[7], ly[D::NPC], la[7], lb[7];
        const double h = 0.01 + 1.5 * (rnd() + 1.0);
        x[0] = h;
        const double sp = (c % 4 == 0) ? 30.0 : 5.0;
        x[1] = (c % 9 == 0) ? 0.0 : sp * rnd() * h;
        x[2] = (c % 9 == 0) ? 0.0 : sp * rnd() * h;
        double m = 0.5 + 0.25 * rnd();
        if (c % 11 == 3) m = 1.0 + 0.05 * (rnd() + 1.0);
        if (c % 13 == 5) m = -0.05 * (rnd() + 1.0);
        x[3] = m * h;
        const double rho = P->rho_s * m + (1.0 - m) * P->rho_f;
        x[4] = (0.6 + 0.5 * rnd()) * rho * P->grav * h;
        if (c % 7 == 2) x[4] = -0.1 * rho * P->grav * h;
        x[5] = (0.5 + 0.6 * rnd()) * h;
        x[6] = 0.1 * rnd();
        for (int k = 0; k < D::NPC; k++) ly[k] = rnd();
        D::PrepF f{0.3 * rnd()};
        A.np = 0;
        D::vjp<7, D::NPC>(A, f, Pz, x, ly, la, nullptr);
        D::ParAdj pa;
        D::pa_zero(pa);
        D::prepare_vjp(*P, x, ly, lb, pa);
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < 6; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        wi = std::max(wi, scale > 0.0 ? diff / scale : diff);
        double ga[NP], gb[NP];
        for (int p0 = 0; p0 < NP; p0 += 4) {
            const int np = std::min(4, NP - p0);
            for (int k = 0; k < np; k++) {
                std::memset(&dP[k], 0, sizeof(dclaw_params));
                double* fld[NP] = {&dP[k].grav, &dP[k].rho_s, &dP[k].rho_f, &dP[k].mu,
                                   &dP[k].phi, &dP[k].kref, &dP[k].mref, &dP[k].delta_kr_order,
                                   &dP[k].alpha_c, &dP[k].sigma_0, &dP[k].delta,
                                   &dP[k].m_crit, &dP[k].c1, &dP[k].speed_limit};
                *fld[p0 + k] = 1.0;
            }
            A.np = np;
            double g4[4] = {0, 0, 0, 0}, tmp[7];
            D::vjp<7, D::NPC>(A, f, Pz, x, ly, tmp, g4);
            for (int k = 0; k < np; k++) {
                ga[p0 + k] = g4[k];
                gb[p0 + k] = D::pa_dot(pa, dP[k]);
            }
        }
        double gs = 0.0;
        for (int k = 0; k < NP; k++) gs = std::max(gs, std::fabs(ga[k] * pv[k]));
        for (int k = 0; k < NP; k++)
            if (gs > 0.0) wp = std::max(wp, std::fabs((ga[k] - gb[k]) * pv[k]) / gs);
    }
    if (worst2) {
        worst2[0] = wi;
        worst2[1] = wp;
    }
    return std::max(wi, wp);
}

// The written-out calc_taudir reverse (taudir_vjp) against the seeded
// product vjp<NTI, 3>(TaudirF) on n random five-cell stencils (kind c % 4:
// all wet and moving, a still centre, a dry centre, dry neighbours; beds,
// surfaces and flows random), the 37 inputs and grav.  Returns the largest
// difference relative to the output adjoints' size times the outputs' own
// scale (|ly| (1 + dx + dy): many entries are 0 in exact arithmetic and
// rounding residue in both forms), or to the largest input adjoint if
// larger; grav as a sensitivity |dL/dg| g against the same scale or the
// largest |dL/dx| |x|.
extern "C" double emu_taudir_vjp_check(const dclaw_params* P, int n, unsigned seed)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    dclaw_params dP[1];
    std::memset(&dP[0], 0, sizeof(dclaw_params));
    dP[0].grav = 1.0;
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    const D::TP Pz = D::zero_params(*P);
    double worst = 0.0;
    for (int c = 0; c < n; c++) {
        const int kind = c % 4;
        double x[D::NTI], ly[3], la[D::NTI], lb[D::NTI];
        D::TaudirF f;
        f.dx = 0.5 + (rnd() + 1.0);
        f.dy = 0.5 + (rnd() + 1.0);
        for (int s = 0; s < 5; s++) {
            f.b[s] = 0.5 * rnd();
            double h = 0.05 + (rnd() + 1.0);
            if ((kind == 2 && s == 0) || (kind == 3 && s > 0 && rnd() > 0.0)) h = 0.5e-3 * (rnd() + 1.0);
            x[7 * s] = h;
            x[7 * s + 1] = (kind == 1 && s == 0) ? 0.0 : 2.0 * rnd() * h;
            x[7 * s + 2] = (kind == 1 && s == 0) ? 0.0 : 2.0 * rnd() * h;
            x[7 * s + 3] = 0.6 * h;
            x[7 * s + 4] = 0.8 * 1500.0 * P->grav * h;
            x[7 * s + 5] = 0.3 * h;
            x[7 * s + 6] = 0.05 * rnd();
        }
        x[35] = 1500.0 + 500.0 * rnd();
        x[36] = 5000.0 * (rnd() + 1.1);
        for (int k = 0; k < 3; k++) ly[k] = rnd();
        A.np = 1;
        double ga[1] = {0.0};
        D::vjp<D::NTI, 3>(A, f, Pz, x, ly, la, ga);
        double gb = 0.0;
        D::taudir_vjp(*P, f.dx, f.dy, f.b, x, ly, lb, gb);
        // scale: the output adjoints times the outputs' own scale (tx, ty ~
        // dx, dy; fsphi ~ 1) per unit of input (depths ~ 1)
        double lmax = 0.0;
        for (int k = 0; k < 3; k++) lmax = std::max(lmax, std::fabs(ly[k]));
        double scale = lmax * (1.0 + f.dx + f.dy), diff = 0.0;
        double sens = std::fabs(ga[0] * P->grav);
        for (int k = 0; k < D::NTI; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
            sens = std::max(sens, std::fabs(la[k] * x[k]));
        }
        sens = std::max(sens, lmax * (1.0 + f.dx + f.dy));
        worst = std::max(worst, diff / scale);
        worst = std::max(worst, std::fabs((ga[0] - gb) * P->grav) / sens);
    }
    return worst;
}

// The hand-written rpn2 adjoint (rpn2_vjp) against the seeded tangent-linear
// vector-Jacobian product (30 side inputs and the grav, rho_f, kappa,
// beta_seg directions) on n random edges of 8 kinds, c % 8: both sides wet
// at moderate (0) and large (1) speeds, a dry right side below (2) and above
// (3, the wall) the left surface, the same on the left (4, 5), both wet at
// rest (6: the static / tausource branches) and near-critical speeds (7).
// Returns the largest difference relative to the largest input adjoint
// (parameters: as sensitivities |dL/dp| |p| against the largest of the
// four); worst[k] gets the same per kind when not null.
extern "C" double emu_rpn2_vjp_check(const dclaw_params* P, int n, unsigned seed, double* worst8)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    dclaw_params dP[4] = {};
    dP[0].grav = 1.0;
    dP[1].rho_f = 1.0;
    dP[2].kappa = 1.0;
    dP[3].beta_seg = 1.0;
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    A.np = 4;
    const D::TP Pz = D::zero_params(*P);
    const double g = P->grav;
    if (worst8)
        for (int k = 0; k < 8; k++) worst8[k] = 0.0;
    double worst = 0.0;
    for (int c = 0; c < n; c++) {
        const int kind = c % 8;
        D::Rpn2F f;
        f.ixy = 1 + ((c / 8) & 1);
        const bool xd = f.ixy == 1;
        double x[2 * D::NSD], ly[D::NW], la[2 * D::NSD], lb[2 * D::NSD];
        const double bedc = 2.0 * rnd();
        for (int s = 0; s < 2; s++) {
            double* X = x + s * D::NSD;
            bool dry = (s == 1 && (kind == 2 || kind == 3)) || (s == 0 && (kind == 4 || kind == 5));
            double h = 0.05 + 1.5 * (rnd() + 1.0);
            const double cs = sqrt(g * h);
            double u = (kind == 1 ? 4.0 : 0.8) * cs * rnd(), v = 0.5 * cs * rnd();
            if (kind == 6) u = v = 0.0;
            if (kind >= 2 && kind <= 5) {
                // the wet side at rest, flowing away from the dry side, or either way
                const int vk = (c / 16) % 4;
                if (vk == 0) u = v = 0.0;
                else if (vk == 1) u = (s == 0 ? -1.0 : 1.0) * 2.0 * cs * std::fabs(rnd());
                else u = 2.0 * cs * rnd();
            }
            if (kind == 7) u = (s == 0 ? 1.0 : -1.0) * cs * (1.0 + 0.02 * rnd());
            if (dry) {
                h = (c % 3 == 0) ? 0.0 : 0.5 * P->dry_tolerance * (rnd() + 1.0);
                u = v = 0.0;
            }
            const double m = dry ? 0.0 : 0.5 + 0.2 * rnd(), chi = dry ? 0.0 : 0.5 + 0.5 * rnd();
            const double rho = m * P->rho_s + (1.0 - m) * P->rho_f;
            X[D::SD_H] = h;
            X[xd ? D::SD_HU : D::SD_HV] = h * u;
            X[xd ? D::SD_HV : D::SD_HU] = h * v;
            X[D::SD_HM] = h * m;
            X[D::SD_P] = dry ? 0.0 : P->rho_f * g * h * (0.8 + 0.3 * rnd());
            X[D::SD_HCHI] = h * chi;
            X[D::SD_U] = u;
            X[D::SD_V] = v;
            X[D::SD_M] = m;
            X[D::SD_CHI] = chi;
            X[D::SD_BDIF] = 0.1 * (rnd() + 1.0);
            X[D::SD_TAUDIR] = rnd();
            X[D::SD_FS] = 0.5 * (rnd() + 1.0);
            X[D::SD_RHO] = dry ? P->rho_f : rho;
            X[D::SD_TAU] = dry ? 0.0 : 2e3 * h * (rnd() + 1.0);
        }
        // beds: the dry side's bed below the wet surface (2, 4) or well above it (3, 5)
        f.bedL = bedc;
        f.bedR = bedc + 0.3 * rnd();
        const double hw = kind == 2 || kind == 3 ? x[D::SD_H] : x[D::NSD + D::SD_H];
        if (kind == 3) f.bedR = bedc + x[D::SD_BDIF] - x[D::NSD + D::SD_BDIF] + hw + 1.0;
        if (kind == 2) f.bedR = bedc - 1.0;
        if (kind == 5) f.bedL = f.bedR + x[D::NSD + D::SD_BDIF] - x[D::SD_BDIF] + hw + 1.0;
        if (kind == 4) f.bedL = f.bedR - 1.0;
        for (int k = 0; k < D::NW; k++) ly[k] = rnd();
        double ga[4] = {0, 0, 0, 0}, gb[4] = {0, 0, 0, 0};
        D::vjp<2 * D::NSD, D::NW>(A, f, Pz, x, ly, la, ga);
        D::rpn2_vjp(A, f, x, ly, lb, gb);
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < 2 * D::NSD; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        double w = scale > 0.0 ? diff / scale : diff;
        // (a parameter that is 0 is scaled as 1)
        const double pm[4] = {std::max(std::fabs(P->grav), 1.0), std::max(std::fabs(P->rho_f), 1.0),
                              std::max(std::fabs(P->kappa), 1.0),
                              std::max(std::fabs(P->beta_seg), 1.0)};
        double gs = 0.0;
        for (int k = 0; k < 4; k++) gs = std::max(gs, std::fabs(ga[k] * pm[k]));
        for (int k = 0; k < 4; k++) {
            const double d = std::fabs((ga[k] - gb[k]) * pm[k]);
            if (gs > 0.0) w = std::max(w, d / gs);
            else w = std::max(w, d);
        }
        worst = std::max(worst, w);
        if (worst8) worst8[kind] = std::max(worst8[kind], w);
    }
    return worst;
}

// The hand-written riemanntype adjoint (riemanntype_vjp: s1m, s2m with
// respect to hL, hR, uL, uR, g) against the seeded tangent-linear
// riemanntype on n random states; counts[4] gets how many fell in each
// branch (a dry side, two rarefactions, two shocks, mixed).  Returns the
// largest difference relative to the largest input adjoint.
extern "C" double emu_riemanntype_vjp_check(double drytol, int n, unsigned seed, int* counts)
{
    namespace T = dclaw_tl;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    for (int k = 0; k < 4; k++) counts[k] = 0;
    double worst = 0.0;
    for (int c = 0; c < n; c++) {
        double v[5];
        v[0] = (c % 11 == 0) ? 0.5 * drytol * (rnd() + 1.0) : 0.02 + 2.0 * (rnd() + 1.0);
        v[1] = (c % 13 == 0) ? 0.0 : 0.02 + 2.0 * (rnd() + 1.0);
        const double g = 9.81 * (0.5 + 0.4 * rnd());
        v[4] = g;
        const double cs = sqrt(g * 0.5 * (v[0] + v[1]));
        v[2] = 3.0 * cs * rnd();
        v[3] = 3.0 * cs * rnd();
        const double s1b = rnd(), s2b = rnd();
        double la[5];
        for (int k = 0; k < 5; k++) {
            T::Dual d[5];
            for (int a = 0; a < 5; a++) d[a] = T::Dual(v[a], a == k ? 1.0 : 0.0);
            const T::RType r = T::riemanntype(d[0], d[1], d[2], d[3], T::Dual(drytol), d[4]);
            la[k] = s1b * r.s1m.d + s2b * r.s2m.d;
        }
        double lb[5] = {0, 0, 0, 0, 0};
        dclaw_adj::riemanntype_vjp(v[0], v[1], v[2], v[3], drytol, g, s1b, s2b, lb[0], lb[1], lb[2],
                                   lb[3], lb[4]);
        // branch of this state (the same tests as riemanntype)
        const double hmin = std::min(v[0], v[1]), hmax = std::max(v[0], v[1]);
        const double delu = v[3] - v[2];
        int br;
        if (hmin <= drytol) br = 0;
        else if (delu + 2.0 * (sqrt(g * hmin) - sqrt(g * hmax)) > 0.0) br = 1;
        else if (delu + (hmax - hmin) * (sqrt(.5 * g * (hmax + hmin) / (hmax * hmin))) <= 0.0) br = 2;
        else br = 3;
        counts[br]++;
        double scale = 0.0, diff = 0.0;
        for (int k = 0; k < 5; k++) {
            scale = std::max(scale, std::fabs(la[k]));
            diff = std::max(diff, std::fabs(la[k] - lb[k]));
        }
        worst = std::max(worst, scale > 0.0 ? diff / scale : diff);
    }
    return worst;
}

// The K-direction products (vjpk) against the one-direction seeded products
// (vjp) of the three stages that use them -- prepare (F1), calc_taudir (F2)
// and src2 -- on n random wet cells with np parameter directions (phi, kref,
// m_crit, mu): returns the number of input or parameter adjoints that differ
// in any bit.
extern "C" int emu_vjpk_check(const dclaw_params* P, int n, int np, unsigned seed)
{
    namespace D = dclaw_adj;
    unsigned long long st = seed * 2862933555777941757ull + 3037000493ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)((st >> 11) & ((1ull << 53) - 1)) / (double)(1ull << 53) * 2.0 - 1.0;
    };
    dclaw_params dP[4] = {};
    dP[0].phi = 1.0;
    dP[1].kref = 1.0;
    dP[2].m_crit = 1.0;
    dP[3].mu = 1.0;
    D::AdjArgs A{};
    A.P = *P;
    A.dP = dP;
    A.np = np;
    const D::TP Pz = D::zero_params(*P);
    const double g = P->grav;
    auto state = [&](double* q, bool wet) {
        const double h = wet ? 0.05 + 2.0 * (rnd() + 1.0) : 0.5 * P->dry_tolerance * (rnd() + 1.0);
        const double m = 0.55 + 0.1 * rnd();
        q[0] = h;
        q[1] = h * 2.0 * rnd();
        q[2] = h * 2.0 * rnd();
        q[3] = h * m;
        q[4] = P->rho_f * g * h * (0.7 + 0.3 * rnd());
        q[5] = h * 0.5 * (rnd() + 1.0);
        q[6] = 0.05 * rnd();
    };
    int bad = 0;
    auto cmp = [&](const double* a, const double* b, int m) {
        for (int k = 0; k < m; k++)
            if (std::memcmp(&a[k], &b[k], sizeof(double)) != 0) bad++;
    };
    for (int c = 0; c < n; c++) {
        double ga[4] = {0, 0, 0, 0}, gb[4] = {0, 0, 0, 0};
        {
            double q[7], ly[D::NPC], la[7], lb[7];
            state(q, c % 5 != 0);
            for (int k = 0; k < D::NPC; k++) ly[k] = rnd();
            D::PrepF f{rnd()};
            D::vjp<7, D::NPC>(A, f, Pz, q, ly, la, ga);
            D::vjpk<7, D::NPC>(A, f, q, ly, lb, gb);
            cmp(la, lb, 7);
        }
        {
            double x[D::NTI], ly[3], la[D::NTI], lb[D::NTI];
            for (int s = 0; s < 5; s++) state(x + 7 * s, s == 0 || c % 3 != 0);
            x[35] = P->rho_f + 1000.0 * (rnd() + 1.0);
            x[36] = 1e3 * (rnd() + 1.0);
            for (int k = 0; k < 3; k++) ly[k] = rnd();
            D::TaudirF f;
            f.dx = 1.0 + 0.5 * rnd();
            f.dy = 1.0 + 0.5 * rnd();
            for (int s = 0; s < 5; s++) f.b[s] = rnd();
            D::vjp<D::NTI, 3>(A, f, Pz, x, ly, la, ga);
            D::vjpk<D::NTI, 3>(A, f, x, ly, lb, gb);
            cmp(la, lb, D::NTI);
        }
        {
            double q[7], ly[7], la[7], lb[7];
            state(q, true);
            for (int k = 0; k < 7; k++) ly[k] = rnd();
            D::Src2F f{0.01 * (rnd() + 1.5), 0.5 * (rnd() + 1.0)};
            D::vjp<7, 7>(A, f, Pz, q, ly, la, ga);
            D::vjpk<7, 7>(A, f, q, ly, lb, gb);
            cmp(la, lb, 7);
        }
        cmp(ga, gb, np);
    }
    return bad;
}
