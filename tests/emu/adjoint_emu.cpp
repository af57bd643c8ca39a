// adjoint_emu.cpp -- the adjoint step's stages (dclaw_adjoint.h) and the
// hand-written reverses run on the host, for the CPU tests: the stage
// sequence over the grid and the checks of each written-out reverse against
// the seeded tangent-linear products.  Built with strip_emu.cpp into one
// library (tests/conftest.py).  Test tool only.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dclaw_adjoint.h"

using namespace dclaw;

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

