/*
 * dclaw_oracle.h -- CPU restatement of D-Claw's per-patch finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (dclaw_amd/).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Arrays follow the reference layout exactly: Fortran column-major, component
 * fastest, q(meqn=7, 1-mbc:mx+mbc, 1-mbc:my+mbc), aux(maux, same).
 * q components (digclaw_module.f90:96-102): 1 h, 2 hu, 3 hv, 4 hm, 5 pb,
 * 6 hchi, 7 bdif.  aux components for coordinate_system=1
 * (digclaw_module.f90:83-110): 1 b, 2 phi, 3 theta, 4 fsphi, 5 taudir_x,
 * 6 taudir_y.
 */
#ifndef DCLAW_ORACLE_H
#define DCLAW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Module globals of geoclaw_module / digclaw_module / amr_module that the
 * reference reads on the step path.  Field order is mirrored by the Python
 * ctypes Structure in oracle/oracle.py. */
typedef struct oracle_params {
    double grav;                /* geoclaw_module grav */
    double dry_tolerance;       /* geoclaw_module dry_tolerance */
    double speed_limit;         /* geoclaw_module speed_limit (b4step2.f90:78) */
    double rho_s, rho_f, m_crit, mref, kref;
    double phi;                 /* radians (digclaw_module.f90:136) */
    double delta, mu, alpha_c, c1, sigma_0;
    double kappa;               /* 1 (digclaw_module.f90:113) */
    double beta_seg, delta_kr_order;
    double manning_coefficient; /* manning_coefficient(1) (src2.f90:82) */
    double friction_depth, manning_max;
    int src2method, alphamethod, rarecorrectortest, segregation;
    int friction_forcing, dd_manning;
    int order;                  /* method(2): 1 Godunov, 2 with limited corrections */
    int transverse;             /* method(3): 0, 1, 2 */
    int mthlim[5];              /* limiter per wave family */
    int mwaves;                 /* 5 */
    /* entrainment (src2.f90:253-263, entrainment.f90:27-100) */
    double entrainment_rate, me, chie;
    int entrainment;            /* 0 or 1 */
    int entrainment_method;     /* 0: ent_dclaw4; 1: no update (the reference's default) */
    int curvature;              /* src2.f90:150-156 */
} oracle_params;

/* ---- point routines (for unit / known-answer tests) ---- */
void oracle_qfix(const oracle_params* p, double* q6, double gz, double* prim5);
void oracle_setvars(const oracle_params* p, const double* in7, double* out6);
void oracle_riemanntype(double hL, double hR, double uL, double uR, int maxiter,
                        double drytol, double g, double* out3, int* rare2);
void oracle_riemann_solver(const oracle_params* p, int ixy, const double* in30,
                           int wallprob, double* sw3, double* fw18);

/* ---- slice routines (1-D arrays with 1-mbc:mx+mbc extents) ---- */
void oracle_rpn2(const oracle_params* p, int ixy, int mx, int mbc, int maux,
                 double* q1d, const double* aux2, double* fwave, double* s,
                 double* amdq, double* apdq);
void oracle_rpt2(const oracle_params* p, int ixy, int imp, int mx, int mbc, int maux,
                 const double* q1d, const double* aux1, const double* aux2,
                 const double* aux3, const double* asdq, double* bmasdq, double* bpasdq);
void oracle_limiter(const oracle_params* p, int mx, int mbc, double* fwave, const double* s);
void oracle_flux2(const oracle_params* p, int ixy, int maxm, int mx, int mbc, int maux,
                  double* q1d, double* dtdx1d, const double* aux1, const double* aux2,
                  const double* aux3, double* faddm, double* faddp, double* gaddm,
                  double* gaddp, double* cfl1d);

/* ---- patch routines ---- */
void oracle_calc_taudir(oracle_params* p, int mx, int my, int mbc, int maux,
                        double dx, double dy, const double* q, double* aux);
void oracle_b4step2(oracle_params* p, int mx, int my, int mbc, int maux,
                    double dx, double dy, double* q, double* aux);
void oracle_step2(const oracle_params* p, int mx, int my, int mbc, int maux,
                  const double* qold, const double* aux, double dx, double dy, double dt,
                  double* cflgrid, double* fm, double* fp, double* gm, double* gp);
void oracle_src2(const oracle_params* p, int mx, int my, int mbc, int maux, double dx,
                 double dy, double* q, const double* aux, double dt);
/* One stepgrid: b4step2, step2, conservative update, src2.  `work` must hold
 * 4*7*(mx+2mbc)*(my+2mbc) doubles (fm, fp, gm, gp).  Returns cflgrid. */
double oracle_stepgrid(oracle_params* p, int mx, int my, int mbc, int maux,
                       double dx, double dy, double dt, double* q, double* aux,
                       double* work);
/* Physical boundary fill, bc codes as Clawpack: 1 extrap, 3 wall.
 * bc[0..3] = xlower, xupper, ylower, yupper. */
void oracle_bc_fill(int mx, int my, int mbc, int meqn, double* q, const int* bc);
/* Fixed-patch time loop: bc fill then stepgrid, nsteps times, with
 * dt_{n+1} = min(dt_n*cfl_desired/cfl_n, dt_max).  dt_inout holds dt_0 on
 * entry and the next dt on return; cfl_hist (nsteps) and dt_hist (nsteps)
 * may be NULL.  Returns the number of steps taken. */
int oracle_run(oracle_params* p, int mx, int my, int mbc, int maux, double dx,
               double dy, const int* bc, double* q, double* aux, int nsteps,
               double* dt_inout, double cfl_desired, double dt_max,
               double* cfl_hist, double* dt_hist, int nthreads);

/* nsteps of oracle_run on each of npatch independent patches (P[k],
 * mx[k], ...; per-patch dt in dt_inout[k]) over nthreads OpenMP threads
 * (dynamic schedule) when the library is built with OpenMP. */
int oracle_run_patches(oracle_params* P, int npatch, const int* mx, const int* my, int mbc,
                       const int* maux, const double* dx, const double* dy, const int* bc,
                       double** q, double** aux, int nsteps, double* dt_inout,
                       double cfl_desired, double dt_max, int nthreads);
/* > 1 if this build runs oracle_run_patches on several threads. */
int oracle_openmp_threads(void);

/* Time loop pieces (device-resident loop of the build, restated):
 * oracle_dt_rule: the dt rule after a step of length dt_used taken at time
 * t_before with CFL number cfl: t_after = t_before + dt_used;
 * dt_next = min(dt_used*cfl_desired/cfl (infinity if cfl = 0), dt_max)
 * (GeoClaw stepgrid dtnew, tick.f:499); then the output-time hit of
 * tick.f:135-144: if t_after < tout and t_after + 1.001*dt_next >= tout,
 * dt_next = tout - t_after and *hit = 1 (else 0). */
void oracle_dt_rule(double dt_used, double cfl, double t_before, double cfl_desired,
                    double dt_max, double tout, double* t_after, double* dt_next, int* hit);

/* Gauges: for each (xg, yg), the 7 q components and eta = h + (aux(1) - bdif),
 * bilinear in the four cells whose centres surround the gauge when all four
 * are wet (h > dry_tolerance), else the values of the cell containing the
 * gauge (the convention of gradient_particle_filter/flume_solver.py:91-99,
 * two cells in 1-D; GeoClaw's gauges_module is not in the reference tree).
 * out: 8 doubles per gauge.  The four cells must be interior: x in
 * [xlower + dx/2, xlower + (mx - 1/2) dx), same in y. */
void oracle_gauges(int mx, int my, int mbc, int maux, const double* q, const double* aux,
                   double xlower, double ylower, double dx, double dy, double dry_tolerance,
                   const double* xg, const double* yg, int ngauges, double* out);

/* AMR transfers (coordinate_system 1, mcapa = 0), arrays 0-based including
 * ghosts, component fastest.  update.f90:112-203, fine -> coarse: coarse cells
 * (ci0 + a, cj0 + b), a < ni, b < nj, from the rx x ry fine blocks at
 * (fi0 + a rx, fj0 + b ry). */
void oracle_amr_average(double* qc, const double* auxc, int mauxc, int mitc, const double* qf,
                        const double* auxf, int mauxf, int mitf, int ci0, int cj0, int ni, int nj,
                        int fi0, int fj0, int rx, int ry, double dry_tolerance);

/* filval.f90:190-381, coarse -> fine: the interpolation loop over the coarse
 * patch enlarged by one cell (mic x mjc) into the fine patch (mitot columns,
 * nghost = 2); needs_set[fine cell] != 0 marks the cells NEEDS_TO_BE_SET. */
void oracle_amr_interp(const double* valc, const double* auxc, int mauxc, int mic, int mjc,
                       double* val, const double* auxf, int mauxf, int mitot,
                       const unsigned char* needs_set, int rx, int ry, double dry_tolerance,
                       double sea_level, double rho_f, double grav);

/* filpatch.f90:182-187: coarse rectangle (iplo, iphi, jplo, jphi) around
 * the unset cells isl..isr x jsb..jst. */
void oracle_amr_coarse_box(int isl, int isr, int jsb, int jst, int rx, int ry, int* box);

/* filpatch.f90:256-520: interpolation of a region's unset cells from the
 * coarse rectangle (mxc x myc from (iplo, jplo)).  0 on success, -1 if an
 * unset cell is outside the rectangle's interior. */
int oracle_amr_filpatch(const double* valcrse, const double* auxcrse, int mauxc, int mxc, int myc,
                        int iplo, int jplo, double* valbig, const double* aux, int maux, int mx,
                        int r0, int c0, int ilo, int ihi, int jlo, int jhi,
                        const unsigned char* flaguse, int rx, int ry, double xlower,
                        double ylower, double dx_fine, double dy_fine, double dx_coarse,
                        double dy_coarse, double sea_level, double dry_tolerance, double rho_f,
                        double grav);

/* AMRClaw steps of a region fill (not in the reference; restated). */
void oracle_amr_intfil(double* q, double* aux, int maux, int mx, int r0, int c0, int ilo,
                       int ihi, int jlo, int jhi, unsigned char* flag, const double* sq,
                       const double* sq_old, double w_old, const double* saux, int smaux,
                       int silo, int sjlo, int smx, int smy, int copy_aux);
int oracle_amr_trimbd(const unsigned char* flag, int ilo, int ihi, int jlo, int jhi, int* box);
void oracle_amr_bc2amr(double* q, double* aux, int maux, int mx, int r0, int c0, int ilo,
                       int ihi, int jlo, int jhi, int iregsz, int jregsz, const int* bc);

/* setaux.f90:59-273 for a patch with 2 ghost cells, the topography integral
 * supplied as bed (used where bed_set is 0). */
void oracle_amr_setaux(double* aux, int maux, int mx, int my, int ilo, int jlo,
                       const unsigned char* bed_set, const double* bed, double xlower,
                       double xupper, double ylower, double yupper, double dx, double dy,
                       double phi, double theta);

#ifdef __cplusplus
}
#endif
#endif
