/*
 * dclaw_hip.h -- C-ABI of the MI355X (gfx950) D-Claw per-patch step.
 *
 * Drop-in boundary for the path GeoClaw's stepgrid drives each step:
 *   bound (ghost fill) -> b4step2 -> step2 (rpn2/rpt2/flux2 sweeps) ->
 *   conservative update -> src2 -> dtnew = dt*cfl_desired/cflgrid.
 *
 * Layout is the Clawpack contract, unchanged: q(meqn=7, 1-mbc:mx+mbc,
 * 1-mbc:my+mbc) and aux(maux, same), Fortran column-major, component
 * fastest (rpn2_dclaw.f:67-74, src2.f90:47-49).  q components: 1 h, 2 hu,
 * 3 hv, 4 hm, 5 pb, 6 hchi, 7 bdif (digclaw_module.f90:96-102).  aux(1) is
 * the bed b; the stepping entry points read aux(1) (stride maux), since phi
 * comes from the parameter record and fsphi/taudir are recomputed inside the
 * step exactly as b4step2 would (b4step2.f90:111).  With entrainment on they
 * also read aux(7) = h_e, the erodible depth (src2.f90:253-263), so such
 * calls need maux >= 7: the device entry points receive the parameters as a
 * device record and cannot check this, the caller must (the host drop-ins,
 * which hold a host copy of the parameters, fail with DCLAW_EINVAL).  mbc
 * must be 2.
 *
 * All pointers passed to the device entry points are device (HBM) pointers;
 * all take an explicit hipStream_t (passed as void*).  Every call returns
 * 0 on success or a negative DCLAW_E* code; dclaw_last_error() gives text.
 * Parameters are an explicit record (the reference's module globals,
 * digclaw_module.f90:15-52, geoclaw_module grav/dry_tolerance/speed_limit,
 * amr_module method/mthlim), not globals.
 */
#ifndef DCLAW_HIP_H
#define DCLAW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCLAW_OK 0
#define DCLAW_EINVAL (-1)
#define DCLAW_EHIP (-2)
#define DCLAW_ENOMEM (-3)

#define DCLAW_MEQN 7
#define DCLAW_MBC 2

/* Clawpack bc codes (bc2amr): */
#define DCLAW_BC_EXTRAP 1
#define DCLAW_BC_WALL 3

typedef struct dclaw_params {
    double grav;                /* geoclaw_module grav */
    double dry_tolerance;       /* geoclaw_module dry_tolerance */
    double speed_limit;         /* geoclaw_module speed_limit (b4step2.f90:78) */
    double rho_s, rho_f, m_crit, mref, kref;
    double phi;                 /* basal friction angle, radians (digclaw_module.f90:136) */
    double delta, mu, alpha_c, c1, sigma_0;
    double kappa;               /* earth-pressure coefficient, 1 (digclaw_module.f90:113) */
    double beta_seg, delta_kr_order;
    double manning_coefficient; /* manning_coefficient(1) (src2.f90:82) */
    double friction_depth, manning_max;
    int src2method;             /* -1, 0 or 1 */
    int alphamethod, rarecorrectortest, segregation;
    int friction_forcing, dd_manning;
    int order;                  /* method(2): 1 or 2 */
    int transverse;             /* method(3): 0, 1 or 2 */
    int mthlim[5];              /* limiter per wave family (0 none, 1 minmod, 2 superbee,
                                   3 van Leer, 4 MC, 5 Beam-Warming) */
    int mwaves;                 /* 5 */
    /* entrainment (src2.f90:253-263, entrainment.f90:27-100) */
    double entrainment_rate, me, chie;
    int entrainment;            /* 0 or 1 */
    int entrainment_method;     /* 0: ent_dclaw4; 1: no update (the reference's default) */
    int curvature;              /* 1: bed-curvature acceleration in src2 (src2.f90:150-156);
                                   only dclaw_stepgrid_curvature_dev and the host drop-ins
                                   apply it (needs aux(3) = theta: maux >= 3) */
} dclaw_params;

/* One patch of a (possibly batched) step.  All pointers are device pointers. */
typedef struct dclaw_patch {
    const double* q_in;         /* q(7, mx+4, my+4): state at t_n, ghosts filled */
    double* q_out;              /* q(7, mx+4, my+4): state at t_n+dt (may not alias q_in) */
    const double* aux;          /* aux(maux, mx+4, my+4); aux(1) = b is read, and
                                   aux(7) = h_e with entrainment on (maux >= 7) */
    int mx, my, maux, member;   /* member indexes params[], dt[], cfl[] */
    double dx, dy;
} dclaw_patch;

/* One patch of a batched adjoint step (dclaw_adjoint_batched). */
typedef struct dclaw_adj_patch {
    const double* q_in;         /* q_n, layout of q */
    const double* aux;          /* aux(maux, mx+4, my+4) */
    const double* lam_out;      /* lambda_{n+1} */
    double* lam_in;             /* lambda_n (written) */
    double* q_out;              /* q_{n+1} (written if not null) */
    double* work;               /* dclaw_adjoint_workspace_doubles(mx, my), zero-filled once */
    double* grad;               /* np doubles (written if np > 0) */
    int mx, my, maux, member;   /* member indexes params[], dt[] and the np dparams */
    double dx, dy;
} dclaw_adj_patch;

const char* dclaw_last_error(void);
int dclaw_version(void);

/* Interior columns per strip of the fused step (one wavefront per strip
 * and row segment). */
int dclaw_strip_width(void);

/* Rows per segment (the tile height) of the single-patch launches; returns
 * the previous value or a negative error.  Results do not depend on it; the
 * autotuner (dclaw_amd/autotune.py) picks it per workload.  No reference
 * counterpart (tiling of the GPU launch). */
int dclaw_set_seg_rows(int seg_rows);

/* Ghost-cell fill of q (bc2amr extrap / wall; bc[4] = xlo, xhi, ylo, yhi),
 * on device memory.  Replaces AMRClaw's bound/bc2amr for one patch. */
int dclaw_bc_fill(double* q, int mx, int my, const int* bc, void* stream);

/* Fused stepgrid for ONE patch on device memory: b4step2 + step2 + update
 * + src2 for interior cells of q_out (ghost cells of q_out receive the
 * b4step2-adjusted values, as the reference leaves them in q).  dt_dev and
 * cfl_dev are device scalars; cfl_dev must be zeroed by the caller (the
 * kernel takes a max into it: cflgrid of stepgrid).  params_dev is a device
 * dclaw_params.  Replaces GeoClaw stepgrid (b4step2.f90:7, flux2fw.f:8,
 * rpn2_dclaw.f:24, rpt2_dclaw.f:16, src2.f90:27). */
int dclaw_stepgrid_dev(const double* q_in, double* q_out, const double* aux, int maux,
                       int mx, int my, double dx, double dy, const double* dt_dev,
                       const dclaw_params* params_dev, double* cfl_dev, void* stream);

/* The same step with the physical boundary conditions applied on load
 * (bc[4] host array: xlo, xhi, ylo, yhi; 1 extrap, 3 wall): ghost cells of
 * q_in are not read, so no separate bound/bc2amr pass is needed.  Result is
 * identical to dclaw_bc_fill followed by dclaw_stepgrid_dev. */
int dclaw_stepgrid_bc_dev(const double* q_in, double* q_out, const double* aux, int maux,
                          int mx, int my, double dx, double dy, const int* bc,
                          const double* dt_dev, const dclaw_params* params_dev,
                          double* cfl_dev, void* stream);

/* Tangent-linear (forward-mode) fused step: the step of dclaw_stepgrid_bc_dev
 * (bc may be NULL: stored ghosts) on q_in together with the tangent dq_in
 * (same layout), with the parameters' tangent record dparams_dev (a device
 * dclaw_params whose floating fields hold the derivative of each parameter
 * along the direction; integer fields ignored).  q_out and cfl_dev equal
 * dclaw_stepgrid_bc_dev's bit for bit; dq_out = (dq_out/dq_in) dq_in +
 * (dq_out/dparams) dparams at dt_dev (dt is not differentiated).  At a branch
 * of the scheme the tangent is the derivative of the branch the values take.
 * aux (the bed, h_e) carries no tangent; aux(4..6) is not written.  No
 * reference counterpart: the reference has no derivative of the step (the
 * gradient-based workflows need it; this is the forward-mode building block,
 * dclaw_adjoint_step_dev the reverse-mode one). */
int dclaw_stepgrid_tl_dev(const double* q_in, const double* dq_in, double* q_out, double* dq_out,
                          const double* aux, int maux, int mx, int my, double dx, double dy,
                          const int* bc, const double* dt_dev, const dclaw_params* params_dev,
                          const dclaw_params* dparams_dev, double* cfl_dev, void* stream);

/* Tiles (58-column strip x segment) of one patch at the current segment
 * height: the smallest tile_flags workspace of the split step below. */
int dclaw_tile_count(int mx, int my);

/* Workspace ints (3 * dclaw_tile_count(mx, my) + 4) with which the split step
 * below runs in the mode set by dclaw_set_split_mode.  No reference
 * counterpart. */
int dclaw_split_workspace_size(int mx, int my);

/* Workspace ints (4 * dclaw_tile_count(mx, my) + 12) of the "resident"
 * wet-first split step: see dclaw_stepgrid_split_dev.  No reference
 * counterpart. */
int dclaw_split_resident_size(int mx, int my);

/* Mode of the split step when its workspace holds dclaw_split_workspace_size
 * ints: 1 = one strip launch with the tiles the previous call flagged first
 * (the workspace carries the flags between calls; any contents give a valid
 * order; the default), 2 = a separate dry pass first, 3 = one strip launch
 * in which the segments of 4-segment groups the previous call flagged run
 * singly and first and the other groups as one tall segment each; with fewer
 * ints (at least dclaw_tile_count) or mode 0: one strip launch in tile order.
 * Results do not depend on it; a library-wide setting like the segment
 * height.  Returns the previous mode or a negative error.  No reference
 * counterpart. */
int dclaw_set_split_mode(int mode);

/* Rows per wavefront of the split step's dry pass (rounded down to whole
 * segments, at least one, at most 32); returns the previous value or a
 * negative error.  Results do not depend on it.  No reference counterpart. */
int dclaw_set_dry_rows(int rows);

/* The step of dclaw_stepgrid_bc_dev (bc may be NULL: stored ghosts) with src2
 * split into a second pass: the strip kernel writes the updated state,
 * zeroes the cells src2 would zero (h <= dry_tolerance, src2.f90:103-111)
 * and flags its tile when a cell needs src2; then src2 runs, at full
 * occupancy, over the interior cells of the flagged tiles only.  Same
 * arithmetic, same result as the fused step.  tile_flags_dev: device
 * workspace of n_flags >= dclaw_tile_count(mx, my) ints, written every call
 * (no initialisation needed); its first dclaw_tile_count ints are the src2
 * flags of the tiles (segment * strips + strip).  With n_flags >=
 * dclaw_split_workspace_size(mx, my) the mode set by dclaw_set_split_mode
 * applies: the dry pass (a light kernel writes the dry state of every tile
 * whose 1-cell neighbourhood is dry and lists the others, and the strip
 * kernel and src2 run over that list only) or the wet-first tile order (the
 * previous call's flagged tiles dispatched first).  The rest of the
 * workspace holds the list or order and counters.  Same result in every
 * mode.  With n_flags >= dclaw_split_resident_size(mx, my), the wet-first
 * order and boundary conditions on load (every bc code nonzero), tiles whose
 * neighbourhood stayed dry on the last three inputs are neither read nor
 * written: for a time loop that swaps q_in and q_out every call and changes
 * neither in between (the library checks the swap and the segment height on
 * the device and starts over when they break; a caller that writes q between
 * calls zeroes the workspace).  No reference counterpart (a GPU scheduling
 * choice). */
int dclaw_stepgrid_split_dev(const double* q_in, double* q_out, const double* aux, int maux,
                             int mx, int my, double dx, double dy, const int* bc,
                             const double* dt_dev, const dclaw_params* params_dev,
                             double* cfl_dev, int* tile_flags_dev, int n_flags, void* stream);

/* The split step of dclaw_stepgrid_split_dev with the tile-flag workspace
 * taken from a static array in the library's device memory instead of a
 * caller buffer: no allocation and no workspace argument.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Limit: dclaw_split_workspace_size(mx, my) (at the current
 * segment height) must not exceed dclaw_static_workspace_ints() (262144
 * ints, 1 MiB); larger patches fail with DCLAW_EINVAL.  Each GPU holds one
 * such array until the process ends and every call on that GPU shares it,
 * so calls that overlap on one GPU (concurrent streams or threads) race;
 * the array keeps the last call's flags, which only reorder the next call's
 * tiles (results never depend on them). */
int dclaw_stepgrid_split_static_dev(const double* q_in, double* q_out, const double* aux,
                                    int maux, int mx, int my, double dx, double dy,
                                    const int* bc, const double* dt_dev,
                                    const dclaw_params* params_dev, double* cfl_dev,
                                    void* stream);

/* Ints of the static workspace of dclaw_stepgrid_split_static_dev. */
int dclaw_static_workspace_ints(void);

/* Handle-based use of the single-patch step: a dclaw_handle carries its own
 * tuning settings (segment height, split mode, dry-pass rows; at creation a
 * copy of the library-wide values that dclaw_set_seg_rows, dclaw_set_split_mode
 * and dclaw_set_dry_rows set) and owns the resident split step's tile
 * workspace, so time loops with different settings run side by side and no
 * caller passes a workspace.  The library-wide calls above stay as they are.
 * Replaces the module-global configuration of the reference's per-patch step
 * (the reference reads it from digclaw_module, digclaw_module.f90:15-52, and
 * GeoClaw's geo_data): the physical parameters still come in the
 * dclaw_params record of each call.  One handle serves one time loop: its
 * calls come from one thread, in order, on one stream.  Every function
 * returns a negative error for a null handle. */
typedef struct dclaw_handle_s* dclaw_handle;

/* New handle in *out (settings copied from the library-wide values; no
 * device memory until its first step). */
int dclaw_create(dclaw_handle* out);

/* Frees the handle and its device workspace (waiting for the device's work);
 * a null handle is accepted. */
int dclaw_destroy(dclaw_handle h);

/* The handle's settings; each returns the previous value (the ranges of the
 * library-wide setters). */
int dclaw_handle_set_seg_rows(dclaw_handle h, int seg_rows);
int dclaw_handle_set_split_mode(dclaw_handle h, int mode);
int dclaw_handle_set_dry_rows(dclaw_handle h, int rows);

/* Tiles of an mx x my patch at the handle's segment height. */
int dclaw_handle_tile_count(dclaw_handle h, int mx, int my);

/* The split step of dclaw_stepgrid_split_dev with the handle's settings and
 * its own resident workspace (allocated, or grown, on the current device at
 * the first step that needs it): the wet-first tile order and, for a loop
 * that swaps q_in and q_out every call with boundary conditions on load
 * (every bc code nonzero), the skipping of tiles that stayed dry.  Same
 * result as every other form of the step.  Replaces GeoClaw stepgrid
 * (b4step2.f90:7, flux2fw.f:8, rpn2_dclaw.f:24, rpt2_dclaw.f:16,
 * src2.f90:27). */
int dclaw_handle_stepgrid_dev(dclaw_handle h, const double* q_in, double* q_out,
                              const double* aux, int maux, int mx, int my, double dx, double dy,
                              const int* bc, const double* dt_dev,
                              const dclaw_params* params_dev, double* cfl_dev, void* stream);

/* Tells the handle that its loop's state was written outside its steps (the
 * dry-tile skipping starts over at the next step).  No reference
 * counterpart. */
int dclaw_handle_state_written(dclaw_handle h, void* stream);

/* The step with src2's curvature term (params.curvature = 1; src2.f90:133-156).
 * The reference sets gz once per src2 call (:87) and adds each wet cell's
 * curvature acceleration max(u^2 b_xx + v^2 b_yy + 2uv b_xy + u^2 dtheta, 0)
 * to it without a reset (:156), so a cell's gravity includes every earlier
 * cell's in src2's loop order (j outer, i inner).  This entry point forms
 * that running sum exactly in that order (one workgroup), between the step
 * without src2 and a parallel src2 that gives each cell its own gz; the
 * other stepping entry points, which cannot, mark a wet cell's pb NaN when
 * curvature is on.  aux(3) = theta is read (maux >= 3), and aux(4..6) are
 * not written.  work: device workspace of
 * dclaw_curvature_workspace_doubles(mx, my) doubles, overwritten.  Curvature
 * with entrainment is rejected (src2 would read beds eroded earlier in the
 * same sweep).  Replaces stepgrid with src2's curvature branch. */
size_t dclaw_curvature_workspace_doubles(int mx, int my);
int dclaw_stepgrid_curvature_dev(const double* q_in, double* q_out, const double* aux, int maux,
                                 int mx, int my, double dx, double dy, const int* bc,
                                 const double* dt_dev, const dclaw_params* params_dev,
                                 double* cfl_dev, double* work, size_t work_doubles,
                                 void* stream);

/* src2 alone with the curvature term (src2.f90:27-306 with curvature = 1) on
 * the interior of q in place; work: mx * my doubles, overwritten. */
int dclaw_src2_curvature_dev(double* q, const double* aux, int maux, int mx, int my, double dx,
                             double dy, const double* dt_dev, const dclaw_params* params_dev,
                             double* work, size_t work_doubles, void* stream);

/* Batched fused step over npatch patches (device array of dclaw_patch).
 * One launch; the work list of (patch, strip, segment) tiles is built on
 * the host by dclaw_worklist_size / dclaw_worklist_build into device memory.
 * params_dev, dt_dev, cfl_dev are indexed by dclaw_patch.member.  bc (host
 * array of 4 codes, or NULL to read stored ghosts) is applied on load for
 * every patch.  tile_flags_dev (NULL: fused step) is a workspace of ntiles
 * ints that selects the split step (src2 in a second pass over the flagged
 * tiles, as in dclaw_stepgrid_split_dev).  No reference counterpart: the
 * reference advances patches one stepgrid call at a time (OpenMP over
 * patches in GeoClaw advanc). */
int dclaw_worklist_size(const int* mx, const int* my, int npatch, int seg_rows);
int dclaw_worklist_build(const int* mx, const int* my, int npatch, int seg_rows,
                         int32_t* tiles_host /* 4*size ints */);
int dclaw_step_batched(const dclaw_patch* patches_dev, int npatch, const int32_t* tiles_dev,
                       int ntiles, int seg_rows, const int* bc, const double* dt_dev,
                       const dclaw_params* params_dev, double* cfl_dev, int* tile_flags_dev,
                       void* stream);

/* dclaw_step_batched with the workspace size given: with n_ws >= 2 * ntiles +
 * 4 ints in tile_flags_dev the tiles the previous call flagged are
 * dispatched first (keep the workspace between steps; any contents give a
 * valid order), else as dclaw_step_batched.  With n_ws >=
 * dclaw_batched_resident_size(ntiles, npatch) (a "resident" workspace) and
 * boundary conditions on load (every bc code nonzero), the tiles of a patch
 * whose neighbourhood stayed dry on that patch's last three inputs are
 * neither read nor written, as in the single-patch split step: for a loop
 * that swaps each patch's q_in and q_out every call and changes neither in
 * between, with the work list of dclaw_worklist_build (the library checks
 * each patch's swap, the segment height and the list's shape on the device
 * and starts over where they break; a caller that writes a state between
 * calls zeroes the workspace, or only its four ints from 2 * ntiles + 4,
 * which keeps the wet-first order).  Same result in every case.  No
 * reference counterpart (a GPU scheduling choice). */
int dclaw_batched_resident_size(int ntiles, int npatch);
int dclaw_step_batched_ws(const dclaw_patch* patches_dev, int npatch, const int32_t* tiles_dev,
                          int ntiles, int seg_rows, const int* bc, const double* dt_dev,
                          const dclaw_params* params_dev, double* cfl_dev, int* tile_flags_dev,
                          int n_ws, void* stream);

/* The tangent-linear step (as dclaw_stepgrid_tl_dev) over a batched work
 * list (as dclaw_step_batched, src2 fused): tangents_dev[p] gives patch p's
 * tangent arrays in its q_in (dq at t_n) and q_out (dq at t_n+dt) fields, the
 * other fields ignored; dparams_dev[member] is the member's direction.  No
 * reference counterpart. */
int dclaw_step_batched_tl(const dclaw_patch* patches_dev, const dclaw_patch* tangents_dev,
                          int npatch, const int32_t* tiles_dev, int ntiles, int seg_rows,
                          const int* bc, const double* dt_dev, const dclaw_params* params_dev,
                          const dclaw_params* dparams_dev, double* cfl_dev, void* stream);

/* Device dt rule for n members: dt[k] <- min(dt[k]*cfl_desired/cfl[k], dt_max)
 * when cfl[k] > 0 (else dt_max), then cfl[k] <- 0 (GeoClaw stepgrid dtnew
 * and tick.f:499 dt_max cap). */
int dclaw_dt_update(double* dt_dev, double* cfl_dev, int n, double cfl_desired,
                    double dt_max, void* stream);

/* dt rule with output times for n members (device arrays): the member has
 * just taken a step of dt[k] at time t[k]; then t[k] += dt[k], dt[k] <-
 * min(dt[k]*cfl_desired/cfl[k] (dt_max if cfl[k] = 0), dt_max), cfl[k] <- 0,
 * and if t[k] < tout[k] <= t[k] + 1.001*dt[k] the next step is shortened to
 * land on tout[k] (dt[k] = tout[k] - t[k], hit[k] = 1, else 0): AMRClaw
 * tick.f:135-144 (output times; tout = +inf for none) and :499 (dt_max). */
int dclaw_dt_update_out(double* dt_dev, double* cfl_dev, double* t_dev, const double* tout_dev,
                        int* hit_dev, int n, double cfl_desired, double dt_max, void* stream);

/* Gauges on device memory: for each gauge (xy_dev: x, y pairs), out_dev gets
 * 8 doubles: the 7 q components and eta = h + (aux(1) - bdif), bilinear in
 * the four cells whose centres surround the gauge when all four are wet
 * (h > dry_tolerance), else the containing cell's values (the convention of
 * gradient_particle_filter/flume_solver.py:91-99 in two dimensions; GeoClaw's
 * gauges_module, which the reference links, is not in its tree).  The four
 * cells must be interior: x in [xlower + dx/2, xlower + (mx - 1/2) dx), same
 * in y; other gauges get NaN. */
int dclaw_gauges_dev(const double* q, const double* aux, int maux, int mx, int my, double xlower,
                     double ylower, double dx, double dy, double dry_tolerance,
                     const double* xy_dev, int ngauges, double* out_dev, void* stream);

/* The same gauges with their tangents: dq is the tangent of q (as from
 * dclaw_stepgrid_tl_dev); dout_dev (8 doubles per gauge) receives the tangent
 * of each gauge value through the branch and weights the values take (bed
 * without tangent).  No reference counterpart (derivatives of gauge
 * records for gradient-based fitting). */
int dclaw_gauges_tl_dev(const double* q, const double* dq, const double* aux, int maux, int mx,
                        int my, double xlower, double ylower, double dx, double dy,
                        double dry_tolerance, const double* xy_dev, int ngauges, double* out_dev,
                        double* dout_dev, void* stream);

/* The adjoint (reverse-mode) fused step: the transpose of the tangent-linear
 * step's Jacobian (dclaw_stepgrid_tl_dev) at q_in, dt fixed.  Given lam_out
 * (the gradient of a loss with respect to q_{n+1}, layout of q, ghosts
 * included), writes lam_in = J_q^T lam_out (with bc set, the ghosts of q_in
 * are not read and get 0: their adjoint goes to the interior cells the
 * boundary map loads them from) and, for the np <= 4 parameter directions
 * dparams_dev[k], grad_dev[k] = lam_out . J_P dparams_dev[k] (device
 * doubles).  If q_out is not null the step's own result q_{n+1} is written
 * there, bit for bit as dclaw_stepgrid_bc_dev (the record pass).  work:
 * caller-owned device scratch of at least dclaw_adjoint_workspace_doubles(mx,
 * my) doubles, zero-filled before its first use; each call leaves it as the
 * next call with the same mx, my needs it (do not modify it in between).
 * q_out and lam_in must not alias q_in, lam_out or each other.  mx, my >= 2.
 * The work is confined to the 16 x 4 tiles within 5 cells of a wet cell (and
 * the border band); elsewhere the step is the dry map, whose adjoint passes
 * only bdif.  The reference has no adjoint; this is the "adjoint step"
 * call of the device API a gradient-based caller (parameter fitting, data
 * assimilation) binds, with rpn2/rpt2/flux2/src2 (rpn2_dclaw.f,
 * rpt2_dclaw.f, flux2fw.f, src2.f90) differentiated as the step computes
 * them. */
size_t dclaw_adjoint_workspace_doubles(int mx, int my);
int dclaw_adjoint_step_dev(const double* q_in, const double* aux, int maux, int mx, int my,
                           double dx, double dy, const int* bc, const double* dt_dev,
                           const dclaw_params* params_dev, const dclaw_params* dparams_dev, int np,
                           const double* lam_out, double* lam_in, double* q_out, double* grad_dev,
                           double* work, size_t work_doubles, void* stream);

/* The adjoint step of many patches of uneven size at once: every pass of
 * dclaw_adjoint_step_dev is one launch for the whole set, over the active
 * tiles of all patches (patches_dev: device array of npatch records,
 * patches_host: the same records in host memory, read for the launch sizes).
 * Patch p uses params_dev[member], dt_dev[member] and the np directions
 * dparams_dev[member * np .. member * np + np - 1]; each patch has its own
 * workspace (as for dclaw_adjoint_step_dev).  batch_work: device scratch of
 * dclaw_adjoint_batch_doubles(patches_host, npatch) doubles (no contents
 * kept between calls).  Results equal dclaw_adjoint_step_dev on each patch
 * bit for bit.  No reference counterpart (ensemble gradients). */
size_t dclaw_adjoint_batch_doubles(const dclaw_adj_patch* patches_host, int npatch);
int dclaw_adjoint_batched(const dclaw_adj_patch* patches_dev, const dclaw_adj_patch* patches_host,
                          int npatch, const int* bc, const double* dt_dev,
                          const dclaw_params* params_dev, const dclaw_params* dparams_dev, int np,
                          double* batch_work, size_t batch_doubles, void* stream);

/* Adjoint of dclaw_gauges_dev: lam_q_dev (layout of q) += the transpose of
 * the gauge map's derivative at q applied to lam_gauges_dev (8 doubles per
 * gauge, the adjoint of each gauge value), through the branch and weights
 * the values take.  Gauges outside the patch contribute nothing.  No
 * reference counterpart. */
int dclaw_gauges_adj_dev(const double* q, const double* aux, int maux, int mx, int my,
                         double xlower, double ylower, double dx, double dy, double dry_tolerance,
                         const double* xy_dev, int ngauges, const double* lam_gauges_dev,
                         double* lam_q_dev, void* stream);

/* AMR transfers between refinement levels, on device memory (the numerical
 * part of the AMR driver's coarse/fine coupling; coordinate_system 1).
 * Arrays: q(7, nx, ny), aux(maux, nx, ny) (aux component 1 = bed), indices
 * 0-based over the whole array including ghost cells.
 *
 * Fine -> coarse (replaces the averaging loop of update.f90:112-203): coarse
 * cells (ci0 + a, cj0 + b), 0 <= a < ni, 0 <= b < nj, of qc (mitc x mjtc) are
 * overwritten by the average of the rx x ry fine cells of qf (mitf x mjtf)
 * from (fi0 + a*rx, fj0 + b*ry): depth from the mean surface of the wet fine
 * cells over the coarse bed, capped by their mean depth; the other
 * components from the wet fine cells, scaled by the new depth over the total
 * fine depth.  The flux-register correction (upbnd) is not part of it. */
int dclaw_amr_average_dev(double* qc, const double* auxc, int mauxc, int mitc, int mjtc,
                          const double* qf, const double* auxf, int mauxf, int mitf, int mjtf,
                          int ci0, int cj0, int ni, int nj, int fi0, int fj0, int rx, int ry,
                          double dry_tolerance, void* stream);

/* Coarse -> fine (replaces the interpolation loop of filval.f90:190-381):
 * valc/auxc is the coarse patch enlarged by one coarse cell all round
 * (mic x mjc, already filled, boundary conditions applied); each interior
 * coarse cell (1..mic-2, 1..mjc-2) fills its rx x ry block of val
 * (mitot x mjtot with 2 ghost cells: block of coarse cell (i, j) starts at
 * ((i-1)*rx + 2, (j-1)*ry + 2)).  needs_set (mitot x mjtot bytes, nonzero =
 * the reference's NEEDS_TO_BE_SET) selects the fine cells whose depth is
 * interpolated (surface by minmod-limited slopes, sea_level for dry coarse
 * cells, minus the fine bed aux(1) of auxf, clipped at 0); the other
 * components are then set over the whole block, as the reference does:
 * limited-linear if no fine value / depth leaves the range of the coarse cell
 * and its wet neighbours, else spread in proportion to the fine depth.
 * use_force_dry and variable_eta_init are not supported (their defaults). */
int dclaw_amr_interp_dev(const double* valc, const double* auxc, int mauxc, int mic, int mjc,
                         double* val, const double* auxf, int mauxf, int mitot, int mjtot,
                         const unsigned char* needs_set, int rx, int ry, double dry_tolerance,
                         double sea_level, double rho_f, double grav, void* stream);

/* ---- AMR ghost fill and regrid on device.  The D-Claw parts are
 * filpatch.f90's interpolation from the enlarged coarse rectangle
 * (dclaw_amr_filpatch_dev), filval.f90's interpolation loop
 * (dclaw_amr_interp_dev) and setaux.f90's per-patch fields
 * (dclaw_amr_setaux_dev); the AMRClaw steps around them (not in the
 * reference: AMRClaw 5.x intfil/icall, trimbd, bc2amr) are restated as
 * dclaw_amr_copy_dev, dclaw_amr_unset_box_dev and dclaw_amr_region_bc_dev.
 * Non-periodic domains, coordinate_system 1, topography fixed in time,
 * force_dry and variable_eta_init off. ---- */

/* A region of one refinement level being filled (filpatch.f90's valbig,
 * aux and flaguse): cells ilo..ihi x jlo..jhi in the level's index space
 * (cell i covers [xlower + i dx, xlower + (i+1) dx)), stored from column r0,
 * row c0 of q(7, mx, my) and aux(maux, mx, my) (device, component fastest);
 * flag: (ihi-ilo+1) x (jhi-jlo+1) device bytes, column fastest, nonzero =
 * set.  Passed by pointer to host memory. */
typedef struct {
    double* q;
    double* aux;
    unsigned char* flag;
    int mx, my, maux;
    int r0, c0;
    int ilo, ihi, jlo, jhi;
} dclaw_amr_region;

/* Start a fill (AMRClaw intfil's marking): flag = 1 on the region's cells
 * outside the level's domain 0..iregsz-1 x 0..jregsz-1 (bc2amr sets them),
 * 0 elsewhere. */
int dclaw_amr_region_init_dev(const dclaw_amr_region* r, int iregsz, int jregsz, void* stream);

/* Copy from one patch of the same level (AMRClaw intfil/icall): the cells of
 * the source's interior (src_ilo.., src_jlo.., src_mx x src_my cells; arrays
 * with 2 ghost cells, interior cell (src_ilo + a, src_jlo + b) at column
 * a + 2, row b + 2) that lie in the region get the source's values and flag
 * 1.  With src_q_old (the source's values at its previous time), the value
 * is q (1 - w_old) + q_old w_old, the linear interpolation in time (w_old =
 * (t_new - t) / (t_new - t_old)); w_old = 0 copies q.  copy_aux != 0 also
 * copies the aux components (min(maux, src_maux) of them).  src_q == r->q
 * marks the cells only (a patch filling its own ghosts). */
int dclaw_amr_copy_dev(const dclaw_amr_region* r, const double* src_q, const double* src_q_old,
                       double w_old, const double* src_aux, int src_maux, int src_ilo,
                       int src_jlo, int src_mx, int src_my, int copy_aux, void* stream);

/* trimbd: the bounding box (isl, isr, jsb, jst) in level index space of the
 * region's unset cells into box_dev[0..3] (device ints); isl > isr when every
 * cell is set. */
int dclaw_amr_unset_box_dev(const dclaw_amr_region* r, int* box_dev, void* stream);

/* The coarse rectangle (iplo, iphi, jplo, jphi) filpatch.f90:182-187 builds
 * around the unset cells isl..isr x jsb..jst at refinement rx, ry (host). */
void dclaw_amr_coarse_box(int isl, int isr, int jsb, int jst, int rx, int ry, int* box);

/* filpatch.f90:256-520 (replaces filrecur's interpolation): the region's
 * unset cells (flag 0) from the filled coarse rectangle valc/auxc
 * (iplo..iphi x jplo..jphi, component fastest, no ghost cells; its bed is
 * auxc component 1): depth from the dry-limited surface by minmod slopes
 * minus the fine bed aux(1), the other components limited-linear within the
 * coarse velocity bounds, else conservatively spread.  Every unset cell
 * must lie in the interior of the coarse rectangle (the rectangle
 * dclaw_amr_coarse_box gives for the unset box does).  dx_coarse, dy_coarse:
 * the coarser level's cell size (hxposs(level-1)). */
int dclaw_amr_filpatch_dev(const dclaw_amr_region* r, const double* valc, const double* auxc,
                           int mauxc, int iplo, int iphi, int jplo, int jphi, int rx, int ry,
                           double xlower, double ylower, double dx_fine, double dy_fine,
                           double dx_coarse, double dy_coarse, double sea_level,
                           double dry_tolerance, double rho_f, double grav, void* stream);

/* bc2amr on a region: its cells outside the domain take q (and all aux
 * components) from the cell the boundary rule names -- extrap: the nearest
 * domain cell; wall (bc 3): the mirror cell with the normal momentum
 * negated -- which must lie in the region.  bc: xlo, xhi, ylo, yhi. */
int dclaw_amr_region_bc_dev(const dclaw_amr_region* r, int iregsz, int jregsz, const int* bc,
                            void* stream);

/* setaux.f90:59-258 for a new patch (interior ilo.., jlo.., mx x my cells, 2
 * ghost cells; aux(maux, mx+4, my+4), maux >= 3): aux(1) = bed where
 * bed_set is 0 and the cell is in the domain (bed: the cell integral of the
 * topography / cell area, cellgridintegrate's result, computed by the
 * caller; bed_set marks cells whose aux(1) was copied from an old patch),
 * then cells outside the domain copy aux(1) of the nearest domain cell
 * (:154-184), then aux(2..maux) = 0, aux(2) = phi (i_phi), aux(3) = theta
 * (i_theta).  bed_set and bed: (mx+4) x (my+4), column fastest.  Auxinit
 * files (aux(i_ent) and others from files) are the caller's to add. */
int dclaw_amr_setaux_dev(double* aux, int maux, int mx, int my, int ilo, int jlo,
                         const unsigned char* bed_set, const double* bed, int iregsz,
                         int jregsz, double phi, double theta, void* stream);

/* ---- Per-stage device entry points (GeoClaw stepgrid's stages one at a
 * time, on device memory; the fused dclaw_stepgrid_dev equals b4step2, step2,
 * the conservative update and src2 in sequence). ---- */

/* check4nans (AMRClaw; called at b4step2.f90:49 and src2.f90:73): the
 * smallest flat index (component fastest, q(m,i,j) -> m-1 + 7*((i+1) +
 * (mx+4)*(j+1))) of a NaN anywhere in q, ghosts included, min-reduced into
 * *first_nan_dev; initialise it to 0x7fffffff (none). */
int dclaw_check4nans_dev(const double* q, int mx, int my, int* first_nan_dev, void* stream);

/* b4step2 (b4step2.f90:7-113) in place: qfix and bdif >= 0 on every cell
 * (:56-64), the speed limit (:74-100), then calc_taudir
 * (digclaw_module.f90:500-719) into aux(4..6) on cells 0..mx+1 x 0..my+1
 * (maux >= 6).  The check4nans of :49 runs first when first_nan_dev is not
 * NULL. */
int dclaw_b4step2_dev(double* q, double* aux, int maux, int mx, int my, double dx, double dy,
                      const dclaw_params* params_dev, int* first_nan_dev, void* stream);

/* step2 (GeoClaw step2.f90 with flux2fw.f:8-286, rpn2_dclaw.f:24, rpt2_dclaw.f:16)
 * on the b4step2 output q and aux (maux >= 6; aux(4..6) read as fsphi and
 * taudir): fm, fp, gm, gp (same shape as q; every entry overwritten, in the
 * reference's summation order) and the max into *cfl_dev (cflgrid). */
int dclaw_step2_dev(const double* q, const double* aux, int maux, int mx, int my, double dx,
                    double dy, const double* dt_dev, const dclaw_params* params_dev, double* fm,
                    double* fp, double* gm, double* gp, double* cfl_dev, void* stream);

/* src2 (src2.f90:27-306) in place on the interior cells; check4nans (:73)
 * first when first_nan_dev is not NULL.  aux is read only with entrainment
 * on, as aux(7) = h_e (src2.f90:253-263; maux >= 7 then). */
int dclaw_src2_dev(double* q, const double* aux, int maux, int mx, int my, const double* dt_dev,
                   const dclaw_params* params_dev, int* first_nan_dev, void* stream);

/* Host-memory drop-ins with the Fortran calling convention (all arguments by
 * reference, q/aux in the Clawpack layout), for AMRClaw executables that
 * swap these objects in through Makefile SOURCES/EXCLUDE_SOURCES
 * (Makefile_template:45-63).  Parameters come from dclaw_set_params()
 * called once after set_dig.  stepgrid_dclaw_ performs b4step2 + step2 +
 * update + src2 in place on q and returns cflgrid; aux(4..6) are written as
 * b4step2/calc_taudir would (digclaw_module.f90:641-713) when maux >= 6.
 * The drop-ins are re-entrant: each calling thread has its own device
 * buffers and HIP stream, so AMRClaw's advanc can step different patches
 * from its OpenMP threads at once; dclaw_set_params waits for the calls in
 * progress. */
int dclaw_set_params(const dclaw_params* params_host);
void stepgrid_dclaw_(double* q, double* aux, const int* mbc, const int* mx, const int* my,
                     const int* maux, const double* dx, const double* dy, const double* dt,
                     double* cflgrid);
/* stepgrid_dclaw_ that also returns the step's fm, fp, gm, gp (host arrays
 * of q's shape, every entry written, in step2's summation order, equal to
 * what the step applied): what AMRClaw's advanc hands to its conservation
 * fix-up (fluxsv on a coarse patch, qad and fluxad on a fine one) after
 * GeoClaw's stepgrid, whose fm..gp arguments this fills (one extra b4step2
 * and step2 pass on the device). */
void stepgrid_fluxes_dclaw_(double* q, double* aux, const int* mbc, const int* mx, const int* my,
                            const int* maux, const double* dx, const double* dy,
                            const double* dt, double* cflgrid, double* fm, double* fp,
                            double* gm, double* gp);

/* Per-stage Fortran drop-ins with the reference's own argument lists, for
 * executables that keep GeoClaw's stepgrid and exclude the D-Claw stage
 * sources instead (Makefile.dclaw:155-173).  Host memory; each call copies
 * the patch to and from the device.  A NaN in q stops the run with
 * check4nans' message, like the reference.
 *   b4step2_: b4step2.f90:7 (setaux is not re-run: aux must be final)
 *   step2_:   GeoClaw step2 (Clawpack 5 form); rpn2/rpt2 are ignored (the
 *             D-Claw solvers are built in); maxm is not needed
 *   src2_:    src2.f90:27 */
void b4step2_(const int* mbc, const int* mx, const int* my, const int* meqn, double* q,
              const double* xlower, const double* ylower, const double* dx, const double* dy,
              const double* t, const double* dt, const int* maux, double* aux,
              const int* actualstep);
void step2_(const int* maxm, const int* meqn, const int* maux, const int* mbc, const int* mx,
            const int* my, const double* qold, const double* aux, const double* dx,
            const double* dy, const double* dt, double* cflgrid, double* fm, double* fp,
            double* gm, double* gp, void* rpn2, void* rpt2);
void src2_(const int* meqn, const int* mbc, const int* mx, const int* my, const double* xlower,
           const double* ylower, const double* dx, const double* dy, double* q, const int* maux,
           const double* aux, const double* t, const double* dt);

#ifdef __cplusplus
}
#endif
#endif
