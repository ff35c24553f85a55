/*
 * compactles_hip.h -- C ABI of the MI355X (gfx950) backend for the CompactLES
 * per-step hot path.
 *
 * Plain C types only (pointers, sizes, doubles); no HIP or torch types in any
 * signature.  Streams are passed as `void*` (a hipStream_t, or NULL for the
 * context's own stream).  Every entry point returns an int status (CLES_OK = 0)
 * and leaves a message for cles_last_error() on failure.
 *
 * Each declaration names the reference interface it replaces
 * (stillyslalom/CompactLES.jl, path:line).
 */
#ifndef COMPACTLES_HIP_H
#define COMPACTLES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLES_ABI_VERSION 5

/* Status codes.  The solver-failure codes mirror SolverFailure reasons
 * (src/stepcontrol.jl:138-164): nonfinite, negative_density, planck, dt_min,
 * dt_collapse. */
enum {
    CLES_OK = 0,
    CLES_ERR_INVALID = 1,        /* bad argument (error(...) in the reference) */
    CLES_ERR_TOO_SMALL = 2,      /* local extent too small for the scheme */
    CLES_ERR_HIP = 3,            /* HIP runtime error */
    CLES_ERR_NONFINITE = 4,
    CLES_ERR_NEGATIVE_DENSITY = 5,
    CLES_ERR_PLANCK = 6,
    CLES_ERR_DT_MIN = 7,
    CLES_ERR_DT_COLLAPSE = 8,
    CLES_ERR_NOMEM = 9,
    CLES_ERR_UNSUPPORTED = 10
};

/* Element types of operator inputs/outputs. */
enum { CLES_F64 = 0, CLES_F32 = 1 };

/* Boundary-condition codes (src/boundary.jl, src/problem.jl:192, src/nscbc.jl). */
enum {
    CLES_BC_PERIODIC = 0,    /* PeriodicBC            boundary.jl:24 */
    CLES_BC_SLIPWALL = 1,    /* SlipWallBC            boundary.jl:32,232-243 */
    CLES_BC_DIRICHLET = 2,   /* DirichletBC, constant state  problem.jl:192-207 */
    CLES_BC_NSCBC_OUT = 3,   /* NSCBCOutflowBC        nscbc.jl:45-154 */
    CLES_BC_NSCBC_IN = 4,    /* NSCBCInflowBC, constant targets  nscbc.jl:204-297 */
    CLES_BC_AXIS = 5,        /* AxisBC, low end of r     boundary.jl:55-66, rhs.jl:155-163 */
    CLES_BC_ORIGIN = 6,      /* OriginBC, low end of r   boundary.jl:68-80, rhs.jl:164-173 */
    CLES_BC_POLE = 7,        /* PoleBC, both ends of theta  boundary.jl:82-92, rhs.jl:174-180 */
    CLES_BC_NOSLIP = 8,      /* NoSlipWallBC (Twall)  boundary.jl:34-43,245-262 */
    CLES_BC_EXTRAP = 9       /* ExtrapolationBC       boundary.jl:45-52,264-276 */
};

#define CLES_MAXCL 8         /* closure rows per edge */
#define CLES_MAXRHS 16       /* closure RHS stencil length */
#define CLES_MAXM 8          /* interior RHS half-width */
#define CLES_MAXSP 8         /* species */

typedef struct cles_ctx cles_ctx;
typedef struct cles_plan cles_plan;
typedef struct cles_solver cles_solver;

/* One compact scheme: the fields of CompactScheme (src/kernels.jl:43-50) and
 * BandedCompactScheme (src/kernels_banded.jl:31-39).  q = LHS half-bandwidth
 * (1: tridiagonal, 2: pentadiagonal); lhs[s-1] is the LHS coefficient at
 * distance s; closure_lhs[j] holds the 2q+1 centred LHS entries of low-edge
 * closure row j (for q = 1: sub, diag, super); closure_rhs[j][0..len-1] are
 * weights on f[1..len] counted inward from the edge.  High-edge rows are
 * mirrored as in the reference (operators.jl:77-83,87-92). */
typedef struct {
    int q;
    double lhs[4];
    double a0;
    int M;
    double coeffs[CLES_MAXM];
    int symmetric;
    int nclosure;
    double closure_lhs[CLES_MAXCL][9];
    int closure_rhs_len[CLES_MAXCL];
    double closure_rhs[CLES_MAXCL][CLES_MAXRHS];
} cles_scheme_desc;

/* ----------------------------------------------------------------------- */
/* Context, errors, device memory                                           */
/* ----------------------------------------------------------------------- */

/* Selects the HIP device and creates the context's stream.  No reference
 * counterpart (the reference is CPU-only, src/CompactLES.jl:1-13). */
int cles_ctx_create(int device, cles_ctx **ctx);
int cles_ctx_destroy(cles_ctx *ctx);
/* Thread-local message of the last failing call on this thread. */
const char *cles_last_error(void);
int cles_abi_version(void);
/* The context's stream as a void* hipStream_t. */
void *cles_ctx_stream(cles_ctx *ctx);

/* Device buffers for hosts without a GPU array package (stand-ins for the
 * reference's `field(decomp)` / `allocate_state`, decomposition.jl:129-134). */
int cles_malloc(cles_ctx *ctx, size_t bytes, void **dptr);
int cles_free(cles_ctx *ctx, void *dptr);
int cles_memcpy_h2d(cles_ctx *ctx, void *dst, const void *src, size_t bytes);
int cles_memcpy_d2h(cles_ctx *ctx, void *dst, const void *src, size_t bytes);
int cles_memset(cles_ctx *ctx, void *dptr, int value, size_t bytes);
int cles_synchronize(cles_ctx *ctx);
/* Device-to-device copy done by a streaming kernel (it also serves as the
 * known-size calibration kernel for hardware byte counters). */
int cles_memcpy_d2d(cles_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ----------------------------------------------------------------------- */
/* Operator plugin: plan_direction / apply_along!                           */
/* ----------------------------------------------------------------------- */

/* plan_direction(decomp, scheme, dim, h; lo_fold, hi_fold) -> DirPlan |
 * BandPlan (src/operators.jl:44-101, src/operators_banded.jl:29-125), one
 * rank: whole-domain lines.  dim is 1-based (1 = x).  n_local are the
 * interior extents; `periodic` selects the cyclic system (otherwise both
 * edges are closed and the closure rows are used).  h is the grid spacing
 * (derivative RHS prescaled by 1/h).  dtype is CLES_F64 or CLES_F32.
 * lo_fold / hi_fold: 0 for `nothing`, or the reference's keyword value
 * sigma_g = +1 / -1, the output's parity across a half-offset mirror at that
 * end of a closed line (operators.jl:55-70, operators_banded.jl:43-73): the
 * LHS coupling to the ghost unknowns moves onto the mirrored interior
 * columns with sign sigma_g.  The reference's RHS reads halos that the caller
 * mirror-fills with the input's sign sigma_f (fold_fill!, folds.jl:83);
 * here halos are never read, and the edge rows read the mirrored interior
 * with sigma_f = sigma_g for a symmetric scheme (filter) and -sigma_g for a
 * derivative, the pairing fold_apply! uses (folds.jl:271-274).  On a
 * periodic line the fold arguments are ignored, as the reference's edge test
 * is false there (decomposition.jl:137-140).  Fails with CLES_ERR_TOO_SMALL on
 * the reference's extent check (operators.jl:50-53). */
int cles_plan_create(cles_ctx *ctx, const cles_scheme_desc *scheme, int dim,
                     const int64_t n_local[3], int periodic, double h, int dtype,
                     int lo_fold, int hi_fold, cles_plan **plan);
int cles_plan_destroy(cles_plan *plan);

/* apply_along!(out, plan, f, decomp) (src/operators.jl:199-215): applies the
 * compact operator along the plan's dimension to every line of f and writes
 * the interior of out.  f and out are device pointers to column-major arrays
 * of extents padded[3] with interior offset halo[3] (Julia's padded `field`
 * layout; pass halo = {0,0,0} for unpadded arrays).  Periodic lines wrap in
 * the interior (equal to reading current halos); halos are never read or
 * written.  out may equal f (in-place filtering).  Asynchronous on `stream`
 * (NULL: the context stream). */
int cles_apply_along(cles_plan *plan, void *out, const void *f,
                     const int64_t padded[3], const int halo[3], void *stream);

/* Times `iters` back-to-back apply_along launches with HIP events on the
 * launch stream and returns the mean milliseconds per launch. */
int cles_apply_along_timed(cles_plan *plan, void *out, const void *f,
                           const int64_t padded[3], const int halo[3],
                           int iters, double *ms_per_launch);

/* ----------------------------------------------------------------------- */
/* Solver: setup / compute_rhs! / step! / max_rate / filter_state! / run!   */
/* ----------------------------------------------------------------------- */

/* One NASA-9 fit interval (Nasa9Interval, src/physics.jl:280-300): cp/R =
 * a1/T^2 + a2/T + a3 + a4 T + a5 T^2 + a6 T^3 + a7 T^4 on [Tmin, Tmax], and
 * h/R = -a1/T + a2 ln T + a3 T + a4 T^2/2 + a5 T^3/3 + a6 T^4/4 + a7 T^5/5 + b1. */
#define CLES_NASA9_MAXINT 3
typedef struct {
    double Tmin, Tmax;
    double a[7];
    double b1;
} cles_nasa9_interval;
/* Nasa9Species (physics.jl:340-375): ordered, contiguous intervals; a
 * temperature above interval i's Tmax uses the next one, the last one
 * beyond its range. */
typedef struct {
    int n_intervals;
    cles_nasa9_interval iv[CLES_NASA9_MAXINT];
} cles_nasa9_species;

/* NSCBC face parameters (src/nscbc.jl:45-54 outflow, :204-216 inflow). */
typedef struct {
    double pinf, sigma, Lref, beta_t;      /* outflow; Lref <= 0: domain length */
    double u[3], T, Y[CLES_MAXSP];         /* inflow targets */
    double eta_u, eta_T, eta_t, eta_Y;
} cles_nscbc_desc;

/* Fields of Problem + Numerics (src/problem.jl:243-306) that reach the hot
 * path, with the EOS as an IdealMixture species table (physics.jl:104-126),
 * Transport (physics.jl:589-593), ArtParams (artificial.jl:50-57) and
 * StepControl's floors and predictor (stepcontrol.jl:116-136). */
typedef struct {
    int64_t n_global[3];
    double L_domain[3];
    double origin[3];
    int n_species;
    double R[CLES_MAXSP], gamma[CLES_MAXSP];
    double mu0, Pr, Sc;
    int art_enabled;
    double C_mu, C_beta, C_kappa, C_D;
    int beta_sensor;                       /* 0 :strain, 1 :gated_strain, 2 :dilatation */
    cles_scheme_desc deriv, filt;
    double cfl;
    int filter_interval;
    int bc[3][2];                          /* CLES_BC_* per (dim, side) */
    double bc_state[3][2][CLES_MAXSP + 4]; /* Dirichlet conserved state */
    cles_nscbc_desc nscbc[3][2];
    double predict, max_growth, dt_min, dt_min_ratio;
    int precision;                         /* CLES_F64, or CLES_F32 = mixed form */
    /* StepControl recovery (stepcontrol.jl:100-114, timestep.jl:288-344) and
     * soft landing (timestep.jl:367-383) */
    int retries;                           /* rollbacks to the last savepoint (0: off) */
    double cfl_backoff;                    /* CFL multiplier per rollback */
    int savepoint_interval;                /* steps between savepoints */
    int landing_steps;                     /* >= 1; 0 selects the default 2 */
    /* EOS: CLES_EOS_IDEAL IdealMixture (R, gamma; physics.jl:104-126),
     * CLES_EOS_NASA9 Nasa9Mixture (R, nasa9, T_guess; physics.jl:402-571) or
     * CLES_EOS_STIFFENED StiffenedGas (one species: gamma[0], R[0] = (gamma -
     * 1) cv, p_inf; physics.jl:160-250) */
    int eos_kind;
    double T_guess;
    cles_nasa9_species nasa9[CLES_MAXSP];
    /* Metric (metric.jl:20-45): coordinates (x, y, z), (r, theta, z) or
     * (r, theta, phi) with theta the polar angle; velocity components are
     * physical components in the local orthonormal basis. */
    int metric;                            /* CLES_METRIC_* */
    /* NoSlipWallBC wall temperature per face: > 0 isothermal (the EOS sets the
     * wall's energy, boundary.jl:203-221), otherwise adiabatic */
    double Twall[3][2];
    /* SwitchableBC (boundary.jl:94-145): a face with switchable != 0 runs
     * bc[d][s] until cles_solver_switch_bc selects bc_after[d][s]; both share
     * the face's parameters (bc_state, nscbc, Twall) */
    int switchable[3][2];
    int bc_after[3][2];
    /* Problem.sources (sources.jl): ConstantBodyForce accelerations, applied
     * in order after the boundary corrections (rhs.jl:692) */
    int n_body_forces;                     /* 0..CLES_MAX_FORCES */
    double body_force[4][3];
    double p_inf;                          /* StiffenedGas cohesive pressure */
} cles_solver_desc;

#define CLES_MAX_FORCES 4

enum { CLES_EOS_IDEAL = 0, CLES_EOS_NASA9 = 1, CLES_EOS_STIFFENED = 2 };
enum { CLES_METRIC_CARTESIAN = 0, CLES_METRIC_CYLINDRICAL = 1, CLES_METRIC_SPHERICAL = 2 };

/* save_checkpoint / load_checkpoint! (src/io.jl:6-75): `prefix`.rNNNN.ckpt
 * with the reference's header (magic, n_cons, n_global, n_local, rank
 * coordinates, t, step; Int64 / Float64 little-endian) and the interior of Q
 * (Float64, x fastest, components last).  One rank: n_local = n_global and
 * coordinates (0, 0, 0).  Loading checks the header as the reference does
 * ("not a CompactLES checkpoint", "conserved layout mismatch", "global grid
 * mismatch", "decomposition mismatch", "rank layout mismatch") and restores
 * Q, t and step. */
int cles_solver_save_checkpoint(cles_solver *solver, const char *prefix, int rank);
int cles_solver_load_checkpoint(cles_solver *solver, const char *prefix, int rank);

/* Device bytes cles_solver_create would allocate for `desc` (an upper bound:
 * every field of the allocation table plus an allowance for the line-plan
 * tables); no device is touched.  The admission measure of cles_server. */
int cles_solver_footprint(const cles_solver_desc *desc, size_t *bytes);
/* setup(prob, num) -> Solver (src/problem.jl:321-345, src/rhs.jl:122-311),
 * one rank.  Allocates the device-resident state and every scratch field. */
int cles_solver_create(cles_ctx *ctx, const cles_solver_desc *desc, cles_solver **solver);
/* setup with Stretch mappings (src/metric.jl:45-58 with CartesianMetric):
 * dxdxi[d] (NULL, or dxdxi itself NULL, for a uniform dimension) holds dx/dxi
 * at the n_global[d] nodes xi = i / (n_global[d] - 1) of a stretched,
 * non-periodic dimension.  That dimension's spacing is then 1 / (n - 1)
 * (rhs.jl:191-201); gradients are scaled by inv_h (rhs.jl:477-483), the flux
 * divergence takes the general path dQ -= inv_J D_d(A_d F_d) (rhs.jl:665-684),
 * max_rate uses inv_h[d] / h[d] (timestep.jl:151), the volume integrals
 * divide by inv_J (diagnostics.jl:73-75) and the NSCBC derivatives take
 * inv_h (nscbc.jl:108-130, :257). */
int cles_solver_create_ex(cles_ctx *ctx, const cles_solver_desc *desc, const double *const dxdxi[3],
                          cles_solver **solver);
/* setup with a metric and Stretch mappings (src/metric.jl:80-269, rhs.jl:
 * 122-311): as cles_solver_create_ex, and x[d] (NULL, or x itself NULL, for
 * a dimension with uniform nodes origin[d] + i h[d]) holds the physical
 * coordinate at the nodes of a stretched dimension whose coordinate the
 * metric reads: r (dimension 1) under CLES_METRIC_CYLINDRICAL and
 * CLES_METRIC_SPHERICAL, theta (dimension 2) under CLES_METRIC_SPHERICAL.
 * A curvilinear metric adds the curvature corrections of the velocity
 * gradients (metric.jl:181-219), the momentum sources (metric.jl:228-269),
 * the discrete-GCL cot(theta)/r of a resolved theta (metric.jl:150-168) and
 * the curvature rate of collapsed angular dimensions (timestep.jl:186-197).
 * Coordinate singularities fold (rhs.jl:143-205, folds.jl): CLES_BC_AXIS
 * at the low end of r (cylindrical; theta collapsed, or resolved over 2 pi
 * with an even count), CLES_BC_ORIGIN at the low end of r (spherical; theta
 * symmetric about pi/2, phi collapsed or even over 2 pi), CLES_BC_POLE at
 * both ends of theta (spherical, theta over (0, pi)).  Folded dimensions are
 * unstretched, from origin 0, on half-offset grids (h = L / (n - 1/2) with
 * one folded end, L / n with two), with parity-folded operators and, where
 * the antipode is another line, the even / odd pairing.  Otherwise every
 * node needs r > 0 and, for the spherical metric, 0 < theta < pi
 * (CLES_ERR_INVALID).  NSCBC
 * faces must have a normal scale factor of one: cylindrical r and z faces,
 * spherical r faces (nscbc.jl:309-316; CLES_ERR_INVALID otherwise). */
int cles_solver_create_geom(cles_ctx *ctx, const cles_solver_desc *desc, const double *const x[3],
                            const double *const dxdxi[3], cles_solver **solver);
int cles_solver_destroy(cles_solver *solver);
/* switch!(bc) (boundary.jl:131-137) on face (dim 0..2, side 0 low / 1 high)
 * of a SwitchableBC: later steps enforce and correct with bc_after.  A second
 * call has no further effect; CLES_ERR_INVALID on a face that is not
 * switchable.  cles_solver_bc_switched reads the state (switched(bc)). */
int cles_solver_switch_bc(cles_solver *solver, int dim, int side);
int cles_solver_bc_switched(cles_solver *solver, int dim, int side, int *switched);

/* Conserved state transfer in Julia's padded layout (Q[..., c], column-major,
 * extents padded[3], interior offset halo[3], n_cons components). */
int cles_solver_upload(cles_solver *solver, const double *Q_host,
                       const int64_t padded[3], const int halo[3]);
int cles_solver_download(cles_solver *solver, double *Q_host,
                         const int64_t padded[3], const int halo[3]);
/* Download a named internal field's interior (as double) into an unpadded
 * host array: "rho","u","v","w","p","T","c","cp","mu_art","beta_art",
 * "kappa_art","Q[c]","dQ[c]","du[c]","grad_u[dj]","grad_T[d]","Y[k]","D_art[k]"
 * (1-based).  For a one-species ideal gas without NSCBC faces, "cp" and "Y[1]"
 * are not stored on the device; they are formed on the host from "rho" by the
 * operations the primitives use (same values). */
int cles_solver_get_field(cles_solver *solver, const char *name, double *host);

int cles_solver_get_time(cles_solver *solver, double *t, int64_t *step);
int cles_solver_set_time(cles_solver *solver, double t, int64_t step);

/* apply_bcs!(solver, Q) (src/boundary.jl:280-286) at the current stage time. */
int cles_apply_bcs(cles_solver *solver);
/* compute_rhs!(solver, Q, dQ, primitives_current) (src/rhs.jl:624-694). */
int cles_compute_rhs(cles_solver *solver, int primitives_current);
/* max_rate(solver, Q) -> (rate, rho_min) (src/timestep.jl:125-175). */
int cles_max_rate(cles_solver *solver, double *rate, double *rho_min);
/* step!(solver, Q, dQ, du, dt, prepared) (src/timestep.jl:72-99). */
int cles_step(cles_solver *solver, double dt, int prepared);
/* filter_state!(solver, Q) (src/timestep.jl:414-426). */
int cles_filter_state(cles_solver *solver);
/* run!(solver, Q; tfinal, nmax) (src/timestep.jl:277-400) with a fresh
 * Workspace and no callbacks; the descriptor's StepControl recovery applies.
 * Returns CLES_OK or a solver-failure code (state left at the failing step).
 * dts (may be NULL) receives up to ndts step sizes, in the order the steps
 * were taken (a rollback's abandoned steps included). */
int cles_run(cles_solver *solver, double tfinal, int64_t nmax, double *dts, int64_t ndts);

/* Step-boundary hooks of run! (src/callbacks.jl), called on the host thread
 * that runs cles_run_ctl, between completed steps (the solver's device state
 * is then consistent and other cles_* calls on it are allowed):
 *  next_time(user, t, step): the next scheduled instant (callback_next_time,
 *    callbacks.jl:215-222), +INFINITY for none; run! clips the step to land on
 *    it softly, in ceil(gap/dt) equal steps when gap < landing_steps * dt
 *    (timestep.jl:357-383);
 *  after_step(user, t, step, dt): after each completed step and its filter
 *    (run_callbacks!, callbacks.jl:233-250); nonzero stops run!;
 *  rewind(user, t, step): after a StepControl rollback to (t, step)
 *    (rewind_callbacks!, callbacks.jl:224-231).
 * Any of the three may be NULL. */
typedef struct {
    void *user;
    double (*next_time)(void *user, double t, int64_t step);
    int (*after_step)(void *user, double t, int64_t step, double dt);
    void (*rewind)(void *user, double t, int64_t step);
} cles_run_hooks;

/* run!(solver, Q; tfinal, nmax, callback, control) (src/timestep.jl:277-400):
 * cles_run with hooks (NULL: none).  With retries > 0 a failing check_step
 * restores the last savepoint (Q, t, step; an extra FP64 state on the device,
 * taken every savepoint_interval steps on a state that has just passed its
 * check, never at or before a step that failed), multiplies the solver's CFL
 * by cfl_backoff (it stays lowered), clears the rate history and dt_seen,
 * calls rewind, and continues; with no retries left the failure is returned.
 * Hooks or retries run every step from the host (the device run modes are
 * used only without them). *retries_used (may be NULL) receives the number of
 * rollbacks. */
int cles_run_ctl(cles_solver *solver, double tfinal, int64_t nmax, const cles_run_hooks *hooks, double *dts,
                 int64_t ndts, int *retries_used);
/* The CFL in effect (Numerics' cfl, lowered by each rollback), and setting it. */
int cles_solver_get_cfl(cles_solver *solver, double *cfl);
int cles_solver_set_cfl(cles_solver *solver, double cfl);
/* Kinetic energy ∫½|ρu|²/ρ dV (examples/taylor_green.jl:51-60) and enstrophy
 * ∫½|ω|² dV with ω from the compact grad_u of the current state
 * (src/io.jl:332-340), both reduced in FP64 in a fixed order
 * (src/diagnostics.jl:61-80 volume_integral). */
int cles_diagnostics(cles_solver *solver, double *ke, double *enstrophy);

/* Per-phase timing of the solver's launches with HIP events on its stream
 * (phases modelled on bench/phases.jl:119-193; adds a few microseconds per
 * launch, so leave it off for throughput runs).  enable = 1 starts a fresh
 * accumulation, 0 stops.  cles_solver_profile_read returns phase `index`
 * (name, total milliseconds, launches, algorithmic bytes moved) and
 * CLES_ERR_INVALID past the last phase. */
int cles_solver_profile(cles_solver *solver, int enable);
int cles_solver_profile_read(cles_solver *solver, int index, char *name, int namelen,
                             double *ms_total, int64_t *launches, double *bytes_total);

/* How run! issues its steps.  CLES_RUN_HOST_DT: every kernel launched from
 * the host and the dt decision of predicted_dt / check_step / the tfinal clip
 * taken on the host each step (one host round trip per step).
 * CLES_RUN_DEVICE_DT: the same kernels, with that decision taken on the
 * device; the host reads it back every few steps.  CLES_RUN_STEP_GRAPH: as
 * CLES_RUN_DEVICE_DT, each step issued as one replay of a captured graph of
 * the whole iteration.  The results are identical in every mode; the default
 * is CLES_RUN_HOST_DT, which profiling (cles_solver_profile) always uses. */
enum { CLES_RUN_HOST_DT = 0, CLES_RUN_DEVICE_DT = 1, CLES_RUN_STEP_GRAPH = 2 };
int cles_solver_set_run_mode(cles_solver *solver, int mode);

/* Device-resident state geometry: interior extents and per-field element
 * count (the solver stores unpadded column-major fields). */
int cles_solver_shape(cles_solver *solver, int64_t n[3], int *n_cons);
/* Device bytes held by a live solver (fields, reduction buffers, plans). */
int cles_solver_memory(cles_solver *solver, size_t *bytes);

/* ----------------------------------------------------------------------- */
/* Batches of independent one-dimensional instances                        */
/* ----------------------------------------------------------------------- */
/* The reference runs one problem per setup/run! (src/problem.jl:321,
 * src/timestep.jl:277); a batch holds many one-dimensional problems
 * (n_global = (n, 1, 1)) that share the physics and numerics of one
 * cles_solver_desc and differ in length, domain, end time, step cap and
 * boundary states.  Each instance follows run! exactly (its own max_rate, dt,
 * check_step floors, tfinal clip and filter cadence); a failing instance is
 * frozen with its status while the others go on.  An instance's result does
 * not depend on the other instances of the batch.  Limits: 2 <= n <= 2048,
 * periodic, slip-wall or Dirichlet (constant state) ends. */
typedef struct cles_batch cles_batch;

typedef struct {
    int64_t n;                             /* points of this instance */
    double L, origin;                      /* domain [origin, origin + L] */
    double tfinal;                         /* run!(...; tfinal) */
    int64_t nmax;                          /* run!(...; nmax) */
    double bc_state[2][CLES_MAXSP + 4];    /* Dirichlet conserved state, low / high end */
} cles_batch_instance;

/* `common`: physics, schemes, art, cfl, filter cadence, step control, x-end
 * condition codes and precision; its n_global, L_domain, origin and bc_state
 * are ignored (they come per instance). */
int cles_batch_create(cles_ctx *ctx, const cles_solver_desc *common, int64_t count,
                      const cles_batch_instance *instances, cles_batch **batch);
int cles_batch_destroy(cles_batch *batch);
int cles_batch_shape(cles_batch *batch, int64_t *count, int64_t *total_points, int *n_cons);
/* Instance i's state in the reference's padded layout (padded = (n + 2H,
 * 1, 1), halo = (H, 0, 0), n_cons components), as cles_solver_upload. */
int cles_batch_upload(cles_batch *batch, int64_t i, const double *Q_host,
                      const int64_t padded[3], const int halo[3]);
int cles_batch_download(cles_batch *batch, int64_t i, double *Q_host,
                        const int64_t padded[3], const int halo[3]);
/* run! on every instance: at most max_steps steps of each; instances stop at
 * their own tfinal / nmax / failure.  dts (count x nrec, row i = instance i)
 * receives each instance's dt sequence of this call when nrec > 0.
 * *steps_taken: batch iterations issued. */
int cles_batch_run(cles_batch *batch, int64_t max_steps, int64_t nrec, double *dts, int64_t *steps_taken);
/* cles_batch_run without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as cles_batch_run; the dt history
 * (count x nrec doubles) is staged in a per-device array of
 * CLES_NOALLOC_DTS_MAX doubles (8 MiB) that each GPU the library runs on holds
 * until the process ends, cleared at the start of every call.  count * nrec
 * beyond CLES_NOALLOC_DTS_MAX is CLES_ERR_INVALID.  All calls on one GPU share
 * that array, so calls that overlap on one GPU (from several threads or
 * batches at once) race. */
#define CLES_NOALLOC_DTS_MAX (1 << 20)
int cles_batch_run_noalloc(cles_batch *batch, int64_t max_steps, int64_t nrec, double *dts,
                           int64_t *steps_taken);
/* How cles_batch_run issues its iterations (the batch always takes every
 * per-instance decision on the device).  CLES_RUN_HOST_DT and
 * CLES_RUN_DEVICE_DT: every kernel launched from the host.
 * CLES_RUN_STEP_GRAPH: groups of 8 iterations captured once as a HIP graph on
 * the context's stream (at the first run; again after a change of nrec or of
 * mode) and replayed; results are bitwise those of host issue.  While a
 * capture is in progress no other thread may issue work to that stream.
 * cles_batch_run_noalloc always issues from the host.  The reference has no
 * counterpart (src/timestep.jl:277 loops on the host). */
int cles_batch_set_run_mode(cles_batch *batch, int mode);
/* Per-instance time, completed steps and status (CLES_OK or the check_step
 * code 4..8 it stopped with); NULL arrays are skipped. */
int cles_batch_state(cles_batch *batch, double *t, int64_t *steps, int *status);
/* compute_rhs! of every instance into dQ (src/rhs.jl:624). */
int cles_batch_compute_rhs(cles_batch *batch, int primitives_current);
/* Instance i's interior of a named field: "rho","u","p","T","c","mu_art",
 * "beta_art","kappa_art","Q[c]","dQ[c]","Y[k]","D_art[k]". */
int cles_batch_get_field(cles_batch *batch, int64_t i, const char *name, double *host);
/* Kinetic energy of every instance (src/diagnostics.jl:61-80 with the
 * closed-edge half weights). */
int cles_batch_kinetic_energy(cles_batch *batch, double *ke);

/* ----------------------------------------------------------------------- */
/* Stacks of same-shape two-dimensional problems                            */
/* ----------------------------------------------------------------------- */
/* `count` problems of one cles_solver_desc (n_global = (nx, ny, 1),
 * Cartesian, uniform grid, no body forces) that differ in their initial
 * state, end time and step cap, held side by side in one solver so that every
 * kernel launch covers all of them.  Each follows run! on its own (its own
 * max_rate, dt, check_step floors, tfinal clip, filter cadence; a failing
 * instance is frozen with its status); an instance's result does not depend
 * on the other instances of the stack.  The reference runs one problem per
 * setup/run! (src/problem.jl:321, src/timestep.jl:277). */
typedef struct cles_stack cles_stack;
int cles_stack_create(cles_ctx *ctx, const cles_solver_desc *desc, int count, cles_stack **stack);
int cles_stack_destroy(cles_stack *stack);
/* Instance i's state in the padded layout of one problem (as cles_solver_upload). */
int cles_stack_upload(cles_stack *stack, int i, const double *Q_host, const int64_t padded[3],
                      const int halo[3]);
int cles_stack_download(cles_stack *stack, int i, double *Q_host, const int64_t padded[3],
                        const int halo[3]);
/* run!(...; tfinal[i], nmax[i]) on every instance, at most max_steps steps of
 * each; *steps_taken: stack iterations issued. */
int cles_stack_run(cles_stack *stack, const double *tfinal, const int64_t *nmax, int64_t max_steps,
                   int64_t *steps_taken);
/* Per-instance time, completed steps and status; NULL arrays are skipped. */
int cles_stack_state(cles_stack *stack, double *t, int64_t *steps, int *status);

/* ----------------------------------------------------------------------- */
/* Serving loop                                                             */
/* ----------------------------------------------------------------------- */
/* Independent solve requests (setup + run! of the reference, one problem
 * each: src/problem.jl:321, src/timestep.jl:277) admitted into device memory
 * by cles_solver_footprint under an HBM budget and run concurrently (up to
 * `runners` at once), each on its own context and stream.  Requests are
 * considered in arrival order; one that does not fit is passed over for
 * smaller ones until it has been passed over 4 times, after which nothing
 * behind it is admitted until it runs.  A request's result is bitwise the
 * same alone and in any mix.  No counterpart in the reference, which runs one
 * problem per process. */
typedef struct cles_server cles_server;

enum { CLES_REQ_QUEUED = 0, CLES_REQ_RUNNING = 1, CLES_REQ_DONE = 2, CLES_REQ_FAILED = 3 };

int cles_server_create(int device, size_t hbm_budget, int runners, cles_server **server);
/* Run mode (cles_solver_set_run_mode) of the solvers of requests admitted
 * from now on; default CLES_RUN_HOST_DT. */
int cles_server_set_run_mode(cles_server *server, int run_mode);
/* Same-shape two-dimensional requests (identical descriptors) queued
 * together are admitted together and run as one stack (cles_stack_*) of at
 * most `limit` requests (default 8; 1: every request alone); a stackable
 * request waits up to window_s (default 0.005) after its submission for the
 * rest of its stack.  Results are bitwise those of each request alone. */
int cles_server_set_stack_limit(cles_server *server, int limit, double window_s);
/* *size: requests of the stack a finished request ran in (1: alone). */
int cles_server_group(cles_server *server, int64_t id, int *size);
/* Marks queued requests CLES_REQ_FAILED and drops them, waits for running
 * ones, frees everything.  No other call on the server (cles_server_wait
 * included) may overlap it. */
int cles_server_destroy(cles_server *server);
/* Copies the padded initial state (as cles_solver_upload); CLES_ERR_NOMEM if
 * the request can never fit the budget. */
int cles_server_submit(cles_server *server, const cles_solver_desc *desc, const double *Q_host,
                       const int64_t padded[3], const int halo[3], double tfinal, int64_t nmax, int64_t *id);
/* Blocks until the request is done or failed, or for timeout_s (< 0: no
 * limit); *state is its CLES_REQ_* state. */
int cles_server_wait(cles_server *server, int64_t id, double timeout_s, int *state);
/* A finished request's state (same layout as submitted; NULL skips it), run!
 * status (CLES_OK or 3..10, message in cles_last_error), time and steps. */
int cles_server_result(cles_server *server, int64_t id, double *Q_host, const int64_t padded[3],
                       const int halo[3], int *status, double *t, int64_t *steps);
/* Seconds since the server started at submission, admission and completion,
 * and the admitted footprint. */
int cles_server_timing(cles_server *server, int64_t id, double *t_submit, double *t_admit, double *t_done,
                       size_t *footprint);
int cles_server_release(cles_server *server, int64_t id);
int cles_server_stats(cles_server *server, size_t *in_use, int *running, int *queued);

/* ----------------------------------------------------------------------- */
/* Line-tile autotuner (no reference counterpart: the reference solves its   */
/* lines serially per rank, src/tridiag.jl:31-57, src/banded.jl:38-80)       */
/* ----------------------------------------------------------------------- */

/* The chunk length of a direction's line plans (points per thread; it fixes
 * the chunks per line and with them the line kernel and its block shape).
 * cles_solver_create takes it from a process-wide table keyed by the grid,
 * precision, periodicity and scheme shapes; shapes the table does not hold
 * get the largest length (x <= 16, y/z <= 32, >= 8) that still launches
 * 131072 threads.  Nothing is tuned implicitly, so a table always gives the
 * same plans and the same results.
 *
 * cles_tune_solver times every candidate length (x: 16, 8, 4; y/z: 32, 16,
 * 8, 4) on the descriptor's grid (one four-field derivative launch and two
 * filter passes per repetition, `reps` repetitions) and stores the fastest
 * in the table; chunks[k] / ms_best[k] (either may be NULL) receive the
 * choice and its milliseconds per repetition, 0 for inactive directions.
 * Solvers created before the call keep their plans. */
int cles_tune_solver(cles_ctx *ctx, const cles_solver_desc *desc, int reps, int chunks[3], double ms_best[3]);
/* The lengths cles_solver_create would use for desc; tuned[k] (may be NULL)
 * is 1 where the table holds the entry.  Touches no device. */
int cles_tune_lookup(const cles_solver_desc *desc, int chunks[3], int tuned[3]);
/* Pins the lengths for desc's grid: chunks[k] > 0 sets direction k (at most
 * 32), 0 removes its entry (back to the rule), < 0 leaves it. */
int cles_tune_set(const cles_solver_desc *desc, const int chunks[3]);
int cles_tune_clear(void);
/* Text table, one entry per line: dim nx ny nz precision periodic deriv_q
 * deriv_M filt_M chunk.  Loading adds to (and overrides) the table. */
int cles_tune_save(const char *path);
int cles_tune_load(const char *path);

#ifdef __cplusplus
}
#endif
#endif
