/*
 * oracle.h -- CPU restatement of CompactLES.jl's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Every function restates the reference algorithm with its operation order
 * (built with -ffp-contract=off so a*b+c is never fused, as in Julia), on the
 * reference's own padded column-major layout (n_local + 2*H per active
 * dimension, H = 4, zero padding on collapsed dimensions) and with P = 1
 * (one rank, whole-domain lines).  Citations are path:line into the
 * reference repository (stillyslalom/CompactLES.jl).
 */
#ifndef CLES_ORACLE_H
#define CLES_ORACLE_H
#include <stdint.h>

#define OR_MAXCL 8      /* closure rows */
#define OR_MAXRHS 16    /* closure RHS length */
#define OR_MAXM 8       /* interior RHS half-width */
#define OR_MAXSP 8      /* species */

/* src/kernels.jl:43-50 (CompactScheme) and src/kernels_banded.jl:31-39
 * (BandedCompactScheme) as one record.  q = LHS half-bandwidth. */
typedef struct {
    int q;
    double lhs[4];                       /* lhs[s-1] at distance s */
    double a0;
    int M;
    double coeffs[OR_MAXM];
    int symmetric;
    int nclosure;
    double cl_lhs[OR_MAXCL][9];          /* 2q+1 centred entries */
    int cl_rhs_len[OR_MAXCL];
    double cl_rhs[OR_MAXCL][OR_MAXRHS];
} or_scheme;

void or_lele_d1_6(or_scheme *s);
void or_pade_d1_4(or_scheme *s);
void or_lele_d1_10(or_scheme *s);
void or_compact_filter(or_scheme *s, double alphaf);

/* One directional plan: src/operators.jl:23-101 / operators_banded.jl:13-125
 * with LineSolver / BandLineSolver (src/tridiag.jl, src/banded.jl) at P = 1. */
typedef struct or_plan or_plan;

/* n_local: interior extents; halo: per-dim pad (H on active dims, 0 otherwise).
 * Returns NULL and writes msg on the reference's validation errors. */
or_plan *or_plan_create(const or_scheme *s, int dim, const int64_t n_local[3],
                        const int halo[3], int periodic, double h,
                        char *msg, int msglen);
/* plan_direction(...; lo_fold, hi_fold) (operators.jl:44-101,
 * operators_banded.jl:29-125): lo_fold / hi_fold = sigma_g (+1 / -1) or 0 for
 * `nothing`.  The RHS reads halos mirror-filled by or_fold_fill_halos. */
or_plan *or_plan_create_fold(const or_scheme *s, int dim, const int64_t n_local[3],
                             const int halo[3], int periodic, double h, int lo_fold, int hi_fold,
                             char *msg, int msglen);
void or_plan_destroy(or_plan *p);
/* folds.jl:83-110 fold_fill!(f, decomp, d, lo, hi, sigma) on a padded scalar
 * field, one rank: halo layer q beyond a folded end of dimension dim (1-based)
 * holds sigma times interior layer q (half-offset mirror). */
void or_fold_fill_halos(double *f, const int64_t n_local[3], const int halo[3], int dim, int lo, int hi,
                        int sigma);
/* apply_along!(out, plan, f, decomp): f, out padded arrays (halos of f current) */
void or_apply_along(const or_plan *p, double *out, const double *f);

/* Fill periodic halos of a padded scalar field along every periodic active dim
 * (src/halo.jl:30-36 wrap_dim!, sequential x, y, z). */
void or_wrap_halos(double *f, const int64_t n_local[3], const int halo[3],
                   const int periodic[3]);

/* --------------------------- solver ------------------------------------- */
enum { OR_BC_PERIODIC = 0, OR_BC_SLIPWALL = 1, OR_BC_DIRICHLET = 2,
       OR_BC_NSCBC_OUT = 3, OR_BC_NSCBC_IN = 4, OR_BC_AXIS = 5, OR_BC_ORIGIN = 6, OR_BC_POLE = 7,
       OR_BC_NOSLIP = 8, OR_BC_EXTRAP = 9 };

typedef struct {
    double pinf, sigma, Lref, beta_t;            /* outflow; Lref <= 0: domain length */
    double u[3], T, Y[OR_MAXSP];                 /* inflow targets */
    double eta_u, eta_T, eta_t, eta_Y;
} or_nscbc;

/* One NASA-9 interval and species (physics.jl:280-375) */
typedef struct {
    double Tmin, Tmax;
    double a[7];
    double b1;
} or_nasa9_interval;
typedef struct {
    int n_intervals;
    or_nasa9_interval iv[3];
} or_nasa9_species;

typedef struct {
    int64_t n_global[3];
    double L_domain[3];
    double origin[3];
    int n_species;
    double R[OR_MAXSP], gamma[OR_MAXSP];
    double mu0, Pr, Sc;
    int art_enabled;
    double C_mu, C_beta, C_kappa, C_D;
    int beta_sensor;                     /* 0 strain, 1 gated, 2 dilatation */
    or_scheme deriv, filt;
    double cfl;
    int filter_interval;
    int bc[3][2];
    double bc_state[3][2][OR_MAXSP + 4]; /* Dirichlet: conserved state */
    /* NSCBC face parameters (src/nscbc.jl:45-54 outflow, :204-216 inflow) */
    or_nscbc nscbc[3][2];
    /* StepControl (src/stepcontrol.jl:116-136) */
    double predict, max_growth, dt_min, dt_min_ratio;
    int retries;
    double cfl_backoff;
    int savepoint_interval;
    int landing_steps;                   /* 0: the default 2 */
    /* EOS: 0 IdealMixture, 1 Nasa9Mixture (physics.jl:280-571) */
    int eos_kind;
    double T_guess;
    or_nasa9_species nasa9[OR_MAXSP];
    int metric;                       /* 0 CartesianMetric, 1 CylindricalMetric, 2 SphericalMetric */
    double Twall[3][2];               /* NoSlipWallBC: > 0 isothermal */
    int switchable[3][2];             /* SwitchableBC: bc until the switch, then bc_after */
    int bc_after[3][2];
    int n_body_forces;                /* ConstantBodyForce accelerations (sources.jl), in order */
    double body_force[4][3];
    double p_inf;                     /* StiffenedGas (eos_kind 2): gamma[0], R[0] = (gamma - 1) cv */
} or_desc;

/* Step-boundary hooks of run! (src/callbacks.jl): next scheduled instant,
 * after each completed step (nonzero stops), after a rollback. */
typedef struct {
    void *user;
    double (*next_time)(void *user, double t, int64_t step);
    int (*after_step)(void *user, double t, int64_t step, double dt);
    void (*rewind)(void *user, double t, int64_t step);
} or_run_hooks;

typedef struct or_solver or_solver;

or_solver *or_solver_create(const or_desc *d, char *msg, int msglen);
/* With Stretch mappings: jac[d] (NULL: uniform) holds dx/dxi at the
 * n_global[d] nodes of a stretched, non-periodic dimension. */
or_solver *or_solver_create_ex(const or_desc *d, const double *const jac[3], char *msg, int msglen);
/* boundary.jl:131-137 switch!: face (dim 0..2, side 0/1) takes bc_after; 0 ok, -1 not switchable */
int or_solver_switch_bc(or_solver *s, int dim, int side);
/* x[d]: physical coordinate at the interior nodes of a stretched dimension
 * whose coordinate the metric reads (r; spherical theta); NULL elsewhere. */
or_solver *or_solver_create_geom(const or_desc *d, const double *const x[3], const double *const jac[3], char *msg,
                                 int msglen);
void or_solver_destroy(or_solver *s);
void or_solver_dims(const or_solver *s, int64_t padded[3], int halo[3], int *n_cons);
/* Q in the padded layout (padded[0]*padded[1]*padded[2]*n_cons doubles). */
double *or_solver_Q(or_solver *s);
double *or_solver_dQ(or_solver *s);
/* Named internal fields, for tests: "rho","u","v","w","p","T","c","cp",
 * "mu_art","beta_art","kappa_art","sensor","strain_mag","grad_u[dj]" (d,j 1-based),
 * "grad_T[d]", "flux[dc]", "D_art[k]", "Y[k]" */
double *or_solver_field(or_solver *s, const char *name);
void or_solver_time(const or_solver *s, double *t, int64_t *step);
void or_solver_set_time(or_solver *s, double t, int64_t step);

/* conserved_from_prim for an ideal mixture (src/problem.jl:76-90):
 * prim = {rho or NaN, u1,u2,u3, p or NaN, T or NaN, Y[0..Ns-1]} */
void or_conserved_from_prim(const or_solver *s, const double *prim, double *q);

void or_apply_bcs(or_solver *s);
void or_compute_rhs(or_solver *s, int primitives_current);
void or_max_rate(or_solver *s, double *rate, double *rho_min);
void or_step(or_solver *s, double dt, int prepared);
void or_filter_state(or_solver *s);
/* run!(solver, Q; tfinal, nmax) with a fresh Workspace.  Returns 0 on success
 * or a SolverFailure code (1 nonfinite, 2 negative density, 3 planck,
 * 4 dt_min, 5 dt collapse).  dts (may be NULL) receives up to ndts dt values. */
int or_run(or_solver *s, double tfinal, int64_t nmax, double *dts, int64_t ndts);
/* run!(solver, Q; tfinal, nmax, callback, control): or_run with hooks (may be
 * NULL) and StepControl recovery (timestep.jl:286-344); *retries_used (may be
 * NULL) receives the number of rollbacks. */
int or_run_ctl(or_solver *s, double tfinal, int64_t nmax, const or_run_hooks *hooks, double *dts, int64_t ndts,
               int *retries_used);
double or_solver_cfl(const or_solver *s);
void or_solver_set_cfl(or_solver *s, double cfl);
/* Kinetic energy and enstrophy (volume integrals; enstrophy from compact grad_u
 * of the current state). */
void or_diagnostics(or_solver *s, double *ke, double *enstrophy);
double or_volume_integral(const or_solver *s, const double *f);

#endif
