// Argument records and launchers of the pointwise / stencil kernels.
#pragma once
#include "cles_internal.h"

namespace cles {

template <typename W>
struct GradPtrs {
    W *gu[9];   // grad_u[d*3 + j] = d u_j / d x_d ; NULL on collapsed d (reads as 0)
};

// Nasa9Mixture tables in device memory (physics.jl:402-571)
struct N9Table {
    int ns;
    double T_guess;
    double R[CLES_MAXSP];
    cles_nasa9_species sp[CLES_MAXSP];
};

// Host side: the EOS checks of a descriptor (physics.jl:312-338 for a
// Nasa9Mixture) and the device table of its Nasa9Mixture.
void validate_eos(const cles_solver_desc &d);
N9Table nasa9_table(const cles_solver_desc &d);

template <typename W>
struct PrimArgs {
    const double *Q;
    int Ns;
    double Rk[CLES_MAXSP], cvk[CLES_MAXSP];
    const N9Table *n9;   // Nasa9Mixture (NULL: IdealMixture)
    int sg;              // StiffenedGas (one species): gamma, p_inf, cv
    double sg_gamma, sg_pinf, sg_cv;
    W *rho, *u, *v, *w, *p, *T, *c, *cp;   // ideal gas: cp may be NULL (not stored)
    W *Y[CLES_MAXSP];                       // ideal gas: Y[k] may be NULL (not stored)
    W *esens;   // optional: internal energy of the kappa* sensor (artificial.jl:319-326)
};

template <typename W>
struct D4Args {
    const W *f;          // MODE 0
    const double *Q;     // MODE 1
    const W *rho;        // MODE 1
    int im1;             // first momentum component
    int64_t n[3];
    int64_t N;
    int active[3], periodic[3];
    double wd[3];        // h_d^wpow
};

// Up to D4_MF fields through one delta4 launch (grid y = field): field i
// reads f[i] and writes out[i] with the weights h_d^wpow[i] (h_d from
// D4Args::wd with wpow = 1, or the weights as given when wpow[i] == 0).
constexpr int D4_MF = 4;
template <typename W>
struct D4Multi {
    int nf;
    const W *f[D4_MF];
    W *out[D4_MF];
    int wpow[D4_MF];
};

template <typename W>
struct FluxArgs {
    const double *Q;
    int Ns;
    int active[3];
    double mu0, Pr, Sc;
    double cpk[CLES_MAXSP];
    double R1, cv1;      // one-species ideal gas: cp and Y_1 formed from rho where NULL
    const N9Table *n9;   // Nasa9Mixture: species enthalpies h_k(T) (NULL: cp_k T)
    const W *rho, *u, *v, *w, *p, *T, *cp;
    const W *mu_art, *beta_art, *kappa_art;
    const W *Y[CLES_MAXSP];
    const W *D_art[CLES_MAXSP];
    GradPtrs<W> g;
    const W *gT[3];
    const W *gY[3][CLES_MAXSP];
    W *flux[3][CLES_MAXSP + 4];
};

// Orthogonal metric of the grid (metric.jl:80-137) from node tables: kind
// CLES_METRIC_*; m[d] the mapping Jacobian dx/dxi at the interior nodes of a
// stretched dimension (NULL: 1); r the radius at the nodes of dimension 1
// (curvilinear); sth, cth sin and cos of theta at the nodes of dimension 2
// (spherical); cotr cot(theta)/r over the interior in its discrete-GCL form
// (spherical with theta resolved, metric.jl:150-168; NULL: cos / (r sin)).
// Scale factors h = (1, r, 1) cylindrical, (1, r, r sin theta) spherical,
// times m_d, clamped to tiny; J = h1 h2 h3, A_d = J / h_d.
struct Metric1D {
    int kind;
    const double *m[3];
    const double *r, *sth, *cth;
    const double *cotr;
    int64_t n[3];
};

// Curvature terms of a curvilinear metric (metric.jl:181-269): corrections
// added to the scaled velocity gradients, and the momentum sources of the
// divergence of Pi.  zero[d]: dimension d is collapsed, its grad_u row holds
// no derivative and takes the correction alone.
template <typename W>
struct CurvArgs {
    Metric1D g;
    int zero[3];
    W *gu[9];
    const W *u, *v, *w, *rho, *p, *mu_art, *beta_art;
    double mu0;
    int im1;     // first momentum component of dQ
    int64_t N;
    W *dQ;
};
// Antipodal pairing of a fold (folds.jl:118-246): the partner of interior
// point (i, j, k) shifts pdim by half its extent and reverses revdim (-1:
// none); canonical points sit in the lower half of sd.
struct PairMap {
    int64_t n[3];
    int pdim, revdim, sd;
};
// e / o butterfly of f into w (sf: the field's antipodal sign)
template <typename T> void launch_pair_forward(const T *f, T *w, const PairMap &pm, double sf, hipStream_t st);
// inverse butterfly of the even-plan result re (canonical slots) and the
// odd-plan result ro (partner slots), applied to dst with epilogue kind epi
// (EPI_STORE, EPI_STORE_NEG, EPI_SUB, EPI_STORE_GATE with gate)
template <typename T>
void launch_pair_backward(const T *re, const T *ro, T *dst, const PairMap &pm, double sf, int epi, const int *gate,
                          hipStream_t st);

template <typename W> void launch_metric_grad(const CurvArgs<W> &a, hipStream_t st);
// sources.jl:44-65 ConstantBodyForce: dQ_m += rho g, dQ_E += (rho u) . g
template <typename W>
void launch_body_force(const double *Q, W *dQ, int Ns, int64_t N, const double (&g)[3], hipStream_t st);
template <typename W> void launch_metric_sources(const CurvArgs<W> &a, hipStream_t st);

template <typename W>
struct RateArgs {
    const double *Q;
    int Ns;
    int active[3];
    double invh[3];
    double h[3];             // spacing of each dimension (stretched: computational)
    Metric1D met;            // stretched dimensions use inv_h[d] / h[d] (timestep.jl:151)
    double mu0, Pr, Sc;
    double R1, cv1;          // one-species ideal gas: cp formed from rho where cp is NULL
    const W *rho, *u, *v, *w, *c, *cp;
    const W *mu_art, *beta_art, *kappa_art;
    const W *D_art[CLES_MAXSP];
};

struct PlaneArgs {
    int dim, side, Ns;
    int64_t n[3];
    int64_t N, count;
    double state[CLES_MAXSP + 4];
    double Twall;                   // NoSlipWallBC: > 0 isothermal
    double cvk[CLES_MAXSP];         // IdealMixture cv_k (wall energy)
    const N9Table *n9;              // Nasa9Mixture: e_k(Twall) (NULL: cv_k Twall)
    int sg;                         // StiffenedGas: rho cv Twall + p_inf (cvk[0] = cv)
    double sg_pinf;
};

// One NSCBC face (nscbc.jl:59-154 outflow, :220-306 inflow, constant targets).
template <typename W>
struct NscbcArgs {
    int dim, side, Ns, inflow;
    int64_t n[3];
    int64_t N, count;
    int act1, act2;                 // transverse dimensions active (outflow)
    const W *rho, *c, *p, *T, *cp, *vel[3];
    const W *Y[CLES_MAXSP], *gY[CLES_MAXSP];   // gY along dim; NULL reads as 0
    const W *gu[9];
    const W *dpn, *dt1, *dt2;       // dp/dx_dim; outflow dp/dx_t1, dp/dx_t2; inflow drho/dx_dim in dt1
    double Lref, pinf, sigma, beta_t;
    double ut[3], Tt, Yt[CLES_MAXSP], eta_u, eta_T, eta_t, eta_Y;
    double Rk[CLES_MAXSP], cvk[CLES_MAXSP];
    const N9Table *n9;              // Nasa9Mixture: cv_k(T) in the inflow's eos_dphi_dY (NULL: cvk)
    double sg_phi;                  // StiffenedGas: phi = 1 / (gamma - 1), dphi/dY = 0 (0: not one)
    Metric1D met;                   // inv_h of the derivatives (stretched or curvilinear grids)
    W *dQ;
};

// ---------------------------------------------------------------------------
// Batches of independent one-dimensional instances laid end to end (instance
// i occupies [off[i], off[i] + n[i]) of every field; N = total points).
// ---------------------------------------------------------------------------
struct BatchGeom {
    int64_t count, N;
    const int64_t *off;
    const int *n;
    const double *h;        // grid spacing per instance
    int periodic;           // every instance is periodic (else closed ends)
};

// run! control per instance (timestep.jl:277-400, stepcontrol.jl:184-204)
struct BatchCtl {
    int64_t count;
    const double *rate, *rhomin, *tfinal;
    double *t, *dt, *dt_prev, *rate_prev, *dt_seen;
    int64_t *step;
    const int64_t *nmax;
    int *status, *active, *stepped, *filt;
    double cfl, predict, max_growth, dt_min, dt_min_ratio;
    int filter_interval;
    double *dts;            // dt history of this run [i * nrec + step - step0 - 1], or NULL
    const int64_t *step0;   // step count of each instance when the run began
    int64_t nrec;
};

struct BatchBC {
    int lo, hi;             // CLES_BC_* of the two ends
    int Ns;
    const double *state;    // Dirichlet conserved states [i][side][Ns + 4]
    const int *mask;        // apply only where nonzero (NULL: every instance)
};

template <typename W>
struct DiagArgs {
    const double *Q;
    int Ns;
    int64_t n[3];
    int64_t N;
    int closed[3];
    int fold_lo[3], fold_hi[3];   // a folded (half-offset) edge takes a full weight (diagnostics.jl:41-47)
    GradPtrs<W> g;
    Metric1D met;            // stretched: integrand / inv_J (diagnostics.jl:73-75)
};

// Metric scalings over nf fields of N points (interior, x fastest):
// mode 0: f *= inv_h[dim] (rhs.jl:477-483 _scale_grad!); mode 1: f *= J / h_dim
// (the area factor A_dim, rhs.jl:672-677); mode 2: f *= inv_J (rhs.jl:680-682);
// mode 3: f = J / h_dim (metric.jl:124-126 area_d).
constexpr int METRIC_MAXF = 12;
template <typename W>
struct MetricFields {
    W *f[METRIC_MAXF];
    int nf;
};
template <typename W>
void launch_metric_scale(const MetricFields<W> &fs, const Metric1D &g, int mode, int dim, int64_t N, hipStream_t st);

template <typename W> void launch_primitives(const PrimArgs<W> &a, int64_t N, hipStream_t st);
template <typename W> void launch_strain(const GradPtrs<W> &g, W *out, int64_t N, hipStream_t st);
template <typename W> void launch_delta4(const D4Args<W> &a, int mode, W *out, hipStream_t st);
template <typename W> void launch_delta4_multi(const D4Args<W> &a, const D4Multi<W> &m, hipStream_t st);
template <typename W> void launch_gate_beta(const GradPtrs<W> &g, W *beta, int64_t N, hipStream_t st);
template <typename W>
void launch_dilatation_prep(const GradPtrs<W> &g, W *div, W *sw, int64_t N, hipStream_t st);
template <typename W>
void launch_dilatation_beta(const W *rho, const W *sensor, const W *sw, W *beta, double C_beta, int64_t N,
                            hipStream_t st);
template <typename W> void launch_fluxes(const FluxArgs<W> &a, int64_t N, hipStream_t st);
template <typename W>
void launch_lsrk(double *Q, double *du, const W *dQ, double A, double B, double dt, const double *dtp,
                 const int *gate, int64_t NC, int64_t N, unsigned ipts, hipStream_t st);
template <typename W>
void launch_rate(const RateArgs<W> &a, double *part, int nb, double *out, int64_t N, hipStream_t st);
template <typename W>
void launch_rate_inst(const RateArgs<W> &a, double *part, int nb, int ni, double *rate, double *rmin, int64_t N,
                      hipStream_t st);
template <typename W>
void launch_nscbc(const NscbcArgs<W> &a, hipStream_t st);
template <typename W>
void launch_nscbc_pair(const NscbcArgs<W> &a0, const NscbcArgs<W> &a1, hipStream_t st);
template <typename W>
void launch_rate_batch(const RateArgs<W> &a, const BatchGeom &g, const int *mask, double *rate, double *rhomin,
                       hipStream_t st);
template <typename W>
void launch_delta4_batch(const W *f, W *out, const BatchGeom &g, int wpow, hipStream_t st);
template <typename W>
void launch_delta4_batch_multi(const D4Multi<W> &m, const BatchGeom &g, hipStream_t st);
void launch_batch_dt(const BatchCtl &c, hipStream_t st);
void launch_batch_post(const BatchCtl &c, hipStream_t st);
void launch_batch_bcs(const BatchGeom &g, const BatchBC &b, double *Q, hipStream_t st);
void launch_batch_count(const int *flags, int64_t count, int *out, hipStream_t st);
void launch_ke_batch(const double *Q, int Ns, const BatchGeom &g, double *ke, hipStream_t st);
void launch_slipwall(const PlaneArgs &a, double *Q, hipStream_t st);
void launch_dirichlet(const PlaneArgs &a, double *Q, hipStream_t st);
void launch_noslip(const PlaneArgs &a, double *Q, hipStream_t st);
void launch_extrapolate(const PlaneArgs &a, double *Q, hipStream_t st);
template <typename W>
void launch_integral(const DiagArgs<W> &a, int kind, double *part, int nb, double scale, double *out,
                     hipStream_t st);
void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
template <typename Dst, typename Src> void launch_convert(Dst *d, const Src *s, int64_t N, hipStream_t st);

}  // namespace cles
