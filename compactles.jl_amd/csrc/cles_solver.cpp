// Device-resident solver: setup, compute_rhs!, step!, max_rate,
// filter_state!, run! and diagnostics on one GPU (whole-domain lines).
//
// Dataflow per RHS follows src/rhs.jl:624-694 with these fusions, none of
// which changes the arithmetic of a value:
//  * the last pass of every sensor smoothing (artificial.jl:98-111) writes the
//    coefficient fields directly (mu*/beta*, kappa*, D*_k epilogues);
//  * the e-sensor's internal energy (artificial.jl:319-326) is evaluated inside
//    the delta4 stencil instead of being stored;
//  * the flux divergence accumulates in place (dQ = -D_1 F_1, dQ -= D_d F_d)
//    and, when no boundary correction follows, the last direction feeds the
//    LSRK update (timestep.jl:85-94) without storing the final dQ.
// Periodic halos are index wraps inside the kernels (halo.jl:27-36).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cles_internal.h"
#include "cles_point.h"

namespace {

const double *const RKA = cles::LSRK_A, *const RKB = cles::LSRK_B, *const RKC = cles::LSRK_C;
const double PLANCK_TIME = 5.391247e-44;   // stepcontrol.jl:63

inline double jmax(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
inline double jmin(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }

}  // namespace

struct cles_solver {
    cles_ctx *ctx = nullptr;
    cles_solver_desc d{};
    int Ns = 1, C = 5, prec = CLES_F64;
    int active[3] = {0, 0, 0}, periodic[3] = {1, 1, 1};
    int64_t n[3] = {1, 1, 1}, N = 1;
    double h[3] = {1, 1, 1};
    double Rk[CLES_MAXSP] = {}, cvk[CLES_MAXSP] = {}, cpk[CLES_MAXSP] = {};
    int nact = 0, adim[3] = {0, 0, 0};
    bool nscbc = false;              // an active face currently runs NSCBC
    bool nscbc_any = false;          // ... or may after a SwitchableBC switch (scratch held)
    bool switched[3][2] = {};
    cles::HostPlan dplan[3], fplan[3], fplanQ[3];
    bool own_fplanQ = false;
    // coordinate-singularity folds (rhs.jl:143-279, folds.jl:56-73) per
    // dimension: folded ends, antipodal pairing (pdim shifted by half, revdim
    // reversed; -1: none, self-paired), antipodal signs of (u, v, w), the area
    // factor's parity, and parity-folded plans (index 0: sigma_g = +1, 1: -1)
    struct Fold {
        bool on = false, lo = false, hi = false;
        int pdim = -1, revdim = -1;
        int sv[3] = {1, 1, 1};
        int area = 1;
        cles::HostPlan dp[2], fp[2], fpQ[2];
        bool paired() const { return pdim >= 0 || revdim >= 0; }
    } fold[3];
    void *pairw = nullptr, *paira = nullptr, *pairb = nullptr;   // pairing scratch, 8 bytes per point
    // Metric and Stretch (metric.jl:45-137): node tables on the device;
    // stretched: some dimension has a mapping, curv: a curvilinear metric
    bool stretched = false, curv = false;
    cles::Metric1D met{};
    bool geom() const { return stretched || curv; }
    const cles::N9Table *n9 = nullptr;   // Nasa9Mixture tables on the device (NULL: IdealMixture)
    std::vector<void *> allocs;
    size_t bytes = 0;                          // device bytes allocated by this solver
    // device state (W-typed pointers are void* here; the Impl casts them)
    double *Q = nullptr, *du = nullptr;
    void *dQ = nullptr;
    void *rho = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr, *T = nullptr, *c = nullptr,
         *cp = nullptr;
    void *Y[CLES_MAXSP] = {}, *gu[9] = {}, *gT[3] = {}, *gY[3][CLES_MAXSP] = {};
    void *mu_art = nullptr, *beta_art = nullptr, *kappa_art = nullptr, *D_art[CLES_MAXSP] = {};
    void *sensor = nullptr, *strain = nullptr, *tmp1 = nullptr, *tmp2 = nullptr, *esens = nullptr;
    void *nsd[3] = {};                         // NSCBC pressure / density derivatives
    // ... or, when they go with the gradient launches (nsd_merged), D_k p and
    // D_k rho for each active dimension k
    void *nsdp[3] = {}, *nsdr[3] = {};
    bool nsd_merged = false;
    void *flux[3][CLES_MAXSP + 4] = {};
    double *part = nullptr, *scal = nullptr;   // reduction partials, device scalars
    int nred = 1024;
    double *host_scal = nullptr;               // pinned
    double t = 0, tstage = 0, dt_prev = 0, rate_prev = 0;
    int64_t step = 0;

    // run! with the dt decision in device memory and each iteration replayed
    // from a captured graph (timestep.jl:345-399 without a host round trip)
    struct Ctl {   // device control block of one instance
        double t, dt, dt_prev, rate_prev, dt_seen, tfinal, pad[2];
        int64_t step, nmax, step0, pad64;
        int status, active, stepped, filt, pad32[4];
    };
    Ctl *ctl = nullptr, *host_ctl = nullptr;   // device / pinned mirror
    double *dts_buf = nullptr;                 // dt history of a run (capacity dts_cap)
    int64_t dts_cap = 0;
    int run_mode = 0;                          // CLES_RUN_HOST_DT, _DEVICE_DT or _STEP_GRAPH
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // set while the iteration is captured: LSRK dt and gates in device memory
    const double *dev_dt = nullptr;
    const int *dev_gate = nullptr, *dev_filt = nullptr;
    // A stack (cles_stack): `stack` two-dimensional instances of the
    // descriptor's grid side by side along z (n[2] = stack planes of ipts
    // points, z inactive).  dev_dt / dev_gate / dev_filt then hold one entry
    // per instance, and the per-step rates go to srate / srmin.  ipts = 0:
    // not a stack.
    int stack = 0;
    unsigned ipts = 0;
    double *srate = nullptr, *srmin = nullptr, *spart = nullptr;

    // the kappa* and D*_k sensors beside `sensor` (mu*, beta*), smoothed in
    // the same launches, and the coefficient epilogue of each (device table)
    void *sensorK = nullptr, *sensD[CLES_MAXSP] = {};
    cles::EpiArgs *art_multi = nullptr;

    // per-phase event timing (cles_solver_profile)
    struct Phase {
        double ms = 0, bytes = 0;
        int64_t launches = 0;
    };
    struct Pending {
        int phase;
        hipEvent_t a, b;
        double bytes;
    };
    bool prof = false;
    std::vector<Phase> phases;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> evpool;

    hipEvent_t event()
    {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        CLES_HIP(hipEventCreate(&e));
        return e;
    }
    template <typename F>
    void timed(int phase, double bytes, F &&launch)
    {
        if (!prof) {
            launch();
            return;
        }
        hipEvent_t a = event(), b = event();
        CLES_HIP(hipEventRecord(a, ctx->stream));
        launch();
        CLES_HIP(hipEventRecord(b, ctx->stream));
        pending.push_back({phase, a, b, bytes});
    }
    void resolve()
    {
        if (pending.empty()) return;
        CLES_HIP(hipStreamSynchronize(ctx->stream));
        for (auto &pd : pending) {
            float ms = 0.f;
            CLES_HIP(hipEventElapsedTime(&ms, pd.a, pd.b));
            Phase &ph = phases[pd.phase];
            ph.ms += ms;
            ph.bytes += pd.bytes;
            ph.launches += 1;
            evpool.push_back(pd.a);
            evpool.push_back(pd.b);
        }
        pending.clear();
    }

    hipStream_t st() const { return ctx->stream; }
    size_t wsize() const { return prec == CLES_F64 ? 8 : 4; }
    // one-species ideal gas that never runs NSCBC: the primitives do not
    // store Y_1 and cp (physics.jl:604-653 stores both per point); the fluxes
    // and the rate form them from rho by the same operations, same bits
    bool const_cp() const { return Ns == 1 && d.eos_kind == CLES_EOS_IDEAL && !nscbc_any; }

    void *alloc(size_t bytes)
    {
        void *ptr = nullptr;
        hipError_t e = hipMalloc(&ptr, bytes ? bytes : 8);
        if (e != hipSuccess) throw cles::Error{CLES_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
        allocs.push_back(ptr);
        this->bytes += bytes ? bytes : 8;
        CLES_HIP(hipMemsetAsync(ptr, 0, bytes ? bytes : 8, st()));
        return ptr;
    }
    void *wfield() { return alloc((size_t)N * wsize()); }
    // Creation-time upload of a host table into memory from alloc(): on the
    // same stream as alloc's zero-fill (a plain hipMemcpy runs on the null
    // stream, which a non-blocking stream does not wait for, so the fill could
    // land after it), and complete before the host buffer goes away.
    void upload(void *dst, const void *src, size_t bytes)
    {
        CLES_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st()));
        CLES_HIP(hipStreamSynchronize(st()));
    }

    ~cles_solver()
    {
        if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (auto &pd : pending) { evpool.push_back(pd.a); evpool.push_back(pd.b); }
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dts_buf) (void)hipFree(dts_buf);
        for (void *ptr : allocs) (void)hipFree(ptr);
        if (host_scal) (void)hipHostFree(host_scal);
        if (host_ctl) (void)hipHostFree(host_ctl);
        for (int k = 0; k < 3; k++) {
            cles::free_line_plan(dplan[k]);
            cles::free_line_plan(fplan[k]);
            if (own_fplanQ) cles::free_line_plan(fplanQ[k]);
        }
        for (int g = 0; g < 6; g++) {
            Fold &fo = fold[g / 2];
            cles::free_line_plan(fo.dp[g % 2]);
            cles::free_line_plan(fo.fp[g % 2]);
            if (own_fplanQ) cles::free_line_plan(fo.fpQ[g % 2]);
        }
        if (ctx) cles::ctx_release(ctx);
    }
};

namespace cles {

// Phases of the per-step path, after bench/phases.jl:119-193.  The five line
// phases are split by sweep direction (x, y, z).
enum Phase {
    PH_DERIV = 0,           // + dim: velocity and temperature gradients
    PH_DIV = 3,             // + dim: flux divergence (store / accumulate)
    PH_DIV_LSRK = 6,        // + dim: last divergence pass fused with the LSRK update
    PH_SMOOTH = 9,          // + dim: sensor smoothing filters
    PH_FILTER_STATE = 12,   // + dim: conserved-state filter
    PH_PRIMITIVES = 15,
    PH_STRAIN,
    PH_DELTA4,
    PH_FLUXES,
    PH_RATE,
    PH_BCS,
    PH_LSRK,
    PH_NSCBC,
    PH_OTHER,
    PH_N
};
static const char *const PHASE_NAMES[PH_N] = {
    "line_deriv_x", "line_deriv_y", "line_deriv_z", "line_div_x", "line_div_y", "line_div_z",
    "line_div_lsrk_x", "line_div_lsrk_y", "line_div_lsrk_z", "line_smooth_x", "line_smooth_y", "line_smooth_z",
    "line_filter_state_x", "line_filter_state_y", "line_filter_state_z", "primitives", "strain", "delta4",
    "fluxes", "max_rate", "bcs", "lsrk", "nscbc", "other"};

template <typename W>
struct Impl {
    cles_solver &s;
    explicit Impl(cles_solver &s_) : s(s_) {}

    W *f(void *ptr) const { return (W *)ptr; }
    View3 view() const
    {
        View3 v;
        for (int k = 0; k < 3; k++) v.n[k] = s.n[k];
        v.s[0] = 1;
        v.s[1] = s.n[0];
        v.s[2] = s.n[0] * s.n[1];
        return v;
    }
    GradPtrs<W> grads() const
    {
        GradPtrs<W> g;
        for (int m = 0; m < 9; m++) g.gu[m] = f(s.gu[m]);
        return g;
    }
    // algorithmic bytes per point of one line operator with epilogue kind
    double line_bytes(const HostPlan &hp, int kind) const
    {
        const double w = hp.dtype == CLES_F64 ? 8.0 : 4.0;
        switch (kind) {
        case EPI_SUB: return 3 * w;                 // read f, read-modify-write out
        case EPI_LSRK:
        case EPI_LSRK_DEV: return 2 * w + 32;       // f, dQ; du and Q read + written
        case EPI_ART_MUBETA: return 4 * w;          // sensor, rho -> mu, beta
        case EPI_ART_KAPPA: return 5 * w;           // sensor, rho, c, T -> kappa
        case EPI_ART_D: return 3 * w;               // sensor, c -> D
        case EPI_ART_MULTI:                         // the mean over mu*/beta*, kappa*, D*_k
            return w * (4.0 + 5.0 + 3.0 * (s.Ns > 1 ? s.Ns : 0)) / (2 + (s.Ns > 1 ? s.Ns : 0));
        default: return 2 * w;                      // read f, write out
        }
    }
    void line(const HostPlan &hp, int d, const void *in, EpiArgs e, int phase) const
    {
        s.timed(phase + d, line_bytes(hp, e.kind) * (double)s.N,
                [&] { launch_line(hp, d + 1, view(), in, e, s.st()); });
    }
    // one launch over nf independent fields: field i reads in[i] and writes
    // out[i] (store kinds) or offsets the LSRK arrays of e by eoff[i] elements
    void lines(const HostPlan &hp, int d, int nf, const void *const *in, void *const *out, const int64_t *eoff,
               EpiArgs e, int phase) const
    {
        LineFields mf;
        mf.nf = nf;
        for (int i = 0; i < nf; i++) {
            mf.in[i] = in[i];
            if (out) mf.out[i] = out[i];
            if (eoff) mf.eoff[i] = eoff[i];
        }
        if (out) e.out = out[0];
        s.timed(phase + d, line_bytes(hp, e.kind) * (double)s.N * nf,
                [&] { launch_line(hp, d + 1, view(), in[0], e, s.st(), &mf); });
    }
    // equations.jl:56-66 parities across the fold of dimension d (the axis:
    // u_r and u_theta odd); flux parity = conserved x normal velocity x area
    // (A_1 = r is odd); plans by field parity, a derivative flipping it and a
    // filter keeping it (folds.jl:271-274)
    int vel_parity(int d, int j) const { return s.fold[d].on ? s.fold[d].sv[j] : 1; }
    int cons_parity(int d, int c) const { return (c >= s.Ns && c < s.Ns + 3) ? vel_parity(d, c - s.Ns) : 1; }
    int flux_parity(int d, int c) const
    {
        return cons_parity(d, c) * vel_parity(d, d) * (s.fold[d].on ? s.fold[d].area : 1);
    }
    enum PlanKind { DERIV, FILT, FILT_Q };
    const HostPlan &plan(PlanKind k, int d, int sf) const
    {
        if (s.fold[d].on) {
            const int g = (k == DERIV ? -sf : sf) > 0 ? 0 : 1;
            return k == DERIV ? s.fold[d].dp[g] : k == FILT ? s.fold[d].fp[g] : s.fold[d].fpQ[g];
        }
        return k == DERIV ? s.dplan[d] : k == FILT ? s.fplan[d] : s.fplanQ[d];
    }
    PairMap pair_map(int d) const
    {
        PairMap pm{};
        for (int k = 0; k < 3; k++) pm.n[k] = s.n[k];
        pm.pdim = s.fold[d].pdim;
        pm.revdim = s.fold[d].revdim;
        pm.sd = pm.pdim >= 0 ? pm.pdim : pm.revdim;
        return pm;
    }
    // folds.jl:276-330 fold_apply! for a paired fold, one field: the e / o
    // butterfly, both parity plans over every line (the e half folds with sf,
    // the o half as odd), and the inverse butterfly into out with e's
    // epilogue kind (store, negated store, subtraction, gated store)
    template <typename T>
    void pair_apply(PlanKind k, int d, const void *in, void *out, int sf, EpiArgs e, int phase) const
    {
        const PairMap pm = pair_map(d);
        T *w = (T *)s.pairw, *a = (T *)s.paira, *b = (T *)s.pairb;
        s.timed(PH_OTHER, 2.0 * 2 * sizeof(T) * (double)s.N,
                [&] { launch_pair_forward<T>((const T *)in, w, pm, (double)sf, s.st()); });
        line(plan(k, d, sf), d, w, store(a), phase);
        line(plan(k, d, -1), d, w, store(b), phase);
        if (!(e.kind == EPI_STORE || e.kind == EPI_STORE_NEG || e.kind == EPI_SUB || e.kind == EPI_STORE_GATE))
            throw Error{CLES_ERR_INVALID, "epilogue not available through a paired fold"};
        s.timed(PH_OTHER, 2.0 * 3 * sizeof(T) * (double)s.N, [&] {
            launch_pair_backward<T>(a, b, (T *)out, pm, (double)sf, e.kind, e.gate, s.st());
        });
    }
    // line() of one field of parity sf along d, through the fold if any
    void line_par(PlanKind k, int d, const void *in, int sf, EpiArgs e, int phase) const
    {
        if (s.fold[d].on && s.fold[d].paired()) {
            if (k == FILT_Q) pair_apply<double>(k, d, in, e.out, sf, e, phase);
            else pair_apply<W>(k, d, in, e.out, sf, e, phase);
            return;
        }
        line(plan(k, d, sf), d, in, e, phase);
    }
    // lines() over fields of parities par[]: one launch per plan when a fold
    // separates them, one pass per field through a paired fold
    void lines_par(PlanKind k, int d, int nf, const void *const *in, void *const *out, const int *par, EpiArgs e,
                   int phase) const
    {
        if (!s.fold[d].on) {
            lines(plan(k, d, 1), d, nf, in, out, nullptr, e, phase);
            return;
        }
        if (s.fold[d].paired()) {
            for (int i = 0; i < nf; i++) {
                EpiArgs ei = e;
                ei.out = out[i];
                line_par(k, d, in[i], par[i], ei, phase);
            }
            return;
        }
        for (int sg : {1, -1}) {
            const void *pin[LINE_MF];
            void *pout[LINE_MF];
            int m = 0;
            for (int i = 0; i < nf; i++)
                if (par[i] == sg) {
                    pin[m] = in[i];
                    pout[m++] = out ? out[i] : nullptr;
                }
            if (m) lines(plan(k, d, sg), d, m, pin, out ? pout : nullptr, nullptr, e, phase);
        }
    }

    EpiArgs store(void *out, int kind = EPI_STORE) const
    {
        EpiArgs e{};
        e.kind = kind;
        e.out = out;
        return e;
    }

    // physics.jl:604-653
    void primitives() const
    {
        PrimArgs<W> a{};
        a.Q = s.Q;
        a.Ns = s.Ns;
        a.n9 = s.n9;
        for (int k = 0; k < s.Ns; k++) { a.Rk[k] = s.Rk[k]; a.cvk[k] = s.cvk[k]; a.Y[k] = f(s.Y[k]); }
        a.sg = s.d.eos_kind == CLES_EOS_STIFFENED;
        a.sg_gamma = s.d.gamma[0];
        a.sg_pinf = s.d.p_inf;
        a.sg_cv = s.cvk[0];
        a.rho = f(s.rho); a.u = f(s.u); a.v = f(s.v); a.w = f(s.w);
        a.p = f(s.p); a.T = f(s.T); a.c = f(s.c); a.cp = f(s.cp);
        a.esens = f(s.esens);
        const bool cc = s.const_cp();
        if (cc) a.cp = a.Y[0] = nullptr;
        const double w = sizeof(W);
        s.timed(PH_PRIMITIVES, (8.0 * s.C + w * ((cc ? 7 : 8 + s.Ns) + (s.esens ? 1 : 0))) * (double)s.N,
                [&] { launch_primitives<W>(a, s.N, s.st()); });
    }

    // rhs.jl:591-606 velocity gradients of the active dimensions; with
    // scalars, also rhs.jl:608-622 the temperature and species gradients,
    // every field of one direction in one launch
    void velocity_gradients(bool scalars = false) const
    {
        for (int k = 0; k < s.nact; k++) gradients_along(k, scalars);
        if (s.curv) {   // rhs.jl:604 metric_correct_gradients!
            const CurvArgs<W> a = curv_args();
            s.timed(PH_OTHER, sizeof(W) * 10.0 * (double)s.N, [&] { launch_metric_grad<W>(a, s.st()); });
        }
    }
    // the gradients along the k-th active dimension
    void gradients_along(int k, bool scalars) const
    {
        {
            const int d = s.adim[k];
            const void *in[LINE_MF] = {s.u, s.v, s.w};
            void *out[LINE_MF] = {s.gu[d * 3], s.gu[d * 3 + 1], s.gu[d * 3 + 2]};
            int par[LINE_MF] = {vel_parity(d, 0), vel_parity(d, 1), vel_parity(d, 2)};
            int nf = 3;
            if (scalars) {
                par[nf] = 1;
                in[nf] = s.T;
                out[nf++] = s.gT[d];
                if (s.Ns > 1)
                    for (int sp = 0; sp < s.Ns; sp++) {
                        par[nf] = 1;
                        in[nf] = s.Y[sp];
                        out[nf++] = s.gY[d][sp];
                    }
                if (s.nscbc && s.nsd_merged) {   // the NSCBC faces' full-field derivatives
                    bool np[3], nr[3];
                    nscbc_needs(np, nr);
                    if (np[d]) {
                        par[nf] = 1;
                        in[nf] = s.p;
                        out[nf++] = s.nsdp[d];
                    }
                    if (nr[d]) {
                        par[nf] = 1;
                        in[nf] = s.rho;
                        out[nf++] = s.nsdr[d];
                    }
                }
            }
            lines_par(DERIV, d, nf, in, out, par, store(out[0]), PH_DERIV);
            if (s.met.m[d] || s.curv) metric(out, nf, 0, d);   // rhs.jl:599,631,634 _scale_grad!
        }
    }


    CurvArgs<W> curv_args() const
    {
        CurvArgs<W> a{};
        a.g = s.met;
        for (int k = 0; k < 3; k++) a.zero[k] = !s.active[k];
        for (int k = 0; k < 9; k++) a.gu[k] = f(s.gu[k]);
        a.u = f(s.u); a.v = f(s.v); a.w = f(s.w); a.rho = f(s.rho); a.p = f(s.p);
        a.mu_art = f(s.mu_art);
        a.beta_art = f(s.beta_art);
        a.mu0 = s.d.mu0;
        a.im1 = s.Ns;
        a.N = s.N;
        a.dQ = f(s.dQ);
        return a;
    }

    // metric.jl:150-168 gcl_cotr!: cot(theta)/r := inv_J D_xi2(A_2), with
    // A_2 = J / h_2 filled over the grid and differentiated along theta
    void gcl_cotr(double *cotr) const
    {
        void *a2 = s.flux[s.adim[0]][0];   // flux scratch, free outside the RHS
        void *da = s.flux[s.adim[0]][1];
        metric(&a2, 1, 3, 1);
        // with the sign of the theta-momentum pressure flux across a pole fold
        line_par(DERIV, 1, a2, flux_parity(1, s.Ns + 1), store(da), PH_OTHER);
        metric(&da, 1, 2, 0);
        launch_convert<double, W>(cotr, f(da), s.N, s.st());
    }

    // metric scalings of nf fields (launch_metric_scale modes)
    void metric(void *const *fields, int nf, int mode, int d) const
    {
        MetricFields<W> fs{};
        fs.nf = nf;
        for (int i = 0; i < nf; i++) fs.f[i] = f(fields[i]);
        s.timed(PH_OTHER, 2.0 * sizeof(W) * nf * (double)s.N,
                [&] { launch_metric_scale<W>(fs, s.met, mode, d, s.N, s.st()); });
    }

    // artificial.jl:98-111 smooth! in place, the last pass with epilogue `last`
    void smooth(void *field, EpiArgs last) const
    {
        for (int k = 0; k < s.nact; k++) {
            const int d = s.adim[k];
            // a paired fold never folds the last active dimension (its pairing
            // dimension comes after it), so the fused last pass stays plain
            if (k < s.nact - 1)
                line_par(FILT, d, field, 1, store(field), PH_SMOOTH);
            else
                line(plan(FILT, d, 1), d, field, last, PH_SMOOTH);
        }
    }

    D4Args<W> d4(int wpow) const
    {
        D4Args<W> a{};
        for (int k = 0; k < 3; k++) {
            a.n[k] = s.n[k];
            a.active[k] = s.active[k];
            a.periodic[k] = s.periodic[k];
            a.wd[k] = wpow == 2 ? s.h[k] * s.h[k] : s.h[k];
        }
        a.N = s.N;
        a.Q = s.Q;
        a.rho = f(s.rho);
        a.im1 = s.Ns;
        return a;
    }

    // artificial.jl:261-359 compute_artificial!
    void artificial() const
    {
        const cles_solver_desc &d = s.d;
        if (!d.art_enabled) return;
        if (d.beta_sensor == 2) {   // the dilatation sensor reuses `sensor` between the two chains
            art_mu_beta();
            art_kappa_d(s.sensor);
            return;
        }
        // the mu*/beta*, kappa* and D*_k sensors, then their smoothing passes
        // in shared launches (the last with each field's own epilogue)
        const double w = sizeof(W);
        s.timed(PH_STRAIN, w * (3 * s.nact + 1) * (double)s.N,
                [&] { launch_strain<W>(grads(), f(s.strain), s.N, s.st()); });
        void *sens[LINE_MF] = {s.sensor, s.sensorK};
        const void *src[LINE_MF] = {s.strain, s.esens};
        int nf = 2;
        if (s.Ns > 1)
            for (int sp = 0; sp < s.Ns; sp++) {
                sens[nf] = s.sensD[sp];
                src[nf++] = s.Y[sp];
            }
        // D4_MF fields per launch: strain with the h^2 weights, the others with h
        for (int i0 = 0; i0 < nf; i0 += D4_MF) {
            D4Args<W> a = d4(1);
            D4Multi<W> m{};
            m.nf = std::min(D4_MF, nf - i0);
            for (int i = 0; i < m.nf; i++) {
                m.f[i] = f((void *)src[i0 + i]);
                m.out[i] = f(sens[i0 + i]);
                m.wpow[i] = i0 + i == 0 ? 2 : 1;
            }
            s.timed(PH_DELTA4, 2 * w * m.nf * (double)s.N, [&] { launch_delta4_multi<W>(a, m, s.st()); });
        }
        for (int k = 0; k < s.nact; k++) {
            const int dd = s.adim[k];
            if (k < s.nact - 1) {
                int par[LINE_MF];
                for (int i = 0; i < nf; i++) par[i] = 1;
                lines_par(FILT, dd, nf, sens, sens, par, store(sens[0]), PH_SMOOTH);
            } else {
                // a paired fold never folds the last active dimension
                EpiArgs m{};
                m.kind = EPI_ART_MULTI;
                m.multi = s.art_multi;
                lines(plan(FILT, dd, 1), dd, nf, sens, nullptr, nullptr, m, PH_SMOOTH);
            }
        }
        if (d.beta_sensor == 1)
            s.timed(PH_OTHER, w * (3 * s.nact + 2) * (double)s.N,
                    [&] { launch_gate_beta<W>(grads(), f(s.beta_art), s.N, s.st()); });
    }
    // mu* and beta* (artificial.jl:267-310): needs the velocity gradients
    void art_mu_beta() const
    {
        const cles_solver_desc &d = s.d;
        const double w = sizeof(W);
        s.timed(PH_STRAIN, w * (3 * s.nact + 1) * (double)s.N,
                [&] { launch_strain<W>(grads(), f(s.strain), s.N, s.st()); });
        D4Args<W> a = d4(2);
        a.f = f(s.strain);
        s.timed(PH_DELTA4, 2 * w * (double)s.N, [&] { launch_delta4<W>(a, 0, f(s.sensor), s.st()); });
        EpiArgs mb{};
        mb.kind = EPI_ART_MUBETA;
        mb.out = s.mu_art;
        mb.out2 = s.beta_art;
        mb.a0 = s.rho;
        mb.c0 = d.C_mu;
        mb.c1 = d.C_beta;
        smooth(s.sensor, mb);
        if (d.beta_sensor == 1) {
            s.timed(PH_OTHER, w * (3 * s.nact + 2) * (double)s.N,
                    [&] { launch_gate_beta<W>(grads(), f(s.beta_art), s.N, s.st()); });
        } else if (d.beta_sensor == 2) {
            s.timed(PH_OTHER, w * (3 * s.nact + 2) * (double)s.N,
                    [&] { launch_dilatation_prep<W>(grads(), f(s.tmp1), f(s.tmp2), s.N, s.st()); });
            D4Args<W> b = d4(2);
            b.f = f(s.tmp1);
            s.timed(PH_DELTA4, 2 * w * (double)s.N, [&] { launch_delta4<W>(b, 0, f(s.sensor), s.st()); });
            smooth(s.sensor, store(s.sensor));
            s.timed(PH_OTHER, 4 * w * (double)s.N, [&] {
                launch_dilatation_beta<W>(f(s.rho), f(s.sensor), f(s.tmp2), f(s.beta_art), d.C_beta, s.N, s.st());
            });
        }
    }
    // kappa* and D*_k (artificial.jl:312-359): needs the primitives only;
    // `sensor` is the chain's scratch field
    void art_kappa_d(void *sensor) const
    {
        const cles_solver_desc &d = s.d;
        const double w = sizeof(W);
        // kappa* from the internal-energy sensor (e stored by the primitives pass)
        D4Args<W> ae = d4(1);
        ae.f = f(s.esens);
        s.timed(PH_DELTA4, 2 * w * (double)s.N, [&] { launch_delta4<W>(ae, 0, f(sensor), s.st()); });
        EpiArgs kp{};
        kp.kind = EPI_ART_KAPPA;
        kp.out = s.kappa_art;
        kp.a0 = s.rho;
        kp.a1 = s.c;
        kp.a2 = s.T;
        kp.c0 = d.C_kappa;
        smooth(sensor, kp);
        if (s.Ns > 1) {
            for (int sp = 0; sp < s.Ns; sp++) {
                D4Args<W> b = d4(1);
                b.f = f(s.Y[sp]);
                s.timed(PH_DELTA4, 2 * w * (double)s.N, [&] { launch_delta4<W>(b, 0, f(sensor), s.st()); });
                EpiArgs dd{};
                dd.kind = EPI_ART_D;
                dd.out = s.D_art[sp];
                dd.a1 = s.c;
                dd.c0 = d.C_D;
                smooth(sensor, dd);
            }
        }
    }

    void fluxes() const
    {
        FluxArgs<W> a{};
        a.Q = s.Q;
        a.Ns = s.Ns;
        a.n9 = s.n9;
        for (int k = 0; k < 3; k++) a.active[k] = s.active[k];
        a.mu0 = s.d.mu0;
        a.Pr = s.d.Pr;
        a.Sc = s.d.Sc;
        for (int k = 0; k < s.Ns; k++) {
            a.cpk[k] = s.cpk[k];
            a.Y[k] = f(s.Y[k]);
            a.D_art[k] = f(s.D_art[k]);
        }
        a.rho = f(s.rho); a.u = f(s.u); a.v = f(s.v); a.w = f(s.w);
        a.p = f(s.p); a.T = f(s.T); a.cp = f(s.cp);
        a.mu_art = f(s.mu_art);
        a.beta_art = f(s.beta_art);
        a.kappa_art = f(s.kappa_art);
        a.g = grads();
        for (int d = 0; d < 3; d++) {
            a.gT[d] = f(s.gT[d]);
            for (int k = 0; k < s.Ns; k++) a.gY[d][k] = f(s.gY[d][k]);
            for (int c = 0; c < s.C; c++) a.flux[d][c] = f(s.flux[d][c]);
        }
        a.R1 = s.Rk[0];
        a.cv1 = s.cvk[0];
        const bool cc = s.const_cp();
        if (cc) a.cp = a.Y[0] = nullptr;
        const double w = sizeof(W);
        const double reads = 8.0 + w * ((cc ? 9 : 10 + s.Ns * (s.Ns > 1 ? 2 + s.nact : 1)) + 3 * s.nact + s.nact);
        s.timed(PH_FLUXES, (reads + w * s.nact * s.C) * (double)s.N, [&] { launch_fluxes<W>(a, s.N, s.st()); });
    }

    W *dQc(int c) const { return f(s.dQ) + (size_t)c * s.N; }

    // rhs.jl:624-694 compute_rhs!; with fuse, the last flux-divergence pass of
    // each component performs the LSRK stage update (A, B, dt).
    void rhs(bool primitives_current, bool fuse, double A, double B, double dt) const
    {
        if (!primitives_current) primitives();
        // the temperature and species gradients (rhs.jl:608-622) are
        // independent of the artificial terms and go with the velocity ones
        velocity_gradients(true);
        artificial();
        fluxes();
        // rhs.jl:624-667 per direction, every component in one launch
        for (int k = 0; k < s.nact; k++) {
            const int d = s.adim[k];
            const bool last = k == s.nact - 1;
            const void *in[LINE_MF];
            void *out[LINE_MF];
            int64_t eoff[LINE_MF];
            for (int c = 0; c < s.C; c++) {
                in[c] = s.flux[d][c];
                out[c] = dQc(c);
                eoff[c] = (int64_t)c * s.N;
            }
            // rhs.jl:672-677: the divergence of A_d F_d (the fluxes are scaled in place)
            if (s.geom()) metric(s.flux[d], s.C, 1, d);
            EpiArgs e{};
            if (fuse && last) {
                e.kind = s.dev_dt ? EPI_LSRK_DEV : EPI_LSRK;
                e.a0 = s.nact > 1 ? (const void *)dQc(0) : nullptr;
                e.du = s.du;
                e.Q = s.Q;
                e.c0 = A;
                e.c1 = B;
                e.dt = dt;
                e.dtp = s.dev_dt;
                e.gate = s.dev_gate;
                e.ipts = s.ipts;
                lines(s.dplan[d], d, s.C, in, nullptr, eoff, e, PH_DIV_LSRK);
            } else {
                e.kind = k == 0 ? EPI_STORE_NEG : EPI_SUB;
                int par[LINE_MF];
                for (int c = 0; c < s.C; c++) par[c] = flux_parity(d, c);
                lines_par(DERIV, d, s.C, in, out, par, e, PH_DIV);
            }
        }
        if (s.geom()) {   // rhs.jl:680-682: dQ -= inv_J D_d(A_d F_d)
            void *dq[LINE_MF];
            for (int c = 0; c < s.C; c++) dq[c] = dQc(c);
            metric(dq, s.C, 2, 0);
        }
        if (s.curv) {   // rhs.jl:687 add_metric_sources!
            const CurvArgs<W> a = curv_args();
            s.timed(PH_OTHER, sizeof(W) * 20.0 * (double)s.N, [&] { launch_metric_sources<W>(a, s.st()); });
        }
    }

    // timestep.jl:125-175
    void max_rate(double *rate, double *rho_min) const
    {
        primitives();
        RateArgs<W> a{};
        a.Q = s.Q;
        a.Ns = s.Ns;
        for (int k = 0; k < 3; k++) {
            a.active[k] = s.active[k];
            a.invh[k] = 1.0 / s.h[k];
            a.h[k] = s.h[k];
        }
        a.met = s.met;
        a.mu0 = s.d.mu0;
        a.Pr = s.d.Pr;
        a.Sc = s.d.Sc;
        a.rho = f(s.rho); a.u = f(s.u); a.v = f(s.v); a.w = f(s.w); a.c = f(s.c); a.cp = f(s.cp);
        a.R1 = s.Rk[0];
        a.cv1 = s.cvk[0];
        if (s.const_cp()) a.cp = nullptr;
        a.mu_art = f(s.mu_art);
        a.beta_art = f(s.beta_art);
        a.kappa_art = f(s.kappa_art);
        for (int k = 0; k < s.Ns; k++) a.D_art[k] = f(s.D_art[k]);
        int nb = (int)std::min<int64_t>(s.nred, (s.N + 255) / 256);
        const double w = sizeof(W);
        s.timed(PH_RATE, (8.0 * s.Ns + w * ((s.const_cp() ? 8 : 9) + (s.Ns > 1 ? s.Ns : 0))) * (double)s.N,
                [&] { launch_rate<W>(a, s.part, nb, s.scal, s.N, s.st()); });
        CLES_HIP(hipMemcpyAsync(s.host_scal, s.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s.st()));
        CLES_HIP(hipStreamSynchronize(s.st()));
        *rate = s.host_scal[0];
        *rho_min = s.host_scal[1];
    }

    void lsrk_all(double A, double B, double dt) const
    {
        s.timed(PH_LSRK, (32.0 + sizeof(W)) * s.C * (double)s.N,
                [&] { launch_lsrk<W>(s.Q, s.du, f(s.dQ), A, B, dt, s.dev_dt, s.dev_gate, (int64_t)s.C * s.N, s.N,
                                     s.ipts, s.st()); });
    }

    // max_rate's kernels only: the rate and the minimum density stay in
    // s.scal[0..1] on the device
    void rate_device() const
    {
        primitives();
        RateArgs<W> a{};
        a.Q = s.Q;
        a.Ns = s.Ns;
        for (int k = 0; k < 3; k++) {
            a.active[k] = s.active[k];
            a.invh[k] = 1.0 / s.h[k];
            a.h[k] = s.h[k];
        }
        a.met = s.met;
        a.mu0 = s.d.mu0;
        a.Pr = s.d.Pr;
        a.Sc = s.d.Sc;
        a.rho = f(s.rho); a.u = f(s.u); a.v = f(s.v); a.w = f(s.w); a.c = f(s.c); a.cp = f(s.cp);
        a.R1 = s.Rk[0];
        a.cv1 = s.cvk[0];
        if (s.const_cp()) a.cp = nullptr;
        a.mu_art = f(s.mu_art);
        a.beta_art = f(s.beta_art);
        a.kappa_art = f(s.kappa_art);
        for (int k = 0; k < s.Ns; k++) a.D_art[k] = f(s.D_art[k]);
        if (s.ipts) {   // a stack: each instance's rate and minimum density
            const int nb = (int)std::min<int64_t>(s.nred, ((int64_t)s.ipts + 255) / 256);
            launch_rate_inst<W>(a, s.spart, nb, s.stack, s.srate, s.srmin, s.N, s.st());
            return;
        }
        int nb = (int)std::min<int64_t>(s.nred, (s.N + 255) / 256);
        launch_rate<W>(a, s.part, nb, s.scal, s.N, s.st());
    }

    // one run! iteration with every decision on the device (the captured graph)
    void iteration_device(const BatchCtl &c) const
    {
        apply_bcs();
        rate_device();
        launch_batch_dt(c, s.st());
        step(0.0, true);   // dt and the update gate come from the control block
        launch_batch_post(c, s.st());
        if (s.d.filter_interval > 0) filter_state();
    }

    // timestep.jl:72-99 step!
    void step(double dt, bool prepared) const
    {
        const bool fuse = !s.nscbc && !s.geom() && s.d.n_body_forces == 0;
        for (int stg = 0; stg < 5; stg++) {
            s.tstage = s.t + RKC[stg] * dt;
            const bool first_prepared = prepared && stg == 0;
            if (!first_prepared) apply_bcs();
            rhs(first_prepared, fuse, RKA[stg], RKB[stg], dt);
            if (!fuse) {
                nscbc_corrections();
                sources();
                lsrk_all(RKA[stg], RKB[stg], dt);
            }
        }
        s.tstage = s.t + dt;
        apply_bcs();
    }

    // rhs.jl:692 add_sources!: the body forces in order, on the assembled RHS
    void sources() const
    {
        for (int i = 0; i < s.d.n_body_forces; i++) {
            const double(&g)[3] = s.d.body_force[i];
            s.timed(PH_OTHER, 8.0 * (s.C + s.Ns) * (double)s.N,
                    [&] { launch_body_force<W>(s.Q, f(s.dQ), s.Ns, s.N, g, s.st()); });
        }
    }

    // rhs.jl:688-690: boundary corrections after the flux divergence, faces in
    // (dim, low, high) order; nscbc.jl:59-154 outflow, :220-306 inflow.
    // the full-field derivatives the NSCBC faces read: D_d p on every NSCBC
    // face's dimension d, D_d rho on inflow faces, the transverse D_t p on
    // outflow faces (nscbc.jl:69-87, :223-224)
    void nscbc_needs(bool (&np)[3], bool (&nr)[3]) const
    {
        for (int k = 0; k < 3; k++) np[k] = nr[k] = false;
        for (int d = 0; d < 3; d++)
            for (int side = 0; side < 2; side++) {
                if (!s.active[d]) continue;
                const int bc = s.d.bc[d][side];
                if (bc != CLES_BC_NSCBC_OUT && bc != CLES_BC_NSCBC_IN) continue;
                np[d] = true;
                if (bc == CLES_BC_NSCBC_IN) {
                    nr[d] = true;
                } else {
                    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                    np[t1] = np[t1] || s.active[t1];
                    np[t2] = np[t2] || s.active[t2];
                }
            }
    }

    void nscbc_corrections() const
    {
        if (!s.nscbc) return;
        void *vel[3] = {s.u, s.v, s.w};
        const double w = sizeof(W);
        for (int d = 0; d < 3; d++) {
            // with the derivatives from the gradient launches, the two faces
            // of a dimension go through one launch
            NscbcArgs<W> held[2];
            bool has[2] = {false, false};
            for (int side = 0; side < 2; side++) {
                if (!s.active[d]) continue;
                const int bc = s.d.bc[d][side];
                if (bc != CLES_BC_NSCBC_OUT && bc != CLES_BC_NSCBC_IN) continue;
                const cles_nscbc_desc &nd = s.d.nscbc[d][side];
                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                NscbcArgs<W> a{};
                a.dim = d;
                a.side = side;
                a.Ns = s.Ns;
                a.inflow = bc == CLES_BC_NSCBC_IN;
                for (int k = 0; k < 3; k++) a.n[k] = s.n[k];
                a.N = s.N;
                a.count = s.N / s.n[d];
                // full-field derivatives (nscbc.jl:69-87, :223-224): from the
                // gradient launches of this RHS, or computed here
                if (!a.inflow) {
                    a.act1 = s.active[t1];
                    a.act2 = s.active[t2];
                }
                if (s.nsd_merged) {
                    a.dpn = f(s.nsdp[d]);
                    a.dt1 = f(a.inflow ? s.nsdr[d] : s.nsdp[t1]);
                    a.dt2 = f(a.inflow ? nullptr : s.nsdp[t2]);
                } else {
                    line_par(DERIV, d, s.p, 1, store(s.nsd[0]), PH_NSCBC);
                    a.dpn = f(s.nsd[0]);
                    if (a.inflow) {
                        line_par(DERIV, d, s.rho, 1, store(s.nsd[1]), PH_NSCBC);
                        a.dt1 = f(s.nsd[1]);
                    } else {
                        if (a.act1) line_par(DERIV, t1, s.p, 1, store(s.nsd[1]), PH_NSCBC);
                        if (a.act2) line_par(DERIV, t2, s.p, 1, store(s.nsd[2]), PH_NSCBC);
                        a.dt1 = f(s.nsd[1]);
                        a.dt2 = f(s.nsd[2]);
                    }
                }
                a.rho = f(s.rho); a.c = f(s.c); a.p = f(s.p); a.T = f(s.T); a.cp = f(s.cp);
                for (int k = 0; k < 3; k++) a.vel[k] = f(vel[k]);
                for (int k = 0; k < 9; k++) a.gu[k] = f(s.gu[k]);
                for (int k = 0; k < s.Ns; k++) {
                    a.Y[k] = f(s.Y[k]);
                    a.gY[k] = f(s.gY[d][k]);   // NULL for one species: d(1)/dx = 0
                    a.Rk[k] = s.Rk[k];
                    a.cvk[k] = s.cvk[k];
                    a.Yt[k] = nd.Y[k];
                }
                a.Lref = nd.Lref > 0 ? nd.Lref : s.d.L_domain[d];
                a.pinf = nd.pinf;
                a.sigma = nd.sigma;
                a.beta_t = nd.beta_t;
                for (int k = 0; k < 3; k++) a.ut[k] = nd.u[k];
                a.Tt = nd.T;
                a.eta_u = nd.eta_u;
                a.eta_T = nd.eta_T;
                a.eta_t = nd.eta_t;
                a.eta_Y = nd.eta_Y;
                a.n9 = s.n9;
                a.sg_phi = s.d.eos_kind == CLES_EOS_STIFFENED ? 1 / (s.d.gamma[0] - 1) : 0.0;
                a.met = s.met;
                a.dQ = f(s.dQ);
                if (s.nsd_merged) {
                    held[side] = a;
                    has[side] = true;
                    continue;
                }
                s.timed(PH_NSCBC, (w * (14 + 2 * s.Ns) + 2 * w * s.C) * (double)a.count,
                        [&] { launch_nscbc<W>(a, s.st()); });
            }
            if (has[0] && has[1]) {
                s.timed(PH_NSCBC, 2 * (w * (14 + 2 * s.Ns) + 2 * w * s.C) * (double)held[0].count,
                        [&] { launch_nscbc_pair<W>(held[0], held[1], s.st()); });
            } else {
                for (int side = 0; side < 2; side++)
                    if (has[side])
                        s.timed(PH_NSCBC, (w * (14 + 2 * s.Ns) + 2 * w * s.C) * (double)held[side].count,
                                [&] { launch_nscbc<W>(held[side], s.st()); });
            }
        }
    }

    // boundary.jl:280-286 apply_bcs!
    void apply_bcs() const
    {
        for (int d = 0; d < 3; d++)
            for (int side = 0; side < 2; side++) {
                if (!s.active[d]) continue;
                const int bc = s.d.bc[d][side];
                if (bc != CLES_BC_SLIPWALL && bc != CLES_BC_DIRICHLET && bc != CLES_BC_NOSLIP && bc != CLES_BC_EXTRAP)
                    continue;
                PlaneArgs a{};
                a.dim = d;
                a.side = side;
                a.Ns = s.Ns;
                for (int k = 0; k < 3; k++) a.n[k] = s.n[k];
                a.N = s.N;
                a.count = s.N / s.n[d];
                for (int c = 0; c < s.C; c++) a.state[c] = s.d.bc_state[d][side][c];
                a.Twall = s.d.Twall[d][side];
                for (int k = 0; k < s.Ns; k++) a.cvk[k] = s.cvk[k];
                a.n9 = s.n9;
                a.sg = s.d.eos_kind == CLES_EOS_STIFFENED;
                a.sg_pinf = s.d.p_inf;
                s.timed(PH_BCS, 8.0 * s.C * (double)a.count, [&] {
                    if (bc == CLES_BC_SLIPWALL)
                        launch_slipwall(a, s.Q, s.st());
                    else if (bc == CLES_BC_DIRICHLET)
                        launch_dirichlet(a, s.Q, s.st());
                    else if (bc == CLES_BC_NOSLIP)
                        launch_noslip(a, s.Q, s.st());
                    else
                        launch_extrapolate(a, s.Q, s.st());
                });
            }
    }

    // timestep.jl:414-426 filter_state! (in place, dimension by dimension, FP64)
    void filter_state() const
    {
        for (int k = 0; k < s.nact; k++) {
            const int d = s.adim[k];
            const void *in[LINE_MF];
            void *out[LINE_MF];
            for (int c = 0; c < s.C; c++) {
                in[c] = s.Q + (size_t)c * s.N;
                out[c] = s.Q + (size_t)c * s.N;
            }
            EpiArgs e{};
            e.kind = s.dev_filt ? EPI_STORE_GATE : EPI_STORE;
            e.gate = s.dev_filt;
            e.ipts = s.ipts;
            int par[LINE_MF];
            for (int c = 0; c < s.C; c++) par[c] = cons_parity(d, c);
            lines_par(FILT_Q, d, s.C, in, out, par, e, PH_FILTER_STATE);
        }
    }

    void diagnostics(double *ke, double *ens) const
    {
        DiagArgs<W> a{};
        a.Q = s.Q;
        a.Ns = s.Ns;
        for (int k = 0; k < 3; k++) {
            a.n[k] = s.n[k];
            a.closed[k] = s.active[k] && !s.periodic[k];
            a.fold_lo[k] = s.fold[k].on && s.fold[k].lo;
            a.fold_hi[k] = s.fold[k].on && s.fold[k].hi;
        }
        a.N = s.N;
        a.met = s.met;
        double dV = 1.0;
        for (int k = 0; k < 3; k++) dV *= s.active[k] ? s.h[k] : 1.0;
        int nb = (int)std::min<int64_t>(s.nred, (s.N + 255) / 256);
        launch_integral<W>(a, 0, s.part, nb, dV, s.scal, s.st());
        primitives();
        velocity_gradients();
        a.g = grads();
        launch_integral<W>(a, 1, s.part, nb, dV, s.scal + 1, s.st());
        CLES_HIP(hipMemcpyAsync(s.host_scal, s.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s.st()));
        CLES_HIP(hipStreamSynchronize(s.st()));
        *ke = s.host_scal[0];
        *ens = s.host_scal[1];
    }
};

}  // namespace cles

namespace {

template <typename Fn>
void dispatch(cles_solver *s, Fn &&fn)
{
    if (s->prec == CLES_F64) {
        cles::Impl<double> im(*s);
        fn(im);
    } else {
        cles::Impl<float> im(*s);
        fn(im);
    }
}

// stepcontrol.jl:184-204 check_step
int check_step(const cles_solver_desc &d, double dt, double rho_min, double dt_seen, std::string &msg)
{
    char buf[256];
    if (!std::isfinite(dt)) { snprintf(buf, sizeof buf, "timestep is %g; the state has gone non-finite", dt); msg = buf; return CLES_ERR_NONFINITE; }
    if (!(rho_min > 0)) { snprintf(buf, sizeof buf, "minimum mixture density is %g; the state is unphysical", rho_min); msg = buf; return CLES_ERR_NEGATIVE_DENSITY; }
    if (!(dt > PLANCK_TIME)) { msg = "timestep is below the Planck time"; return CLES_ERR_PLANCK; }
    if (!(dt >= d.dt_min)) { msg = "timestep is below the configured dt_min"; return CLES_ERR_DT_MIN; }
    if (d.dt_min_ratio > 0 && dt_seen > 0 && !(dt >= d.dt_min_ratio * dt_seen)) {
        snprintf(buf, sizeof buf, "timestep has collapsed to %g of the largest seen this run", dt / dt_seen);
        msg = buf;
        return CLES_ERR_DT_COLLAPSE;
    }
    return CLES_OK;
}

}  // namespace

namespace cles {

void validate_eos(const cles_solver_desc &d)
{
    if (d.eos_kind != CLES_EOS_IDEAL && d.eos_kind != CLES_EOS_NASA9 && d.eos_kind != CLES_EOS_STIFFENED)
        throw cles::Error{CLES_ERR_INVALID, "eos_kind must be CLES_EOS_IDEAL, CLES_EOS_NASA9 or CLES_EOS_STIFFENED"};
    if (d.eos_kind == CLES_EOS_STIFFENED) {   // physics.jl:198-205
        if (d.n_species != 1) throw cles::Error{CLES_ERR_INVALID, "StiffenedGas is single-component"};
        if (!(d.gamma[0] > 1) || !(d.R[0] > 0) || !std::isfinite(d.p_inf))
            throw cles::Error{CLES_ERR_INVALID, "StiffenedGas needs gamma > 1, cv > 0 and a finite p_inf"};
    }
    if (d.eos_kind == CLES_EOS_NASA9) {   // physics.jl:312-338 _validate_nasa9_intervals
        if (!(d.T_guess > 0)) throw cles::Error{CLES_ERR_INVALID, "Nasa9Mixture: T_guess must be positive"};
        for (int k = 0; k < d.n_species; k++) {
            const cles_nasa9_species &sp = d.nasa9[k];
            if (sp.n_intervals < 1 || sp.n_intervals > CLES_NASA9_MAXINT)
                throw cles::Error{CLES_ERR_INVALID, "NASA-9 species " + std::to_string(k + 1) +
                                                        ": 1..3 intervals are supported"};
            if (!(d.R[k] > 0)) throw cles::Error{CLES_ERR_INVALID, "NASA-9 species has non-positive R"};
            for (int i = 0; i < sp.n_intervals; i++)
                if (!(sp.iv[i].Tmin > 0 && sp.iv[i].Tmin < sp.iv[i].Tmax))
                    throw cles::Error{CLES_ERR_INVALID, "invalid NASA-9 interval"};
        }
    }
}

N9Table nasa9_table(const cles_solver_desc &d)
{
    N9Table t{};
    t.ns = d.n_species;
    t.T_guess = d.T_guess;
    for (int k = 0; k < d.n_species; k++) {
        t.R[k] = d.R[k];
        t.sp[k] = d.nasa9[k];
    }
    return t;
}

}  // namespace cles

namespace {

void validate_desc(const cles_solver_desc &d)
{
    if (d.n_species < 1 || d.n_species > CLES_MAXSP) throw cles::Error{CLES_ERR_INVALID, "n_species must be 1..8"};
    for (int k = 0; k < 3; k++) {
        if (d.n_global[k] < 1) throw cles::Error{CLES_ERR_INVALID, "n_global must be >= 1"};
        if (!(d.L_domain[k] > 0)) throw cles::Error{CLES_ERR_INVALID, "domain extents must be positive"};
        const int lo = d.bc[k][0], hi = d.bc[k][1];
        if ((lo == CLES_BC_PERIODIC) != (hi == CLES_BC_PERIODIC))
            throw cles::Error{CLES_ERR_INVALID, "dimension mixes periodic and non-periodic conditions"};
        if (d.n_global[k] == 1 && lo != CLES_BC_PERIODIC)
            throw cles::Error{CLES_ERR_INVALID, "collapsed dimension must use (PeriodicBC(), PeriodicBC())"};
        for (int sd = 0; sd < 2; sd++)
            if (d.bc[k][sd] < 0 || d.bc[k][sd] > CLES_BC_EXTRAP)
                throw cles::Error{CLES_ERR_INVALID, "unknown boundary-condition code"};
        // boundary.jl:111-121 SwitchableBC: both arms agree on periodicity and
        // neither is a fold
        for (int sd = 0; sd < 2; sd++) {
            if (!d.switchable[k][sd]) continue;
            const int b0 = d.bc[k][sd], b1 = d.bc_after[k][sd];
            if (b1 < 0 || b1 > CLES_BC_EXTRAP) throw cles::Error{CLES_ERR_INVALID, "unknown boundary-condition code"};
            if ((b0 == CLES_BC_PERIODIC) != (b1 == CLES_BC_PERIODIC))
                throw cles::Error{CLES_ERR_INVALID, "SwitchableBC: both conditions must agree on periodicity"};
            for (int b : {b0, b1})
                if (b == CLES_BC_AXIS || b == CLES_BC_ORIGIN || b == CLES_BC_POLE)
                    throw cles::Error{CLES_ERR_INVALID, "SwitchableBC cannot wrap a fold condition (AxisBC, OriginBC, "
                                                        "PoleBC); setup detects those by type"};
        }
    }
    if (d.beta_sensor < 0 || d.beta_sensor > 2)
        throw cles::Error{CLES_ERR_INVALID, "art.beta_sensor must be :strain, :gated_strain or :dilatation"};
    if (d.precision != CLES_F64 && d.precision != CLES_F32)
        throw cles::Error{CLES_ERR_INVALID, "precision must be CLES_F64 or CLES_F32"};
    if (!(d.cfl > 0)) throw cles::Error{CLES_ERR_INVALID, "cfl must be positive"};
    // stepcontrol.jl:116-136
    if (d.retries < 0) throw cles::Error{CLES_ERR_INVALID, "StepControl: retries must be >= 0"};
    if (d.retries > 0 && !(d.cfl_backoff > 0))
        throw cles::Error{CLES_ERR_INVALID, "StepControl: cfl_backoff must be positive"};
    if (d.savepoint_interval < 0) throw cles::Error{CLES_ERR_INVALID, "StepControl: savepoint_interval must be >= 0"};
    cles::validate_eos(d);
    if (d.n_body_forces < 0 || d.n_body_forces > CLES_MAX_FORCES)
        throw cles::Error{CLES_ERR_INVALID, "n_body_forces must be 0..4"};
    if (d.metric < CLES_METRIC_CARTESIAN || d.metric > CLES_METRIC_SPHERICAL)
        throw cles::Error{CLES_ERR_INVALID, "metric must be CLES_METRIC_CARTESIAN, _CYLINDRICAL or _SPHERICAL"};
    if (d.landing_steps < 0)
        throw cles::Error{CLES_ERR_INVALID, "StepControl: landing_steps must be >= 1 (1 is a hard clip onto the "
                                            "scheduled time)"};
    if ((double)d.n_global[0] * (double)d.n_global[1] * (double)d.n_global[2] >= 2147483648.0)
        throw cles::Error{CLES_ERR_UNSUPPORTED, "grids of 2^31 points or more are not supported on one GPU"};
}

}  // namespace

extern "C" {

// io.jl:6-75 save_checkpoint / load_checkpoint! byte format, one rank:
// magic, n_cons, n_global[3], n_local[3], coords[3] (Int64), t (Float64),
// step (Int64), then the interior of Q as Float64 (x fastest, components
// last) -- the device layout of Q.
namespace {
constexpr uint64_t CKPT_MAGIC = 0x434c4553434b5054ULL;   // "CLES CKPT"

std::string ckpt_name(const char *prefix, int rank)
{
    char r[16];
    std::snprintf(r, sizeof r, ".r%04d.ckpt", rank);
    return std::string(prefix) + r;
}
}  // namespace

int cles_solver_save_checkpoint(cles_solver *s, const char *prefix, int rank)
{
    CLES_TRY({
        CLES_REQUIRE(s && prefix, "NULL argument");
        CLES_REQUIRE(rank >= 0, "rank must be >= 0");
        std::vector<double> buf((size_t)s->N * s->C);
        CLES_HIP(hipMemcpyAsync(buf.data(), s->Q, buf.size() * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
        const std::string name = ckpt_name(prefix, rank);
        FILE *fp = std::fopen(name.c_str(), "wb");
        if (!fp) throw cles::Error{CLES_ERR_INVALID, "cannot open " + name + " for writing"};
        int64_t hdr[10] = {s->C, s->n[0], s->n[1], s->n[2], s->n[0], s->n[1], s->n[2], 0, 0, 0};
        const int64_t step = s->step;
        bool ok = std::fwrite(&CKPT_MAGIC, 8, 1, fp) == 1 && std::fwrite(hdr, 8, 10, fp) == 10 &&
                  std::fwrite(&s->t, 8, 1, fp) == 1 && std::fwrite(&step, 8, 1, fp) == 1 &&
                  std::fwrite(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
        ok = (std::fclose(fp) == 0) && ok;
        if (!ok) throw cles::Error{CLES_ERR_INVALID, "short write to " + name};
    })
}

int cles_solver_load_checkpoint(cles_solver *s, const char *prefix, int rank)
{
    CLES_TRY({
        CLES_REQUIRE(s && prefix, "NULL argument");
        const std::string name = ckpt_name(prefix, rank);
        FILE *fp = std::fopen(name.c_str(), "rb");
        if (!fp) throw cles::Error{CLES_ERR_INVALID, "cannot open " + name};
        struct Closer {
            FILE *f;
            ~Closer() { std::fclose(f); }
        } closer{fp};
        uint64_t magic = 0;
        int64_t hdr[10];
        double t = 0;
        int64_t step = 0;
        if (std::fread(&magic, 8, 1, fp) != 1 || magic != CKPT_MAGIC)
            throw cles::Error{CLES_ERR_INVALID, "not a CompactLES checkpoint"};
        if (std::fread(hdr, 8, 10, fp) != 10) throw cles::Error{CLES_ERR_INVALID, "truncated checkpoint header"};
        if (hdr[0] != s->C) throw cles::Error{CLES_ERR_INVALID, "conserved layout mismatch"};
        for (int d = 0; d < 3; d++)
            if (hdr[1 + d] != s->n[d]) throw cles::Error{CLES_ERR_INVALID, "global grid mismatch"};
        for (int d = 0; d < 3; d++)
            if (hdr[4 + d] != s->n[d]) throw cles::Error{CLES_ERR_INVALID, "decomposition mismatch"};
        for (int d = 0; d < 3; d++)
            if (hdr[7 + d] != 0) throw cles::Error{CLES_ERR_INVALID, "rank layout mismatch"};
        if (std::fread(&t, 8, 1, fp) != 1 || std::fread(&step, 8, 1, fp) != 1)
            throw cles::Error{CLES_ERR_INVALID, "truncated checkpoint header"};
        std::vector<double> buf((size_t)s->N * s->C);
        if (std::fread(buf.data(), sizeof(double), buf.size(), fp) != buf.size())
            throw cles::Error{CLES_ERR_INVALID, "truncated checkpoint data"};
        CLES_HIP(hipMemcpyAsync(s->Q, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
        s->t = t;
        s->step = step;
    })
}

int cles_solver_footprint(const cles_solver_desc *desc, size_t *bytes)
{
    CLES_TRY({
        CLES_REQUIRE(desc && bytes, "NULL argument");
        validate_desc(*desc);
        const cles_solver_desc &d = *desc;
        const int Ns = d.n_species, C = Ns + 4;
        const double w = d.precision == CLES_F64 ? 8.0 : 4.0;
        double N = 1.0, plans = 0.0;
        int nact = 0;
        bool nscbc = false;
        for (int k = 0; k < 3; k++) {
            N *= (double)d.n_global[k];
            if (d.n_global[k] > 1) {
                nact++;
                // three line plans at most per dimension: row tables (<= 12 values of
                // 8 + 4 bytes per node), PCR blocks and 256-byte-aligned parts
                plans += 3.0 * ((double)d.n_global[k] * 12.0 * 12.0 + 64.0 * 1024.0);
                for (int sd = 0; sd < 2; sd++)
                    nscbc = nscbc || d.bc[k][sd] == CLES_BC_NSCBC_IN || d.bc[k][sd] == CLES_BC_NSCBC_OUT;
            }
        }
        // the allocation table of cles_solver_create
        double fields = 8 + Ns;                                   // primitives, Y
        fields += nact * (3 + 1 + (Ns > 1 ? Ns : 0) + C);         // grad_u, grad_T, grad_Y, fluxes
        fields += 3 + (Ns > 1 ? Ns : 0);                          // mu*, beta*, kappa*, D*
        if (d.art_enabled) fields += 4 + (Ns > 1 ? Ns : 0) + (d.beta_sensor == 2 ? 2 : 0);
        for (int k = 0; k < 3; k++)   // a SwitchableBC arm that may run NSCBC holds the scratch too
            for (int sd = 0; sd < 2; sd++)
                nscbc = nscbc || (d.switchable[k][sd] && d.n_global[k] > 1 &&
                                  (d.bc_after[k][sd] == CLES_BC_NSCBC_IN || d.bc_after[k][sd] == CLES_BC_NSCBC_OUT));
        if (nscbc) fields += std::max(3, 2 * nact);   // NSCBC derivatives
        {   // pairing scratch of a paired fold (folds.jl:276-330): three 8-byte fields
            const bool paired = (d.bc[0][0] == CLES_BC_AXIS && d.n_global[1] > 1) ||
                                (d.bc[0][0] == CLES_BC_ORIGIN && (d.n_global[1] > 1 || d.n_global[2] > 1)) ||
                                (d.bc[1][0] == CLES_BC_POLE && d.n_global[2] > 1);
            if (paired) fields += 3.0 * 8.0 / w;
        }
        if (d.metric != CLES_METRIC_CARTESIAN) {   // grad_u rows of collapsed angular dims, GCL cot/r
            fields += 3.0 * ((d.n_global[1] == 1) + (d.metric == CLES_METRIC_SPHERICAL && d.n_global[2] == 1));
            if (d.metric == CLES_METRIC_SPHERICAL && d.n_global[1] > 1) fields += 8.0 / w;
        }
        const double total = N * C * 16.0 + N * C * w + N * w * fields + 1024.0 * 2 * 8 + 16 * 8 + plans;
        *bytes = (size_t)total;
    })
}

int cles_solver_memory(cles_solver *s, size_t *bytes)
{
    CLES_TRY({
        CLES_REQUIRE(s && bytes, "NULL argument");
        size_t b = s->bytes;
        for (int k = 0; k < 3; k++) {
            b += s->dplan[k].dbytes + s->fplan[k].dbytes;
            if (s->own_fplanQ) b += s->fplanQ[k].dbytes;
        }
        for (int g = 0; g < 6; g++) {
            const cles_solver::Fold &fo = s->fold[g / 2];
            b += fo.dp[g % 2].dbytes + fo.fp[g % 2].dbytes;
            if (s->own_fplanQ) b += fo.fpQ[g % 2].dbytes;
        }
        *bytes = b;
    })
}

int cles_solver_create(cles_ctx *ctx, const cles_solver_desc *desc, cles_solver **out)
{
    return cles_solver_create_ex(ctx, desc, nullptr, out);
}

int cles_solver_create_ex(cles_ctx *ctx, const cles_solver_desc *desc, const double *const dxdxi[3],
                          cles_solver **out)
{
    return cles_solver_create_geom(ctx, desc, nullptr, dxdxi, out);
}

}  // extern "C"

namespace {

// cles_solver_create_geom (stack 0), or the solver of a cles_stack of
// `stack` instances
void create_solver(cles_ctx *ctx, const cles_solver_desc *desc, const double *const x[3],
                   const double *const dxdxi[3], int stack, cles_solver **out)
{
    {
        CLES_REQUIRE(ctx && desc && out, "NULL argument");
        validate_desc(*desc);
        for (int k = 0; k < 3; k++) {
            if (!dxdxi || !dxdxi[k]) continue;
            // rhs.jl:186-187
            CLES_REQUIRE(desc->n_global[k] > 1 && desc->bc[k][0] != CLES_BC_PERIODIC,
                         "dimension " + std::to_string(k + 1) + ": stretched dimensions must be non-periodic");
            for (int64_t i = 0; i < desc->n_global[k]; i++)
                CLES_REQUIRE(dxdxi[k][i] > 0 && std::isfinite(dxdxi[k][i]),
                             "stretch mapping for dim " + std::to_string(k + 1) + " is not monotone increasing");
        }
        const int kind = desc->metric;
        // rhs.jl:143-190: coordinate-singularity folds recognised from the BCs
        for (int k = 0; k < 3; k++)
            for (int sd = 0; sd < 2; sd++) {
                const int bc = desc->bc[k][sd];
                if (bc == CLES_BC_AXIS && !(k == 0 && sd == 0))
                    throw cles::Error{CLES_ERR_INVALID, "AxisBC goes at the low end of dimension 1"};
                if (bc == CLES_BC_ORIGIN && !(k == 0 && sd == 0))
                    throw cles::Error{CLES_ERR_INVALID, "OriginBC goes at the low end of dimension 1"};
                if (bc == CLES_BC_POLE && k != 1)
                    throw cles::Error{CLES_ERR_INVALID, "PoleBC goes on dimension 2 (theta)"};
            }
        const bool axis = desc->bc[0][0] == CLES_BC_AXIS, orig = desc->bc[0][0] == CLES_BC_ORIGIN;
        const bool pole_l = desc->bc[1][0] == CLES_BC_POLE, pole_h = desc->bc[1][1] == CLES_BC_POLE;
        CLES_REQUIRE(pole_l == pole_h, "PoleBC must be applied at both ends of theta");
        const int64_t *ng = desc->n_global;
        const double *Ld = desc->L_domain;
        auto even_period = [&](int k) { return ng[k] == 1 || (ng[k] % 2 == 0 && std::fabs(Ld[k] - 2 * M_PI) <= 1e-10); };
        if (axis) {
            CLES_REQUIRE(kind == CLES_METRIC_CYLINDRICAL, "AxisBC requires CylindricalMetric");
            CLES_REQUIRE(!(dxdxi && dxdxi[0]), "folded dimensions cannot be stretched");
            CLES_REQUIRE(std::fabs(desc->origin[0]) < 1e-14, "AxisBC requires origin[1] = 0");
            CLES_REQUIRE(even_period(1), "resolved-theta AxisBC requires an even theta point count over 2 pi");
        }
        if (orig) {
            CLES_REQUIRE(kind == CLES_METRIC_SPHERICAL, "OriginBC requires SphericalMetric");
            CLES_REQUIRE(!(dxdxi && dxdxi[0]), "folded dimensions cannot be stretched");
            CLES_REQUIRE(std::fabs(desc->origin[0]) < 1e-14, "OriginBC requires origin[1] = 0");
            CLES_REQUIRE(even_period(2), "resolved-phi OriginBC requires an even phi point count over 2 pi");
            CLES_REQUIRE(ng[1] == 1 || std::fabs(2 * desc->origin[1] + Ld[1] - M_PI) <= 1e-10,
                         "OriginBC requires a theta range symmetric about pi/2");
            CLES_REQUIRE(!(ng[1] > 1 && ng[2] == 1 && ng[1] % 2),
                         "OriginBC with a collapsed phi needs an even theta point count");
        }
        if (pole_l) {
            CLES_REQUIRE(kind == CLES_METRIC_SPHERICAL, "PoleBC requires SphericalMetric");
            CLES_REQUIRE(!(dxdxi && dxdxi[1]), "folded dimensions cannot be stretched");
            CLES_REQUIRE(std::fabs(desc->origin[1]) < 1e-14 && std::fabs(Ld[1] - M_PI) <= 1e-10,
                         "PoleBC requires the theta domain (0, pi)");
            CLES_REQUIRE(even_period(2), "resolved-phi PoleBC requires an even phi point count over 2 pi");
        }
        // nscbc.jl:309-316 _validate_nscbc_face: the LODI analysis needs a face
        // whose normal scale factor is one (metric.jl:93-95 unit_scalefactor)
        for (int k = 0; k < 3; k++) {
            const bool unit = kind == CLES_METRIC_CARTESIAN || (kind == CLES_METRIC_CYLINDRICAL && k != 1) ||
                              (kind == CLES_METRIC_SPHERICAL && k == 0);
            for (int sd = 0; sd < 2 * 2; sd++) {   // both arms of a SwitchableBC (boundary.jl:199-201)
                if (sd >= 2 && !desc->switchable[k][sd - 2]) continue;
                const int bc = sd < 2 ? desc->bc[k][sd] : desc->bc_after[k][sd - 2];
                if ((bc == CLES_BC_NSCBC_IN || bc == CLES_BC_NSCBC_OUT) && desc->n_global[k] > 1 && !unit)
                    throw cles::Error{CLES_ERR_INVALID,
                                      std::string(bc == CLES_BC_NSCBC_IN ? "NSCBCInflowBC" : "NSCBCOutflowBC") +
                                          " on dimension " + std::to_string(k + 1) +
                                          ": the LODI formulation requires a face whose normal metric scale "
                                          "factor is one (a Cartesian face, or a cylindrical r/z or spherical r "
                                          "face)"};
            }
        }
        // the metric reads r (dimension 1) and, spherical, theta (dimension 2)
        auto reads = [&](int k) {
            return (kind != CLES_METRIC_CARTESIAN && k == 0) || (kind == CLES_METRIC_SPHERICAL && k == 1);
        };
        for (int k = 0; k < 3; k++)
            if (reads(k) && dxdxi && dxdxi[k] && !(x && x[k]))
                throw cles::Error{CLES_ERR_INVALID, "dimension " + std::to_string(k + 1) +
                                                        ": a stretched curvilinear coordinate needs its node "
                                                        "coordinates x"};
        CLES_HIP(hipSetDevice(ctx->device));
        auto *s = new cles_solver;
        try {
            s->ctx = ctx;
            cles::ctx_retain(ctx);
            s->d = *desc;
            s->Ns = desc->n_species;
            s->C = s->Ns + 4;
            s->prec = desc->precision;
            s->N = 1;
            for (int k = 0; k < 3; k++) {
                s->n[k] = desc->n_global[k];
                s->N *= s->n[k];
                s->active[k] = s->n[k] > 1;
                s->periodic[k] = s->active[k] ? desc->bc[k][0] == CLES_BC_PERIODIC : 1;
                // rhs.jl:198-204
                s->h[k] = !s->active[k] ? 1.0
                          : s->periodic[k] ? desc->L_domain[k] / (double)desc->n_global[k]
                                           : desc->L_domain[k] / (double)(desc->n_global[k] - 1);
                if (s->active[k]) s->adim[s->nact++] = k;
                for (int sd = 0; sd < 2; sd++) {
                    auto is_ns = [](int b) { return b == CLES_BC_NSCBC_IN || b == CLES_BC_NSCBC_OUT; };
                    if (is_ns(desc->bc[k][sd])) s->nscbc = s->active[k] || s->nscbc;
                    if (desc->switchable[k][sd] && is_ns(desc->bc_after[k][sd]))
                        s->nscbc_any = s->active[k] || s->nscbc_any;
                }
                if (dxdxi && dxdxi[k]) {   // rhs.jl:191-201: computational spacing on [0, 1]
                    s->h[k] = 1.0 / (double)(desc->n_global[k] - 1);
                    s->stretched = true;
                }
            }
            if (stack > 0) {   // the instances as planes along the inactive z
                CLES_REQUIRE(desc->n_global[2] == 1 && s->nact >= 1, "a stack holds two-dimensional problems");
                CLES_REQUIRE(desc->metric == CLES_METRIC_CARTESIAN && !(x || dxdxi) && desc->n_body_forces == 0,
                             "a stack holds Cartesian problems on uniform grids without body forces");
                CLES_REQUIRE(s->N * stack < (int64_t)1 << 31, "a stack holds fewer than 2^31 points");
                s->stack = stack;
                s->ipts = (unsigned)s->N;
                s->n[2] = stack;
                s->N *= stack;
            }
            // rhs.jl:256-279 foldspec: folded ends, pairing, sigvel, area parity
            auto set_fold = [&](int k, bool hi, int pdim, int revdim, int s0, int s1, int s2, int area) {
                cles_solver::Fold &fo = s->fold[k];
                fo.on = true;
                fo.lo = true;
                fo.hi = hi;
                fo.pdim = pdim >= 0 && s->active[pdim] ? pdim : -1;
                fo.revdim = revdim >= 0 && s->active[revdim] ? revdim : -1;
                fo.sv[0] = s0; fo.sv[1] = s1; fo.sv[2] = s2;
                fo.area = area;
            };
            if (axis && s->active[0]) set_fold(0, false, 1, -1, -1, -1, 1, -1);   // A_1 = r
            if (orig && s->active[0]) set_fold(0, false, 2, 1, -1, 1, -1, 1);    // A_1 = r^2 sin(theta)
            if (pole_l && s->active[1]) set_fold(1, true, 2, -1, 1, -1, -1, -1); // A_2 = r sin(theta)
            for (int k = 0; k < 3; k++)   // rhs.jl:196-204: half-offset grids
                if (s->fold[k].on)
                    s->h[k] = s->fold[k].hi ? Ld[k] / (double)ng[k] : Ld[k] / ((double)ng[k] - 0.5);
            s->curv = kind != CLES_METRIC_CARTESIAN;
            for (int k = 0; k < s->Ns; k++) {   // physics.jl:111-115
                s->Rk[k] = desc->R[k];
                s->cvk[k] = desc->eos_kind == CLES_EOS_NASA9 ? 0.0 : desc->R[k] / (desc->gamma[k] - 1);
                s->cpk[k] = s->Rk[k] + s->cvk[k];
            }
            const int wdt = s->prec;
            auto chunk_of = [&](int k) { return cles::line_chunk(k + 1, *desc); };
            for (int k = 0; k < 3; k++) {
                if (!s->active[k]) continue;
                for (cles::HostPlan *hp : {&s->dplan[k], &s->fplan[k]}) {
                    hp->dim = k + 1;
                    hp->dtype = wdt;
                    for (int j = 0; j < 3; j++) hp->nloc[j] = s->n[j];
                }
                cles::build_line_plan(s->dplan[k], desc->deriv, (int)s->n[k], s->periodic[k], s->h[k], chunk_of(k));
                cles::build_line_plan(s->fplan[k], desc->filt, (int)s->n[k], s->periodic[k], s->h[k], chunk_of(k));
            }
            for (int k = 0; k < 3; k++) {   // rhs.jl:256-261: plans for sigma_g = +1, -1
                cles_solver::Fold &fo = s->fold[k];
                if (!fo.on) continue;
                for (int g = 0; g < 2; g++) {
                    for (cles::HostPlan *hp : {&fo.dp[g], &fo.fp[g], &fo.fpQ[g]}) {
                        hp->dim = k + 1;
                        hp->dtype = wdt;
                        for (int j = 0; j < 3; j++) hp->nloc[j] = s->n[j];
                    }
                    const int sg = g == 0 ? 1 : -1, lo = sg, hi = fo.hi ? sg : 0;
                    cles::build_line_plan(fo.dp[g], desc->deriv, (int)s->n[k], 0, s->h[k], chunk_of(k), lo, hi);
                    cles::build_line_plan(fo.fp[g], desc->filt, (int)s->n[k], 0, s->h[k], chunk_of(k), lo, hi);
                    if (s->prec == CLES_F64) {
                        fo.fpQ[g] = fo.fp[g];
                    } else {
                        fo.fpQ[g].dtype = CLES_F64;
                        cles::build_line_plan(fo.fpQ[g], desc->filt, (int)s->n[k], 0, s->h[k], chunk_of(k), lo, hi);
                    }
                }
            }
            if (s->prec == CLES_F64) {
                for (int k = 0; k < 3; k++) s->fplanQ[k] = s->fplan[k];
            } else {
                s->own_fplanQ = true;
                for (int k = 0; k < 3; k++) {
                    if (!s->active[k]) continue;
                    s->fplanQ[k].dim = k + 1;
                    s->fplanQ[k].dtype = CLES_F64;
                    for (int j = 0; j < 3; j++) s->fplanQ[k].nloc[j] = s->n[j];
                    cles::build_line_plan(s->fplanQ[k], desc->filt, (int)s->n[k], s->periodic[k], s->h[k], chunk_of(k));
                }
            }
            const size_t N = (size_t)s->N;
            s->Q = (double *)s->alloc(N * s->C * sizeof(double));
            s->du = (double *)s->alloc(N * s->C * sizeof(double));
            s->dQ = s->alloc(N * s->C * s->wsize());
            if (s->fold[0].paired() || s->fold[1].paired() || s->fold[2].paired())
                for (void **fp : {&s->pairw, &s->paira, &s->pairb}) *fp = s->alloc(N * sizeof(double));
            for (void **fp : {&s->rho, &s->u, &s->v, &s->w, &s->p, &s->T, &s->c, &s->cp}) *fp = s->wfield();
            for (int k = 0; k < s->Ns; k++) s->Y[k] = s->wfield();
            for (int dd = 0; dd < 3; dd++) {
                // a collapsed angular dimension still carries the curvature
                // terms of grad_u (metric.jl:181-219)
                if (!s->active[dd] && s->curv && (dd == 1 || (dd == 2 && kind == CLES_METRIC_SPHERICAL)))
                    for (int j = 0; j < 3; j++) s->gu[dd * 3 + j] = s->wfield();
                if (!s->active[dd]) continue;
                for (int j = 0; j < 3; j++) s->gu[dd * 3 + j] = s->wfield();
                s->gT[dd] = s->wfield();
                if (s->Ns > 1)
                    for (int k = 0; k < s->Ns; k++) s->gY[dd][k] = s->wfield();
                for (int c = 0; c < s->C; c++) s->flux[dd][c] = s->wfield();
            }
            // artificial coefficients stay zero when disabled (artificial.jl:263-265)
            s->mu_art = s->wfield();
            s->beta_art = s->wfield();
            s->kappa_art = s->wfield();
            if (s->Ns > 1)
                for (int k = 0; k < s->Ns; k++) s->D_art[k] = s->wfield();
            if (desc->art_enabled) {
                s->sensor = s->wfield();
                s->sensorK = s->wfield();
                if (s->Ns > 1)
                    for (int k = 0; k < s->Ns; k++) s->sensD[k] = s->wfield();
                s->strain = s->wfield();
                s->esens = s->wfield();
                if (desc->beta_sensor == 2) {
                    s->tmp1 = s->wfield();
                    s->tmp2 = s->wfield();
                }
            }
            s->nscbc_any = s->nscbc_any || s->nscbc;
            if (s->nscbc_any) {
                // Cartesian uniform grids (no metric scaling of the gradients)
                // compute the NSCBC derivatives in the gradient launches
                s->nsd_merged = !s->geom() && 4 + (s->Ns > 1 ? s->Ns : 0) + 2 <= cles::LINE_MF;
                if (s->nsd_merged) {
                    for (int k = 0; k < 3; k++)
                        if (s->active[k]) {
                            s->nsdp[k] = s->wfield();
                            s->nsdr[k] = s->wfield();
                        }
                } else {
                    for (void *&fp : s->nsd) fp = s->wfield();
                }
            }
            if (desc->eos_kind == CLES_EOS_NASA9) {
                const cles::N9Table t = cles::nasa9_table(*desc);
                auto *dt = (cles::N9Table *)s->alloc(sizeof(cles::N9Table));
                s->upload(dt, &t, sizeof t);
                s->n9 = dt;
            }
            for (int k = 0; k < 3; k++) {
                s->met.n[k] = s->n[k];
                if (!(dxdxi && dxdxi[k])) continue;
                double *m = (double *)s->alloc((size_t)s->n[k] * sizeof(double));
                s->upload(m, dxdxi[k], (size_t)s->n[k] * sizeof(double));
                s->met.m[k] = m;
            }
            s->met.kind = kind;
            if (s->curv) {
                // metric.jl:94-103: node coordinates origin + i h, or the mapping's
                auto node = [&](int k, int64_t i) {
                    const double shift = s->fold[k].on ? s->h[k] / 2 : 0.0;   // rhs.jl:205 coord_shift
                    return (x && x[k]) ? x[k][i] : desc->origin[k] + shift + (double)i * s->h[k];
                };
                std::vector<double> r((size_t)s->n[0]);
                for (int64_t i = 0; i < s->n[0]; i++) {
                    r[i] = node(0, i);
                    CLES_REQUIRE(r[i] > 0 && std::isfinite(r[i]),
                                 "curvilinear metric: every node needs r > 0 (coordinate singularities are not "
                                 "treated on the GPU backend)");
                }
                double *rd = (double *)s->alloc(r.size() * sizeof(double));
                s->upload(rd, r.data(), r.size() * sizeof(double));
                s->met.r = rd;
                if (kind == CLES_METRIC_SPHERICAL) {
                    std::vector<double> sc(2 * (size_t)s->n[1]);
                    for (int64_t j = 0; j < s->n[1]; j++) {
                        const double th = node(1, j);
                        CLES_REQUIRE(th > 0 && th < M_PI,
                                     "SphericalMetric: every node needs 0 < theta < pi (the poles are not treated "
                                     "on the GPU backend)");
                        sc[j] = std::sin(th);
                        sc[s->n[1] + j] = std::cos(th);
                    }
                    double *sd = (double *)s->alloc(sc.size() * sizeof(double));
                    s->upload(sd, sc.data(), sc.size() * sizeof(double));
                    s->met.sth = sd;
                    s->met.cth = sd + s->n[1];
                }
            }
            if (desc->art_enabled) {
                std::vector<cles::EpiArgs> ep(2 + (s->Ns > 1 ? s->Ns : 0));
                ep[0].kind = cles::EPI_ART_MUBETA;
                ep[0].out = s->mu_art;
                ep[0].out2 = s->beta_art;
                ep[0].a0 = s->rho;
                ep[0].c0 = desc->C_mu;
                ep[0].c1 = desc->C_beta;
                ep[1].kind = cles::EPI_ART_KAPPA;
                ep[1].out = s->kappa_art;
                ep[1].a0 = s->rho;
                ep[1].a1 = s->c;
                ep[1].a2 = s->T;
                ep[1].c0 = desc->C_kappa;
                for (size_t k = 2; k < ep.size(); k++) {
                    ep[k].kind = cles::EPI_ART_D;
                    ep[k].out = s->D_art[k - 2];
                    ep[k].a1 = s->c;
                    ep[k].c0 = desc->C_D;
                }
                s->art_multi = (cles::EpiArgs *)s->alloc(ep.size() * sizeof(cles::EpiArgs));
                s->upload(s->art_multi, ep.data(), ep.size() * sizeof(cles::EpiArgs));
            }
            s->part = (double *)s->alloc((size_t)s->nred * 2 * sizeof(double));
            s->scal = (double *)s->alloc(16 * sizeof(double));
            CLES_HIP(hipHostMalloc((void **)&s->host_scal, 16 * sizeof(double), hipHostMallocDefault));
            s->ctl = (cles_solver::Ctl *)s->alloc(sizeof(cles_solver::Ctl));
            CLES_HIP(hipHostMalloc((void **)&s->host_ctl, sizeof(cles_solver::Ctl), hipHostMallocDefault));
            if (kind == CLES_METRIC_SPHERICAL && s->active[1]) {
                double *cotr = (double *)s->alloc((size_t)s->N * sizeof(double));
                dispatch(s, [&](auto &im) { im.gcl_cotr(cotr); });
                s->met.cotr = cotr;
            }
            if (stack > 0) {   // per-instance rates and rate partials of a stack
                s->srate = (double *)s->alloc((size_t)stack * sizeof(double));
                s->srmin = (double *)s->alloc((size_t)stack * sizeof(double));
                s->spart = (double *)s->alloc((size_t)stack * s->nred * 2 * sizeof(double));
            }
            CLES_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    }
}

}  // namespace

extern "C" {

int cles_solver_create_geom(cles_ctx *ctx, const cles_solver_desc *desc, const double *const x[3],
                            const double *const dxdxi[3], cles_solver **out)
{
    CLES_TRY({ create_solver(ctx, desc, x, dxdxi, 0, out); })
}

int cles_solver_destroy(cles_solver *s)
{
    CLES_TRY({ delete s; })
}

int cles_solver_shape(cles_solver *s, int64_t n[3], int *n_cons)
{
    CLES_TRY({
        CLES_REQUIRE(s && n && n_cons, "NULL argument");
        for (int k = 0; k < 3; k++) n[k] = s->n[k];
        *n_cons = s->C;
    })
}

static void check_layout(const cles_solver *s, const int64_t padded[3], const int halo[3])
{
    for (int k = 0; k < 3; k++)
        if (halo[k] < 0 || padded[k] < s->n[k] + 2 * (int64_t)halo[k])
            throw cles::Error{CLES_ERR_INVALID, "padded extents smaller than n + 2*halo"};
}

int cles_solver_upload(cles_solver *s, const double *Qh, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(s && Qh && padded && halo, "NULL argument");
        check_layout(s, padded, halo);
        const int64_t P0 = padded[0], P1 = padded[1], NP = padded[0] * padded[1] * padded[2];
        std::vector<double> buf((size_t)s->N * s->C);
        for (int c = 0; c < s->C; c++)
            for (int64_t k = 0; k < s->n[2]; k++)
                for (int64_t j = 0; j < s->n[1]; j++)
                    std::memcpy(&buf[(size_t)c * s->N + (size_t)s->n[0] * (j + s->n[1] * k)],
                                Qh + c * NP + halo[0] + P0 * ((j + halo[1]) + P1 * (k + halo[2])),
                                sizeof(double) * s->n[0]);
        CLES_HIP(hipMemcpyAsync(s->Q, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

int cles_solver_download(cles_solver *s, double *Qh, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(s && Qh && padded && halo, "NULL argument");
        check_layout(s, padded, halo);
        std::vector<double> buf((size_t)s->N * s->C);
        CLES_HIP(hipMemcpyAsync(buf.data(), s->Q, buf.size() * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
        const int64_t P0 = padded[0], P1 = padded[1], NP = padded[0] * padded[1] * padded[2];
        for (int c = 0; c < s->C; c++)
            for (int64_t k = 0; k < s->n[2]; k++)
                for (int64_t j = 0; j < s->n[1]; j++)
                    std::memcpy(Qh + c * NP + halo[0] + P0 * ((j + halo[1]) + P1 * (k + halo[2])),
                                &buf[(size_t)c * s->N + (size_t)s->n[0] * (j + s->n[1] * k)],
                                sizeof(double) * s->n[0]);
    })
}

int cles_solver_get_field(cles_solver *s, const char *name, double *host)
{
    CLES_TRY({
        CLES_REQUIRE(s && name && host, "NULL argument");
        std::string nm(name);
        void *src = nullptr;
        bool is_double = false;
        int a = 0, b = 0;
        if (nm == "rho") src = s->rho;
        else if (nm == "u") src = s->u;
        else if (nm == "v") src = s->v;
        else if (nm == "w") src = s->w;
        else if (nm == "p") src = s->p;
        else if (nm == "T") src = s->T;
        else if (nm == "c") src = s->c;
        else if (nm == "cp") src = s->cp;
        else if (nm == "mu_art") src = s->mu_art;
        else if (nm == "beta_art") src = s->beta_art;
        else if (nm == "kappa_art") src = s->kappa_art;
        else if (std::sscanf(name, "dQ[%d]", &a) == 1 && a >= 1 && a <= s->C)
            src = (char *)s->dQ + (size_t)(a - 1) * s->N * s->wsize();
        else if (std::sscanf(name, "Q[%d]", &a) == 1 && a >= 1 && a <= s->C) {
            src = s->Q + (size_t)(a - 1) * s->N;
            is_double = true;
        } else if (std::sscanf(name, "du[%d]", &a) == 1 && a >= 1 && a <= s->C) {
            src = s->du + (size_t)(a - 1) * s->N;
            is_double = true;
        } else if (std::sscanf(name, "grad_u[%1d%1d]", &a, &b) == 2 && a >= 1 && a <= 3 && b >= 1 && b <= 3)
            src = s->gu[(a - 1) * 3 + (b - 1)];
        else if (std::sscanf(name, "grad_T[%d]", &a) == 1 && a >= 1 && a <= 3)
            src = s->gT[a - 1];
        else if (std::sscanf(name, "Y[%d]", &a) == 1 && a >= 1 && a <= s->Ns)
            src = s->Y[a - 1];
        else if (std::sscanf(name, "D_art[%d]", &a) == 1 && a >= 1 && a <= s->Ns)
            src = s->D_art[a - 1];
        if (!src) throw cles::Error{CLES_ERR_INVALID, "unknown or unallocated field '" + nm + "'"};
        if (s->const_cp() && (nm == "cp" || nm == "Y[1]")) {
            // not stored for a one-species ideal gas: formed from rho as the
            // primitives would, in the working precision
            std::vector<double> r((size_t)s->N);
            CLES_REQUIRE(cles_solver_get_field(s, "rho", r.data()) == CLES_OK, "rho download failed");
            for (size_t i = 0; i < r.size(); i++) {
                double cp = 1.0, y = 1.0;
                if (r[i] > 0) {
                    const double ri = 1 / r[i];
                    double Rm = r[i] * s->Rk[0], cvm = r[i] * s->cvk[0];
                    Rm *= ri;
                    cvm *= ri;
                    cp = cvm + Rm;
                    y = r[i] * ri;
                }
                const double v = nm == "cp" ? cp : y;
                host[i] = s->prec == CLES_F64 ? v : (double)(float)v;
            }
            return CLES_OK;
        }
        std::vector<double> buf((size_t)s->N);
        if (is_double || s->prec == CLES_F64) {
            CLES_HIP(hipMemcpyAsync(buf.data(), src, s->N * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        } else {
            std::vector<float> tmp((size_t)s->N);
            CLES_HIP(hipMemcpyAsync(tmp.data(), src, s->N * sizeof(float), hipMemcpyDeviceToHost, s->st()));
            CLES_HIP(hipStreamSynchronize(s->st()));
            for (size_t i = 0; i < tmp.size(); i++) buf[i] = tmp[i];
        }
        CLES_HIP(hipStreamSynchronize(s->st()));
        std::memcpy(host, buf.data(), buf.size() * sizeof(double));
    })
}

int cles_solver_get_time(cles_solver *s, double *t, int64_t *step)
{
    CLES_TRY({
        CLES_REQUIRE(s && t && step, "NULL argument");
        *t = s->t;
        *step = s->step;
    })
}

int cles_solver_set_time(cles_solver *s, double t, int64_t step)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        s->t = t;
        s->step = step;
    })
}

int cles_apply_bcs(cles_solver *s)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        dispatch(s, [](auto &im) { im.apply_bcs(); });
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

int cles_compute_rhs(cles_solver *s, int primitives_current)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        dispatch(s, [&](auto &im) {
            im.rhs(primitives_current != 0, false, 0, 0, 0);
            im.nscbc_corrections();
            im.sources();
        });
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

int cles_max_rate(cles_solver *s, double *rate, double *rho_min)
{
    CLES_TRY({
        CLES_REQUIRE(s && rate && rho_min, "NULL argument");
        dispatch(s, [&](auto &im) { im.max_rate(rate, rho_min); });
    })
}

int cles_step(cles_solver *s, double dt, int prepared)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        dispatch(s, [&](auto &im) { im.step(dt, prepared != 0); });
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

int cles_filter_state(cles_solver *s)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        dispatch(s, [&](auto &im) { im.filter_state(); });
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

// timestep.jl:277-400 run! with the per-step decisions on the device
namespace {

cles::BatchCtl device_ctl(cles_solver *s)
{
    cles::BatchCtl c{};
    cles_solver::Ctl *k = s->ctl;
    c.count = 1;
    c.rate = s->scal;
    c.rhomin = s->scal + 1;
    c.tfinal = &k->tfinal;
    c.t = &k->t;
    c.dt = &k->dt;
    c.dt_prev = &k->dt_prev;
    c.rate_prev = &k->rate_prev;
    c.dt_seen = &k->dt_seen;
    c.step = &k->step;
    c.nmax = &k->nmax;
    c.status = &k->status;
    c.active = &k->active;
    c.stepped = &k->stepped;
    c.filt = &k->filt;
    c.cfl = s->d.cfl;
    c.predict = s->d.predict;
    c.max_growth = s->d.max_growth;
    c.dt_min = s->d.dt_min;
    c.dt_min_ratio = s->d.dt_min_ratio;
    c.filter_interval = s->d.filter_interval;
    c.dts = s->dts_buf;
    c.step0 = &k->step0;
    c.nrec = s->dts_cap;
    return c;
}

void drop_graph(cles_solver *s)
{
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
}

// capture one device-decided run! iteration into s->gexec
void capture_iteration(cles_solver *s)
{
    drop_graph(s);
    const cles::BatchCtl c = device_ctl(s);
    s->dev_dt = &s->ctl->dt;
    s->dev_gate = &s->ctl->stepped;
    s->dev_filt = &s->ctl->filt;
    hipError_t cap = hipStreamBeginCapture(s->st(), hipStreamCaptureModeThreadLocal);
    if (cap != hipSuccess) {
        s->dev_dt = nullptr;
        s->dev_gate = s->dev_filt = nullptr;
        throw cles::Error{CLES_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(cap)};
    }
    std::string err;
    try {
        dispatch(s, [&](auto &im) { im.iteration_device(c); });
    } catch (const cles::Error &e) {
        err = e.msg;
    }
    hipGraph_t g = nullptr;
    hipError_t end = hipStreamEndCapture(s->st(), &g);
    s->dev_dt = nullptr;
    s->dev_gate = s->dev_filt = nullptr;
    if (!err.empty() || end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw cles::Error{CLES_ERR_HIP, err.empty() ? std::string("hipStreamEndCapture: ") + hipGetErrorString(end) : err};
    }
    s->graph = g;
    hipError_t inst = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
    if (inst != hipSuccess) {
        drop_graph(s);
        throw cles::Error{CLES_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(inst)};
    }
}

// timestep.jl:277-400 with every per-step decision on the device: the host
// replays the captured iteration and reads the control block back every few
// steps (fewer near tfinal or nmax); iterations after the run has stopped
// change nothing (their updates are gated off).
int run_graph(cles_solver *s, double tfinal, int64_t nmax, double *dts, int64_t ndts, std::string &msg, bool graphs)
{
    if (!(s->t < tfinal && s->step < nmax)) return CLES_OK;
    if (ndts > s->dts_cap) {
        drop_graph(s);
        if (s->dts_buf) CLES_HIP(hipFree(s->dts_buf));
        s->dts_buf = nullptr;
        s->dts_cap = 0;
        CLES_HIP(hipMalloc((void **)&s->dts_buf, (size_t)ndts * sizeof(double)));
        s->dts_cap = ndts;
    }
    if (graphs && !s->gexec) capture_iteration(s);
    cles_solver::Ctl h{};
    h.t = s->t;
    h.dt_prev = s->dt_prev;
    h.rate_prev = s->rate_prev;
    h.tfinal = tfinal;
    h.step = s->step;
    h.nmax = nmax;
    h.step0 = s->step;
    h.active = 1;
    *s->host_ctl = h;
    // fresh workspace (timestep.jl:402-405)
    CLES_HIP(hipMemsetAsync(s->du, 0, (size_t)s->N * s->C * sizeof(double), s->st()));
    CLES_HIP(hipMemsetAsync(s->dQ, 0, (size_t)s->N * s->C * s->wsize(), s->st()));
    CLES_HIP(hipMemcpyAsync(s->ctl, s->host_ctl, sizeof(h), hipMemcpyHostToDevice, s->st()));
    CLES_HIP(hipStreamSynchronize(s->st()));
    for (;;) {
        const cles_solver::Ctl &c = *s->host_ctl;
        int64_t k = std::min<int64_t>(16, nmax - c.step);
        if (c.dt_prev > 0) {
            const double est = std::ceil((tfinal - c.t) / c.dt_prev);
            if (est >= 1 && est < (double)k) k = (int64_t)est;
        }
        if (k < 1) k = 1;
        if (graphs) {
            for (int64_t i = 0; i < k; i++) CLES_HIP(hipGraphLaunch(s->gexec, s->st()));
        } else {
            const cles::BatchCtl bc = device_ctl(s);
            s->dev_dt = &s->ctl->dt;
            s->dev_gate = &s->ctl->stepped;
            s->dev_filt = &s->ctl->filt;
            try {
                for (int64_t i = 0; i < k; i++) dispatch(s, [&](auto &im) { im.iteration_device(bc); });
            } catch (...) {
                s->dev_dt = nullptr;
                s->dev_gate = s->dev_filt = nullptr;
                throw;
            }
            s->dev_dt = nullptr;
            s->dev_gate = s->dev_filt = nullptr;
        }
        CLES_HIP(hipMemcpyAsync(s->host_ctl, s->ctl, sizeof(cles_solver::Ctl), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
        if (!s->host_ctl->active) break;
    }
    const cles_solver::Ctl &c = *s->host_ctl;
    const int64_t taken = c.step - h.step0;
    if (ndts > 0 && taken > 0)
        CLES_HIP(hipMemcpy(dts, s->dts_buf, (size_t)std::min(taken, ndts) * sizeof(double), hipMemcpyDeviceToHost));
    s->t = c.t;
    s->step = c.step;
    s->dt_prev = c.dt_prev;
    s->rate_prev = c.rate_prev;
    if (c.status != CLES_OK) {
        // the message of check_step (stepcontrol.jl:184-204) for the rejected dt
        CLES_HIP(hipMemcpy(s->host_scal, s->scal, 2 * sizeof(double), hipMemcpyDeviceToHost));
        std::string m;
        const int rc = check_step(s->d, c.dt, s->host_scal[1], c.dt_seen, m);
        msg = rc == c.status ? m : std::string("run! stopped by check_step (status ") + std::to_string(c.status) + ")";
        return c.status;
    }
    return CLES_OK;
}

}  // namespace

int cles_solver_set_run_mode(cles_solver *s, int mode)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        CLES_REQUIRE(mode >= CLES_RUN_HOST_DT && mode <= CLES_RUN_STEP_GRAPH, "unknown run mode");
        s->run_mode = mode;
    })
}

int cles_run(cles_solver *s, double tfinal, int64_t nmax, double *dts, int64_t ndts)
{
    return cles_run_ctl(s, tfinal, nmax, nullptr, dts, ndts, nullptr);
}

int cles_run_ctl(cles_solver *s, double tfinal, int64_t nmax, const cles_run_hooks *hooks, double *dts,
                 int64_t ndts, int *retries_used)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        if (retries_used) *retries_used = 0;
        const bool hooked = hooks && (hooks->next_time || hooks->after_step || hooks->rewind);
        const bool recover = s->d.retries > 0;
        int rc = CLES_OK;
        std::string msg;
        if (s->run_mode != CLES_RUN_HOST_DT && !s->prof && !hooked && !recover) {
            rc = run_graph(s, tfinal, nmax, dts, ndts, msg, s->run_mode == CLES_RUN_STEP_GRAPH);
            if (rc != CLES_OK) throw cles::Error{rc, msg};
            return CLES_OK;
        }
        // timestep.jl:286: the savepoint (an FP64 copy of Q, t, step)
        double *save_q = nullptr;
        double save_t = 0.0;
        int64_t save_step = 0;
        const size_t qbytes = (size_t)s->N * s->C * sizeof(double);
        if (recover) {
            CLES_HIP(hipMalloc((void **)&save_q, qbytes));
            save_t = s->t;
            save_step = s->step;
        }
        int attempts = 0;
        try {
            CLES_HIP(hipSetDevice(s->ctx->device));
            if (recover) CLES_HIP(hipMemcpyAsync(save_q, s->Q, qbytes, hipMemcpyDeviceToDevice, s->st()));
            const int landing = s->d.landing_steps > 0 ? s->d.landing_steps : 2;
            dispatch(s, [&](auto &im) {
                // fresh Workspace (timestep.jl:402-405)
                CLES_HIP(hipMemsetAsync(s->du, 0, (size_t)s->N * s->C * sizeof(double), s->st()));
                CLES_HIP(hipMemsetAsync(s->dQ, 0, (size_t)s->N * s->C * s->wsize(), s->st()));
                double dt_seen = 0.0;
                int64_t taken = 0;
                int64_t guard_step = -1;
                while (s->t < tfinal && s->step < nmax) {
                    s->tstage = s->t;
                    im.apply_bcs();
                    double rate = 0, rho_min = 0;
                    im.max_rate(&rate, &rho_min);
                    double r = rate;   // timestep.jl:29-42 predicted_dt
                    if (s->d.predict > 0 && s->rate_prev > 0) r = jmax(r, r + s->d.predict * (r - s->rate_prev));
                    double dt = s->d.cfl / r;
                    if (s->d.max_growth > 0 && s->dt_prev > 0) dt = jmin(dt, s->d.max_growth * s->dt_prev);
                    std::string m;
                    const int fail = check_step(s->d, dt, rho_min, dt_seen, m);
                    if (fail != CLES_OK) {
                        if (!recover || attempts >= s->d.retries) {
                            rc = fail;
                            msg = m;
                            return;
                        }
                        // timestep.jl:318-344: roll back, lower the CFL (compounding)
                        attempts++;
                        CLES_HIP(hipMemcpyAsync(s->Q, save_q, qbytes, hipMemcpyDeviceToDevice, s->st()));
                        guard_step = s->step;
                        s->t = save_t;
                        s->step = save_step;
                        s->d.cfl *= s->d.cfl_backoff;
                        s->dt_prev = 0.0;
                        s->rate_prev = 0.0;
                        dt_seen = 0.0;
                        if (hooks && hooks->rewind) {
                            CLES_HIP(hipStreamSynchronize(s->st()));
                            hooks->rewind(hooks->user, save_t, save_step);
                        }
                        continue;
                    }
                    // timestep.jl:345-351: bank the state that has just passed its check
                    if (recover && s->d.savepoint_interval > 0 && s->step > guard_step &&
                        s->step % s->d.savepoint_interval == 0) {
                        CLES_HIP(hipMemcpyAsync(save_q, s->Q, qbytes, hipMemcpyDeviceToDevice, s->st()));
                        save_t = s->t;
                        save_step = s->step;
                    }
                    dt_seen = jmax(dt_seen, dt);
                    dt = jmin(dt, tfinal - s->t);
                    // timestep.jl:367-383 soft landing on the next scheduled instant
                    if (hooks && hooks->next_time) {
                        const double next = hooks->next_time(hooks->user, s->t, s->step);
                        const double gap = next - s->t;
                        if (next <= tfinal && gap > 0 && gap < dt * landing) dt = gap / std::ceil(gap / dt);
                    }
                    im.step(dt, true);
                    if (dts && taken < ndts) dts[taken] = dt;
                    taken++;
                    s->t += dt;
                    s->step += 1;
                    s->dt_prev = dt;
                    s->rate_prev = rate;
                    if (s->d.filter_interval > 0 && s->step % s->d.filter_interval == 0) im.filter_state();
                    if (hooks && hooks->after_step) {
                        CLES_HIP(hipStreamSynchronize(s->st()));
                        if (hooks->after_step(hooks->user, s->t, s->step, dt)) break;
                    }
                }
                CLES_HIP(hipStreamSynchronize(s->st()));
            });
        } catch (...) {
            if (save_q) {
                (void)hipStreamSynchronize(s->st());
                (void)hipFree(save_q);
            }
            throw;
        }
        if (save_q) {
            CLES_HIP(hipStreamSynchronize(s->st()));
            CLES_HIP(hipFree(save_q));
        }
        if (retries_used) *retries_used = attempts;
        if (attempts > 0) drop_graph(s);   // the CFL was lowered
        if (rc != CLES_OK) throw cles::Error{rc, msg};
    })
}

int cles_solver_get_cfl(cles_solver *s, double *cfl)
{
    CLES_TRY({
        CLES_REQUIRE(s && cfl, "NULL argument");
        *cfl = s->d.cfl;
    })
}

int cles_solver_switch_bc(cles_solver *s, int dim, int side)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        CLES_REQUIRE(dim >= 0 && dim < 3 && side >= 0 && side < 2, "dim must be 0..2 and side 0..1");
        CLES_REQUIRE(s->d.switchable[dim][side], "this face is not a SwitchableBC");
        if (s->switched[dim][side]) return CLES_OK;   // boundary.jl:131-137: repeated calls do nothing more
        s->switched[dim][side] = true;
        s->d.bc[dim][side] = s->d.bc_after[dim][side];
        s->nscbc = false;
        for (int k = 0; k < 3; k++)
            for (int sd = 0; sd < 2; sd++)
                if (s->d.bc[k][sd] == CLES_BC_NSCBC_IN || s->d.bc[k][sd] == CLES_BC_NSCBC_OUT)
                    s->nscbc = s->active[k] || s->nscbc;
        drop_graph(s);   // a captured step holds the face's condition
    })
}

int cles_solver_bc_switched(cles_solver *s, int dim, int side, int *out)
{
    CLES_TRY({
        CLES_REQUIRE(s && out, "NULL argument");
        CLES_REQUIRE(dim >= 0 && dim < 3 && side >= 0 && side < 2, "dim must be 0..2 and side 0..1");
        *out = s->switched[dim][side] ? 1 : 0;
    })
}

int cles_solver_set_cfl(cles_solver *s, double cfl)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        CLES_REQUIRE(cfl > 0, "cfl must be positive");
        s->d.cfl = cfl;
        drop_graph(s);   // a captured iteration holds the CFL as a kernel argument
    })
}

int cles_solver_profile(cles_solver *s, int enable)
{
    CLES_TRY({
        CLES_REQUIRE(s, "NULL argument");
        s->resolve();
        if (enable) s->phases.assign(cles::PH_N, cles_solver::Phase{});
        s->prof = enable != 0;
    })
}

int cles_solver_profile_read(cles_solver *s, int index, char *name, int namelen, double *ms, int64_t *launches,
                             double *bytes)
{
    CLES_TRY({
        CLES_REQUIRE(s && name && namelen > 0 && ms && launches && bytes, "NULL argument");
        s->resolve();
        if (index < 0 || index >= cles::PH_N || s->phases.empty())
            throw cles::Error{CLES_ERR_INVALID, "no such phase"};
        std::snprintf(name, (size_t)namelen, "%s", cles::PHASE_NAMES[index]);
        *ms = s->phases[index].ms;
        *launches = s->phases[index].launches;
        *bytes = s->phases[index].bytes;
    })
}

int cles_diagnostics(cles_solver *s, double *ke, double *ens)
{
    CLES_TRY({
        CLES_REQUIRE(s && ke && ens, "NULL argument");
        dispatch(s, [&](auto &im) { im.diagnostics(ke, ens); });
    })
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Stacks of same-shape two-dimensional problems (cles_stack_*): one solver
// whose z planes are the instances, so that every launch covers all of them,
// and run! per instance through the device control of the batch form
// (launch_batch_dt / launch_batch_post on per-instance arrays).  An instance's
// arithmetic is that of the same problem alone: pointwise kernels and line
// solves act on its own plane, its rate and minimum density are reduced over
// its own points, and its dt, LSRK gate and filter flag are its own.
// ---------------------------------------------------------------------------
struct cles_stack {
    cles_solver *s = nullptr;
    int count = 0;
    double *t = nullptr, *dt = nullptr, *dt_prev = nullptr, *rate_prev = nullptr, *dt_seen = nullptr,
           *tfinal = nullptr;
    int64_t *step = nullptr, *nmax = nullptr, *step0 = nullptr;
    int *status = nullptr, *active = nullptr, *stepped = nullptr, *filt = nullptr, *cnt = nullptr;
    int *host_cnt = nullptr;
    ~cles_stack()
    {
        if (host_cnt) (void)hipHostFree(host_cnt);
        if (s) delete s;
    }
};

namespace {

void check_plane(const cles_stack *st, int i, const int64_t padded[3], const int halo[3])
{
    CLES_REQUIRE(i >= 0 && i < st->count, "instance index out of range");
    for (int k = 0; k < 2; k++)
        CLES_REQUIRE(halo[k] >= 0 && padded[k] >= st->s->n[k] + 2 * (int64_t)halo[k],
                     "padded extents smaller than n + 2*halo");
    CLES_REQUIRE(halo[2] >= 0 && padded[2] >= 1 + 2 * (int64_t)halo[2], "padded extents smaller than n + 2*halo");
}

// instance i's plane <-> the padded host layout of one problem (cles_solver_upload)
void copy_plane(cles_stack *st, int i, double *host, const int64_t padded[3], const int halo[3], bool up)
{
    cles_solver *s = st->s;
    const int64_t nx = s->n[0], ny = s->n[1], Ni = (int64_t)s->ipts;
    const int64_t P0 = padded[0], P1 = padded[1], NP = padded[0] * padded[1] * padded[2];
    std::vector<double> buf((size_t)Ni * s->C);
    if (!up) {
        for (int c = 0; c < s->C; c++)
            CLES_HIP(hipMemcpyAsync(&buf[(size_t)c * Ni], s->Q + (size_t)c * s->N + (size_t)i * Ni,
                                    (size_t)Ni * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
    }
    for (int c = 0; c < s->C; c++)
        for (int64_t j = 0; j < ny; j++) {
            double *h = host + c * NP + halo[0] + P0 * ((j + halo[1]) + P1 * (int64_t)halo[2]);
            double *b = &buf[(size_t)c * Ni + (size_t)nx * j];
            if (up) std::memcpy(b, h, sizeof(double) * nx);
            else std::memcpy(h, b, sizeof(double) * nx);
        }
    if (up) {
        for (int c = 0; c < s->C; c++)
            CLES_HIP(hipMemcpyAsync(s->Q + (size_t)c * s->N + (size_t)i * Ni, &buf[(size_t)c * Ni],
                                    (size_t)Ni * sizeof(double), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
    }
}

}  // namespace

extern "C" {

int cles_stack_create(cles_ctx *ctx, const cles_solver_desc *desc, int count, cles_stack **out)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && desc && out, "NULL argument");
        CLES_REQUIRE(count >= 1 && count <= 4096, "a stack holds 1..4096 instances");
        auto *st = new cles_stack;
        try {
            create_solver(ctx, desc, nullptr, nullptr, count, &st->s);
            st->count = count;
            cles_solver *s = st->s;
            const size_t cnt = (size_t)count;
            for (double **fp : {&st->t, &st->dt, &st->dt_prev, &st->rate_prev, &st->dt_seen, &st->tfinal})
                *fp = (double *)s->alloc(cnt * sizeof(double));
            for (int64_t **fp : {&st->step, &st->nmax, &st->step0}) *fp = (int64_t *)s->alloc(cnt * sizeof(int64_t));
            for (int **fp : {&st->status, &st->active, &st->stepped, &st->filt})
                *fp = (int *)s->alloc(cnt * sizeof(int));
            st->cnt = (int *)s->alloc(sizeof(int));
            CLES_HIP(hipHostMalloc((void **)&st->host_cnt, sizeof(int), hipHostMallocDefault));
            CLES_HIP(hipStreamSynchronize(s->st()));
        } catch (...) {
            delete st;
            throw;
        }
        *out = st;
    })
}

int cles_stack_destroy(cles_stack *st)
{
    CLES_TRY({ delete st; })
}

int cles_stack_upload(cles_stack *st, int i, const double *Q_host, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(st && Q_host && padded && halo, "NULL argument");
        check_plane(st, i, padded, halo);
        copy_plane(st, i, const_cast<double *>(Q_host), padded, halo, true);
    })
}

int cles_stack_download(cles_stack *st, int i, double *Q_host, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(st && Q_host && padded && halo, "NULL argument");
        check_plane(st, i, padded, halo);
        copy_plane(st, i, Q_host, padded, halo, false);
    })
}

int cles_stack_run(cles_stack *st, const double *tfinal, const int64_t *nmax, int64_t max_steps, int64_t *taken)
{
    CLES_TRY({
        CLES_REQUIRE(st && tfinal && nmax, "NULL argument");
        cles_solver *s = st->s;
        const int planes = st->count;
        const size_t cnt = (size_t)planes;
        // the run! entry of every instance (timestep.jl:281-344)
        std::vector<double> t(cnt), tf(tfinal, tfinal + cnt);
        std::vector<int64_t> stp(cnt), nmx(nmax, nmax + cnt);
        std::vector<int> status(cnt), active(cnt, 0);
        CLES_HIP(hipMemcpyAsync(t.data(), st->t, cnt * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipMemcpyAsync(stp.data(), st->step, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipMemcpyAsync(status.data(), st->status, cnt * sizeof(int), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
        int nact = 0;
        for (int i = 0; i < st->count; i++) {
            active[i] = status[i] == CLES_OK && t[i] < tf[i] && stp[i] < nmx[i];
            nact += active[i];
        }
        CLES_HIP(hipMemcpyAsync(st->tfinal, tf.data(), cnt * sizeof(double), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipMemcpyAsync(st->nmax, nmx.data(), cnt * sizeof(int64_t), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipMemcpyAsync(st->active, active.data(), cnt * sizeof(int), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipMemcpyAsync(st->step0, stp.data(), cnt * sizeof(int64_t), hipMemcpyHostToDevice, s->st()));
        CLES_HIP(hipMemsetAsync(st->dt_seen, 0, cnt * sizeof(double), s->st()));
        // fresh workspace (timestep.jl:402-405)
        CLES_HIP(hipMemsetAsync(s->du, 0, (size_t)s->N * s->C * sizeof(double), s->st()));
        CLES_HIP(hipMemsetAsync(s->dQ, 0, (size_t)s->N * s->C * s->wsize(), s->st()));
        cles::BatchCtl c{};
        c.count = planes;
        c.rate = s->srate;
        c.rhomin = s->srmin;
        c.tfinal = st->tfinal;
        c.t = st->t;
        c.dt = st->dt;
        c.dt_prev = st->dt_prev;
        c.rate_prev = st->rate_prev;
        c.dt_seen = st->dt_seen;
        c.step = st->step;
        c.nmax = st->nmax;
        c.status = st->status;
        c.active = st->active;
        c.stepped = st->stepped;
        c.filt = st->filt;
        c.cfl = s->d.cfl;
        c.predict = s->d.predict;
        c.max_growth = s->d.max_growth;
        c.dt_min = s->d.dt_min;
        c.dt_min_ratio = s->d.dt_min_ratio;
        c.filter_interval = s->d.filter_interval;
        c.dts = nullptr;
        c.step0 = st->step0;
        c.nrec = 0;
        int64_t steps = 0;
        s->dev_dt = st->dt;
        s->dev_gate = st->stepped;
        s->dev_filt = st->filt;
        try {
            const int64_t poll = 8;
            while (nact > 0 && steps < max_steps) {
                dispatch(s, [&](auto &im) { im.iteration_device(c); });
                steps++;
                if (steps % poll == 0 || steps == max_steps) {
                    cles::launch_batch_count(st->active, planes, st->cnt, s->st());
                    CLES_HIP(hipMemcpyAsync(st->host_cnt, st->cnt, sizeof(int), hipMemcpyDeviceToHost, s->st()));
                    CLES_HIP(hipStreamSynchronize(s->st()));
                    nact = *st->host_cnt;
                }
            }
            CLES_HIP(hipStreamSynchronize(s->st()));
        } catch (...) {
            s->dev_dt = nullptr;
            s->dev_gate = s->dev_filt = nullptr;
            throw;
        }
        s->dev_dt = nullptr;
        s->dev_gate = s->dev_filt = nullptr;
        if (taken) *taken = steps;
    })
}

int cles_stack_state(cles_stack *st, double *t, int64_t *steps, int *status)
{
    CLES_TRY({
        CLES_REQUIRE(st, "NULL argument");
        cles_solver *s = st->s;
        const size_t cnt = (size_t)st->count;
        if (t) CLES_HIP(hipMemcpyAsync(t, st->t, cnt * sizeof(double), hipMemcpyDeviceToHost, s->st()));
        if (steps) CLES_HIP(hipMemcpyAsync(steps, st->step, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, s->st()));
        if (status) CLES_HIP(hipMemcpyAsync(status, st->status, cnt * sizeof(int), hipMemcpyDeviceToHost, s->st()));
        CLES_HIP(hipStreamSynchronize(s->st()));
    })
}

}  // extern "C"
