// Batched independent one-dimensional instances: many copies of the per-step
// path of run! (timestep.jl:277-400) advanced together, each instance with its
// own grid length, spacing, end time, step cap, boundary states and dt.
//
// Layout: the instances are laid end to end in every field (instance i at
// [off[i], off[i] + n[i])), so the pointwise kernels of the single-domain
// solver run unchanged over the total point count.  Line operators use the
// ragged-line kernel (instances grouped by chunk count, several per wave); the
// delta4 stencil, the CFL reduction and the kinetic energy use one block per
// instance.  The dt decision of run! (predicted_dt, check_step, the tfinal
// clip) runs on the device per instance, so the host issues whole steps
// without reading anything back; it polls the number of active instances
// every few steps.  An instance that fails check_step is frozen with its
// status and the others go on.
//
// Every instance is computed from its own data only (its own lines, its own
// blocks, fixed-order reductions), so its result is the same alone and in any
// mix of instances.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "cles_internal.h"
#include "cles_point.h"

struct cles_batch {
    cles_ctx *ctx = nullptr;
    cles_solver_desc d{};
    int Ns = 1, C = 5, prec = CLES_F64;
    int64_t count = 0, N = 0;
    int periodic = 0;
    std::vector<int64_t> off;
    std::vector<int> n;
    std::vector<double> h, L, origin;
    std::vector<void *> allocs;
    double Rk[CLES_MAXSP] = {}, cvk[CLES_MAXSP] = {}, cpk[CLES_MAXSP] = {};
    const cles::N9Table *n9 = nullptr;   // Nasa9Mixture tables on the device (NULL: IdealMixture)

    // line plans (one device buffer of tables) and wave work lists
    void *tables = nullptr;
    cles::LineDev *dplans = nullptr, *fplans = nullptr;
    int *order_d = nullptr, *items_d = nullptr, *order_f = nullptr, *items_f = nullptr;
    int nitems_d = 0, nitems_f = 0, nshort_d = 0, nshort_f = 0;
    int dq = 1, dM = 2, dsym = 0, fq = 1, fM = 4, fsym = 1;

    // geometry on the device
    int64_t *g_off = nullptr;
    int *g_n = nullptr;
    double *g_h = nullptr;

    // state
    double *Q = nullptr, *du = nullptr;
    void *dQ = nullptr;
    void *rho = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr, *T = nullptr, *c = nullptr,
         *cp = nullptr;
    void *Y[CLES_MAXSP] = {}, *gu[3] = {}, *gT = nullptr, *gY[CLES_MAXSP] = {};
    void *mu_art = nullptr, *beta_art = nullptr, *kappa_art = nullptr, *D_art[CLES_MAXSP] = {};
    void *sensor = nullptr, *strain = nullptr, *esens = nullptr, *tmp1 = nullptr, *tmp2 = nullptr;
    // the sensors of mu*/beta*, kappa* and D*_k side by side (stride N), and
    // their coefficient epilogues (device table) for one smoothing launch
    int nsens = 0;
    cles::EpiArgs *art_multi = nullptr;
    void *flux[CLES_MAXSP + 4] = {};

    // run! control per instance
    double *t = nullptr, *dt = nullptr, *dt_prev = nullptr, *rate_prev = nullptr, *dt_seen = nullptr,
           *rate = nullptr, *rhomin = nullptr, *tfinal = nullptr, *bcst = nullptr, *ke = nullptr;
    int64_t *step = nullptr, *nmax = nullptr, *step0 = nullptr;
    int *status = nullptr, *active = nullptr, *stepped = nullptr, *filt = nullptr, *cnt = nullptr;
    int *host_cnt = nullptr;

    // CLES_RUN_STEP_GRAPH: GRAPH_ITERS run! iterations captured once and
    // replayed; the captured kernels record dts into gdts (nrec = g_nrec)
    int run_mode = CLES_RUN_HOST_DT;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double *gdts = nullptr;
    size_t gdts_cap = 0;
    int64_t g_nrec = -1;
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    hipStream_t st() const { return ctx->stream; }
    size_t wsize() const { return prec == CLES_F64 ? 8 : 4; }
    void *alloc(size_t bytes)
    {
        void *ptr = nullptr;
        hipError_t e = hipMalloc(&ptr, bytes ? bytes : 8);
        if (e != hipSuccess) throw cles::Error{CLES_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
        allocs.push_back(ptr);
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
    template <typename X>
    X *upload_vec(const std::vector<X> &v)
    {
        X *d = (X *)alloc(v.size() * sizeof(X));
        if (!v.empty()) upload(d, v.data(), v.size() * sizeof(X));
        return d;
    }

    ~cles_batch()
    {
        if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        drop_graph();
        if (gdts) (void)hipFree(gdts);
        for (void *ptr : allocs) (void)hipFree(ptr);
        if (host_cnt) (void)hipHostFree(host_cnt);
        if (ctx) cles::ctx_release(ctx);
    }
};

namespace cles {

namespace {

// Builds every instance's plan for one scheme (deduplicated by (n, h)), packs
// the tables into `blob`, and returns the per-instance host LineDev (pointers
// as offsets into the blob, bound after upload) and the wave work list.
struct PlanSet {
    std::vector<LineDev> dev;                       // per instance, pointers are blob offsets
    std::vector<int> order, items;
    int nshort = 0;                                 // leading items of the short class
};

void build_plan_set(PlanSet &ps, std::vector<unsigned char> &blob, const cles_scheme_desc &s,
                    const std::vector<int> &n, const std::vector<double> &h, int periodic)
{
    const int64_t count = (int64_t)n.size();
    std::map<std::pair<int, double>, std::pair<LineDev, size_t>> cache;   // (n, h) -> (dev, blob offset)
    ps.dev.resize(count);
    std::vector<int> P(count);
    for (int64_t i = 0; i < count; i++) {
        const double hk = s.symmetric ? 1.0 : h[i];   // filters do not scale with h
        auto key = std::make_pair(n[i], hk);
        auto it = cache.find(key);
        if (it == cache.end()) {
            // lines of up to 64 x 16 points are split into chunks of at most
            // 16 (the short class: a kernel with half the registers and LDS
            // per wave and sweeps half as long); longer ones into up to 32
            LineTables t;
            const int mc = n[i] <= 64 * RAG_SHORT_CHUNK ? RAG_SHORT_CHUNK : RAG_MAX_CHUNK;
            build_line_tables(t, s, n[i], periodic, h[i], mc, 1, true);
            if (t.dev.P > 64)
                throw Error{CLES_ERR_UNSUPPORTED, "batched instances are limited to " +
                                                      std::to_string(64 * RAG_MAX_CHUNK) + " points"};
            const size_t base = blob.size();
            blob.insert(blob.end(), t.image.begin(), t.image.end());
            LineDev d = t.dev;
            // offsets relative to the blob; rebased after upload
            d.coef64 = (const double *)(uintptr_t)(base);
            d.coef32 = (const float *)(uintptr_t)(base + t.o_coef32);
            d.row64 = (const double *)(uintptr_t)(base + t.o_row64);
            d.row32 = (const float *)(uintptr_t)(base + t.o_row32);
            d.pcr = (const double *)(uintptr_t)(base + t.o_pcr);
            d.rowt64 = (const double *)(uintptr_t)(base + t.o_rowt64);
            d.rowt32 = (const float *)(uintptr_t)(base + t.o_rowt32);
            it = cache.emplace(key, std::make_pair(d, base)).first;
        }
        ps.dev[i] = it->second.first;
        P[i] = ps.dev[i].P;
    }
    // instances grouped by class (short first), then chunk count; one wave
    // holds 64/P of them
    std::vector<int> cls(count);
    for (int64_t i = 0; i < count; i++)
        cls[i] = ps.dev[i].m + (ps.dev[i].extra ? 1 : 0) <= RAG_SHORT_CHUNK ? 0 : 1;
    ps.order.resize(count);
    for (int64_t i = 0; i < count; i++) ps.order[i] = (int)i;
    std::stable_sort(ps.order.begin(), ps.order.end(),
                     [&](int a, int b) { return cls[a] != cls[b] ? cls[a] < cls[b] : P[a] < P[b]; });
    ps.items.clear();
    ps.nshort = 0;
    int64_t k = 0;
    while (k < count) {
        const int i0 = ps.order[k];
        const int Pk = P[i0], ck = cls[i0];
        const int per = 64 / Pk;
        int cnt = 0;
        while (k + cnt < count && cnt < per && P[ps.order[k + cnt]] == Pk && cls[ps.order[k + cnt]] == ck) cnt++;
        ps.items.push_back(Pk);
        ps.items.push_back((int)k);
        ps.items.push_back(cnt);
        if (ck == 0) ps.nshort++;
        k += cnt;
    }
}

void rebase(std::vector<LineDev> &v, unsigned char *base)
{
    for (LineDev &d : v) {
        d.coef64 = (const double *)(base + (uintptr_t)d.coef64);
        d.coef32 = (const float *)(base + (uintptr_t)d.coef32);
        d.row64 = (const double *)(base + (uintptr_t)d.row64);
        d.row32 = (const float *)(base + (uintptr_t)d.row32);
        d.pcr = (const double *)(base + (uintptr_t)d.pcr);
        d.rowt64 = (const double *)(base + (uintptr_t)d.rowt64);
        d.rowt32 = (const float *)(base + (uintptr_t)d.rowt32);
    }
}

bool rag_shape_ok(const cles_scheme_desc &s)
{
    if (s.symmetric) return s.q == 1 && s.M <= 4;
    return (s.q == 1 && (s.M == 2 || s.M == 1)) || (s.q == 2 && s.M == 3);
}

template <typename W>
struct BImpl {
    cles_batch &b;
    explicit BImpl(cles_batch &b_) : b(b_) {}
    W *f(void *ptr) const { return (W *)ptr; }
    const W *f(const void *ptr) const { return (const W *)ptr; }

    BatchGeom geom() const
    {
        BatchGeom g{};
        g.count = b.count;
        g.N = b.N;
        g.off = b.g_off;
        g.n = b.g_n;
        g.h = b.g_h;
        g.periodic = b.periodic;
        return g;
    }
    RagLines rag(bool filter, const double *dtv, const int *mask) const
    {
        RagLines r{};
        r.plans = filter ? b.fplans : b.dplans;
        r.off = b.g_off;
        r.order = filter ? b.order_f : b.order_d;
        r.items = filter ? b.items_f : b.items_d;
        r.nitems = filter ? b.nitems_f : b.nitems_d;
        r.nshort = filter ? b.nshort_f : b.nshort_d;
        r.dt = dtv;
        r.mask = mask;
        r.q = filter ? b.fq : b.dq;
        r.M = filter ? b.fM : b.dM;
        r.sym = filter ? b.fsym : b.dsym;
        r.nf = 1;
        r.fstride = b.N;
        return r;
    }
    // nf fields laid out at stride N, one launch
    void lines(bool filter, int nf, const void *in, const EpiArgs &e, const double *dtv = nullptr,
               const int *mask = nullptr, int dtype = -1) const
    {
        RagLines r = rag(filter, dtv, mask);
        r.nf = nf;
        launch_line_rag(r, dtype < 0 ? b.prec : dtype, in, e, b.st());
    }
    void line(bool filter, const void *in, const EpiArgs &e, const int *mask = nullptr) const
    {
        launch_line_rag(rag(filter, nullptr, mask), b.prec, in, e, b.st());
    }
    EpiArgs store(void *out, int kind = EPI_STORE) const
    {
        EpiArgs e{};
        e.kind = kind;
        e.out = out;
        return e;
    }
    GradPtrs<W> grads() const
    {
        GradPtrs<W> g{};
        for (int j = 0; j < 3; j++) g.gu[j] = f(b.gu[j]);
        return g;
    }

    // physics.jl:604-653
    void primitives() const
    {
        PrimArgs<W> a{};
        a.Q = b.Q;
        a.Ns = b.Ns;
        for (int k = 0; k < b.Ns; k++) { a.Rk[k] = b.Rk[k]; a.cvk[k] = b.cvk[k]; a.Y[k] = f(b.Y[k]); }
        a.n9 = b.n9;
        a.sg = b.d.eos_kind == CLES_EOS_STIFFENED;
        a.sg_gamma = b.d.gamma[0];
        a.sg_pinf = b.d.p_inf;
        a.sg_cv = b.cvk[0];
        a.rho = f(b.rho); a.u = f(b.u); a.v = f(b.v); a.w = f(b.w);
        a.p = f(b.p); a.T = f(b.T); a.c = f(b.c); a.cp = f(b.cp);
        a.esens = f(b.esens);
        launch_primitives<W>(a, b.N, b.st());
    }

    // artificial.jl:261-359 along the single active dimension: one smoothing
    // pass, which is also the last and carries the coefficient epilogue.  The
    // mu*/beta*, kappa* and D*_k sensors are smoothed together in one launch
    // (each field with its own epilogue); the dilatation sensor's beta* goes
    // after them, as it overwrites beta*.
    void artificial() const
    {
        const cles_solver_desc &d = b.d;
        if (!d.art_enabled) return;
        const BatchGeom g = geom();
        const size_t ws = (size_t)b.N * b.wsize();
        auto sens = [&](int i) { return f((char *)b.sensor + (size_t)i * ws); };
        launch_strain<W>(grads(), f(b.strain), b.N, b.st());
        for (int i0 = 0; i0 < b.nsens; i0 += D4_MF) {   // D4_MF fields per launch
            D4Multi<W> m4{};
            m4.nf = std::min(D4_MF, b.nsens - i0);
            for (int k = 0; k < m4.nf; k++) {
                const int i = i0 + k;
                m4.f[k] = i == 0 ? f(b.strain) : i == 1 ? f(b.esens) : f(b.Y[i - 2]);
                m4.out[k] = sens(i);
                m4.wpow[k] = i == 0 ? 2 : 1;
            }
            launch_delta4_batch_multi<W>(m4, g, b.st());
        }
        EpiArgs m{};
        m.kind = EPI_ART_MULTI;
        m.multi = b.art_multi;
        lines(true, b.nsens, b.sensor, m);
        if (d.beta_sensor == 1) {
            launch_gate_beta<W>(grads(), f(b.beta_art), b.N, b.st());
        } else if (d.beta_sensor == 2) {
            launch_dilatation_prep<W>(grads(), f(b.tmp1), f(b.tmp2), b.N, b.st());
            launch_delta4_batch<W>(f(b.tmp1), sens(0), g, 2, b.st());
            line(true, b.sensor, store(b.sensor));
            launch_dilatation_beta<W>(f(b.rho), sens(0), f(b.tmp2), f(b.beta_art), d.C_beta, b.N, b.st());
        }
    }

    void fluxes() const
    {
        FluxArgs<W> a{};
        a.Q = b.Q;
        a.Ns = b.Ns;
        a.active[0] = 1;
        a.mu0 = b.d.mu0;
        a.Pr = b.d.Pr;
        a.Sc = b.d.Sc;
        for (int k = 0; k < b.Ns; k++) {
            a.cpk[k] = b.cpk[k];
            a.Y[k] = f(b.Y[k]);
            a.D_art[k] = f(b.D_art[k]);
        }
        a.n9 = b.n9;
        a.rho = f(b.rho); a.u = f(b.u); a.v = f(b.v); a.w = f(b.w);
        a.p = f(b.p); a.T = f(b.T); a.cp = f(b.cp);
        a.mu_art = f(b.mu_art);
        a.beta_art = f(b.beta_art);
        a.kappa_art = f(b.kappa_art);
        a.g = grads();
        a.gT[0] = f(b.gT);
        for (int k = 0; k < b.Ns; k++) a.gY[0][k] = f(b.gY[k]);
        for (int c = 0; c < b.C; c++) a.flux[0][c] = f(b.flux[c]);
        launch_fluxes<W>(a, b.N, b.st());
    }

    W *dQc(int c) const { return f(b.dQ) + (size_t)c * b.N; }

    // rhs.jl:624-694 for one active dimension; with fuse, the divergence pass
    // performs the LSRK stage update of the instances that step (their own dt)
    void rhs(bool primitives_current, bool fuse, double A, double B) const
    {
        if (!primitives_current) primitives();
        lines(false, 3, b.u, store(b.gu[0]));                                  // u, v, w -> grad_u
        artificial();
        lines(false, b.Ns > 1 ? 1 + b.Ns : 1, b.T, store(b.gT));              // T, Y_k -> gradients
        fluxes();
        if (fuse) {
            EpiArgs e{};
            e.kind = EPI_LSRK;
            e.a0 = nullptr;   // dQ = -D_x F_x
            e.du = b.du;
            e.Q = b.Q;
            e.c0 = A;
            e.c1 = B;
            lines(false, b.C, b.flux[0], e, b.dt, b.stepped);
        } else {
            lines(false, b.C, b.flux[0], store(dQc(0), EPI_STORE_NEG));
        }
    }

    void bcs(const int *mask) const
    {
        const int lo = b.d.bc[0][0], hi = b.d.bc[0][1];
        if (lo == CLES_BC_PERIODIC) return;
        BatchBC bc{};
        bc.lo = lo;
        bc.hi = hi;
        bc.Ns = b.Ns;
        bc.state = b.bcst;
        bc.mask = mask;
        launch_batch_bcs(geom(), bc, b.Q, b.st());
    }

    // timestep.jl:125-175 per instance
    void rates(const int *mask) const
    {
        primitives();
        RateArgs<W> a{};
        a.Q = b.Q;
        a.Ns = b.Ns;
        a.active[0] = 1;
        a.mu0 = b.d.mu0;
        a.Pr = b.d.Pr;
        a.Sc = b.d.Sc;
        a.rho = f(b.rho); a.u = f(b.u); a.v = f(b.v); a.w = f(b.w); a.c = f(b.c); a.cp = f(b.cp);
        a.mu_art = f(b.mu_art);
        a.beta_art = f(b.beta_art);
        a.kappa_art = f(b.kappa_art);
        for (int k = 0; k < b.Ns; k++) a.D_art[k] = f(b.D_art[k]);
        launch_rate_batch<W>(a, geom(), mask, b.rate, b.rhomin, b.st());
    }

    BatchCtl ctl(double *dts, int64_t nrec) const
    {
        BatchCtl c{};
        c.count = b.count;
        c.rate = b.rate;
        c.rhomin = b.rhomin;
        c.tfinal = b.tfinal;
        c.t = b.t;
        c.dt = b.dt;
        c.dt_prev = b.dt_prev;
        c.rate_prev = b.rate_prev;
        c.dt_seen = b.dt_seen;
        c.step = b.step;
        c.nmax = b.nmax;
        c.status = b.status;
        c.active = b.active;
        c.stepped = b.stepped;
        c.filt = b.filt;
        c.cfl = b.d.cfl;
        c.predict = b.d.predict;
        c.max_growth = b.d.max_growth;
        c.dt_min = b.d.dt_min;
        c.dt_min_ratio = b.d.dt_min_ratio;
        c.filter_interval = b.d.filter_interval;
        c.dts = dts;
        c.step0 = b.step0;
        c.nrec = nrec;
        return c;
    }

    // timestep.jl:414-426 on the FP64 state of the instances due a filter
    void filter_state() const
    {
        EpiArgs e{};
        e.kind = EPI_STORE;
        e.out = b.Q;
        lines(true, b.C, b.Q, e, nullptr, b.filt, CLES_F64);
    }

    // one run! iteration of every active instance (timestep.jl:345-399)
    void iteration(const BatchCtl &c) const
    {
        bcs(b.active);
        rates(b.active);
        launch_batch_dt(c, b.st());
        for (int stg = 0; stg < 5; stg++) {
            if (stg > 0) bcs(b.stepped);
            rhs(stg == 0, true, LSRK_A[stg], LSRK_B[stg]);
        }
        bcs(b.stepped);
        launch_batch_post(c, b.st());
        filter_state();
    }
};

template <typename Fn>
void bdispatch(cles_batch *b, Fn &&fn)
{
    if (b->prec == CLES_F64) {
        BImpl<double> im(*b);
        fn(im);
    } else {
        BImpl<float> im(*b);
        fn(im);
    }
}

void check_inst(const cles_batch *b, int64_t i)
{
    if (i < 0 || i >= b->count) throw Error{CLES_ERR_INVALID, "instance index out of range"};
}

}  // namespace
}  // namespace cles

extern "C" {

int cles_batch_create(cles_ctx *ctx, const cles_solver_desc *desc, int64_t count, const cles_batch_instance *inst,
                      cles_batch **out)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && desc && out && (inst || count == 0), "NULL argument");
        CLES_REQUIRE(count >= 1, "a batch needs at least one instance");
        const cles_solver_desc &d = *desc;
        CLES_REQUIRE(d.n_species >= 1 && d.n_species <= CLES_MAXSP, "n_species must be 1..8");
        CLES_REQUIRE(d.precision == CLES_F64 || d.precision == CLES_F32, "precision must be CLES_F64 or CLES_F32");
        CLES_REQUIRE(d.cfl > 0, "cfl must be positive");
        CLES_REQUIRE(d.beta_sensor >= 0 && d.beta_sensor <= 2,
                     "art.beta_sensor must be :strain, :gated_strain or :dilatation");
        const int lo = d.bc[0][0], hi = d.bc[0][1];
        CLES_REQUIRE((lo == CLES_BC_PERIODIC) == (hi == CLES_BC_PERIODIC),
                     "dimension mixes periodic and non-periodic conditions");
        for (int sd = 0; sd < 2; sd++)
            if (d.bc[0][sd] != CLES_BC_PERIODIC && d.bc[0][sd] != CLES_BC_SLIPWALL && d.bc[0][sd] != CLES_BC_DIRICHLET)
                throw cles::Error{CLES_ERR_UNSUPPORTED,
                                  "batched instances take periodic, slip-wall or Dirichlet ends"};
        for (int k = 1; k < 3; k++)
            CLES_REQUIRE(d.bc[k][0] == CLES_BC_PERIODIC && d.bc[k][1] == CLES_BC_PERIODIC,
                         "batched instances are one-dimensional: dimensions 2 and 3 are collapsed and periodic");
        cles::validate_eos(d);
        if (d.metric != CLES_METRIC_CARTESIAN)
            throw cles::Error{CLES_ERR_UNSUPPORTED, "batched instances take the Cartesian metric"};
        if (d.n_body_forces != 0)
            throw cles::Error{CLES_ERR_UNSUPPORTED, "batched instances take no sources"};
        if (d.retries != 0)
            throw cles::Error{CLES_ERR_UNSUPPORTED, "StepControl retries are not available for batched instances "
                                                    "(each instance stops at its own failure)"};
        if (!cles::rag_shape_ok(d.deriv) || !cles::rag_shape_ok(d.filt))
            throw cles::Error{CLES_ERR_UNSUPPORTED, "scheme shape not available for batched instances"};
        CLES_HIP(hipSetDevice(ctx->device));
        auto *b = new cles_batch;
        try {
            b->ctx = ctx;
            cles::ctx_retain(ctx);
            b->d = d;
            b->Ns = d.n_species;
            b->C = b->Ns + 4;
            b->prec = d.precision;
            b->count = count;
            b->periodic = lo == CLES_BC_PERIODIC;
            for (int k = 0; k < b->Ns; k++) {   // physics.jl:111-115
                b->Rk[k] = d.R[k];
                b->cvk[k] = d.eos_kind == CLES_EOS_NASA9 ? 0.0 : d.R[k] / (d.gamma[k] - 1);
                b->cpk[k] = b->Rk[k] + b->cvk[k];
            }
            if (d.eos_kind == CLES_EOS_NASA9) {
                const cles::N9Table t = cles::nasa9_table(d);
                auto *dt = (cles::N9Table *)b->alloc(sizeof(cles::N9Table));
                b->upload(dt, &t, sizeof t);
                b->n9 = dt;
            }
            b->off.resize(count);
            b->n.resize(count);
            b->h.resize(count);
            b->L.resize(count);
            b->origin.resize(count);
            std::vector<double> tfin(count), bcs((size_t)count * 2 * b->C);
            std::vector<int64_t> nmx(count);
            int64_t N = 0;
            for (int64_t i = 0; i < count; i++) {
                const cles_batch_instance &in = inst[i];
                if (in.n < 2 || in.n > 64 * cles::RAG_MAX_CHUNK)
                    throw cles::Error{CLES_ERR_INVALID, "instance " + std::to_string(i) + ": n must be 2.." +
                                                           std::to_string(64 * cles::RAG_MAX_CHUNK)};
                if (!(in.L > 0)) throw cles::Error{CLES_ERR_INVALID, "instance domain lengths must be positive"};
                b->off[i] = N;
                b->n[i] = (int)in.n;
                b->L[i] = in.L;
                b->origin[i] = in.origin;
                // rhs.jl:198-204
                b->h[i] = b->periodic ? in.L / (double)in.n : in.L / (double)(in.n - 1);
                tfin[i] = in.tfinal;
                nmx[i] = in.nmax;
                for (int sd = 0; sd < 2; sd++)
                    for (int c = 0; c < b->C; c++) bcs[((size_t)i * 2 + sd) * b->C + c] = in.bc_state[sd][c];
                N += in.n;
            }
            if (N >= 2147483648LL) throw cles::Error{CLES_ERR_UNSUPPORTED, "batches of 2^31 points or more"};
            b->N = N;

            // plans and work lists
            std::vector<unsigned char> blob;
            cles::PlanSet pd, pf;
            cles::build_plan_set(pd, blob, d.deriv, b->n, b->h, b->periodic);
            cles::build_plan_set(pf, blob, d.filt, b->n, b->h, b->periodic);
            b->tables = b->alloc(blob.size());
            CLES_HIP(hipMemcpyAsync(b->tables, blob.data(), blob.size(), hipMemcpyHostToDevice, b->st()));
            cles::rebase(pd.dev, (unsigned char *)b->tables);
            cles::rebase(pf.dev, (unsigned char *)b->tables);
            b->dplans = b->upload_vec(pd.dev);
            b->fplans = b->upload_vec(pf.dev);
            b->order_d = b->upload_vec(pd.order);
            b->items_d = b->upload_vec(pd.items);
            b->nitems_d = (int)(pd.items.size() / 3);
            b->nshort_d = pd.nshort;
            b->order_f = b->upload_vec(pf.order);
            b->items_f = b->upload_vec(pf.items);
            b->nitems_f = (int)(pf.items.size() / 3);
            b->nshort_f = pf.nshort;
            b->dq = d.deriv.q; b->dM = d.deriv.M; b->dsym = d.deriv.symmetric ? 1 : 0;
            b->fq = d.filt.q; b->fM = d.filt.M; b->fsym = d.filt.symmetric ? 1 : 0;

            b->g_off = b->upload_vec(b->off);
            b->g_n = b->upload_vec(b->n);
            b->g_h = b->upload_vec(b->h);

            const size_t NN = (size_t)N;
            b->Q = (double *)b->alloc(NN * b->C * sizeof(double));
            b->du = (double *)b->alloc(NN * b->C * sizeof(double));
            b->dQ = b->alloc(NN * b->C * b->wsize());
            // fields that one line launch treats together are contiguous:
            // u | v | w, T | Y_1.. | , grad_u, grad_T | grad_Y, fluxes
            const size_t ws = NN * b->wsize();
            auto blockf = [&](int k) { return (char *)b->alloc(ws * (size_t)k); };
            for (void **fp : {&b->rho, &b->p, &b->c, &b->cp}) *fp = b->wfield();
            char *vel = blockf(3);
            b->u = vel;
            b->v = vel + ws;
            b->w = vel + 2 * ws;
            char *ty = blockf(1 + b->Ns);
            b->T = ty;
            for (int k = 0; k < b->Ns; k++) b->Y[k] = ty + (size_t)(1 + k) * ws;
            char *gv = blockf(3);
            for (int j = 0; j < 3; j++) b->gu[j] = gv + (size_t)j * ws;
            char *gty = blockf(b->Ns > 1 ? 1 + b->Ns : 1);
            b->gT = gty;
            if (b->Ns > 1)
                for (int k = 0; k < b->Ns; k++) b->gY[k] = gty + (size_t)(1 + k) * ws;
            char *fl = blockf(b->C);
            for (int c = 0; c < b->C; c++) b->flux[c] = fl + (size_t)c * ws;
            b->mu_art = b->wfield();
            b->beta_art = b->wfield();
            b->kappa_art = b->wfield();
            if (b->Ns > 1)
                for (int k = 0; k < b->Ns; k++) b->D_art[k] = b->wfield();
            if (d.art_enabled) {
                b->nsens = 2 + (b->Ns > 1 ? b->Ns : 0);
                b->sensor = blockf(b->nsens);
                b->strain = b->wfield();
                b->esens = b->wfield();
                if (d.beta_sensor == 2) {
                    b->tmp1 = b->wfield();
                    b->tmp2 = b->wfield();
                }
            }
            if (d.art_enabled) {
                std::vector<cles::EpiArgs> ep((size_t)b->nsens);
                ep[0].kind = cles::EPI_ART_MUBETA;
                ep[0].out = b->mu_art;
                ep[0].out2 = b->beta_art;
                ep[0].a0 = b->rho;
                ep[0].c0 = d.C_mu;
                ep[0].c1 = d.C_beta;
                ep[1].kind = cles::EPI_ART_KAPPA;
                ep[1].out = b->kappa_art;
                ep[1].a0 = b->rho;
                ep[1].a1 = b->c;
                ep[1].a2 = b->T;
                ep[1].c0 = d.C_kappa;
                for (int k = 2; k < b->nsens; k++) {
                    ep[k].kind = cles::EPI_ART_D;
                    ep[k].out = b->D_art[k - 2];
                    ep[k].a1 = b->c;
                    ep[k].c0 = d.C_D;
                }
                b->art_multi = b->upload_vec(ep);
            }
            const size_t cnt = (size_t)count;
            for (double **fp : {&b->t, &b->dt, &b->dt_prev, &b->rate_prev, &b->dt_seen, &b->rate, &b->rhomin,
                                &b->ke})
                *fp = (double *)b->alloc(cnt * sizeof(double));
            b->tfinal = b->upload_vec(tfin);
            b->nmax = b->upload_vec(nmx);
            b->bcst = b->upload_vec(bcs);
            b->step = (int64_t *)b->alloc(cnt * sizeof(int64_t));
            b->step0 = (int64_t *)b->alloc(cnt * sizeof(int64_t));
            for (int **fp : {&b->status, &b->active, &b->stepped, &b->filt}) *fp = (int *)b->alloc(cnt * sizeof(int));
            b->cnt = (int *)b->alloc(sizeof(int));
            CLES_HIP(hipHostMalloc((void **)&b->host_cnt, sizeof(int), hipHostMallocDefault));
            CLES_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    })
}

int cles_batch_destroy(cles_batch *b)
{
    CLES_TRY({ delete b; })
}

int cles_batch_shape(cles_batch *b, int64_t *count, int64_t *total_points, int *n_cons)
{
    CLES_TRY({
        CLES_REQUIRE(b && count && total_points && n_cons, "NULL argument");
        *count = b->count;
        *total_points = b->N;
        *n_cons = b->C;
    })
}

int cles_batch_upload(cles_batch *b, int64_t i, const double *Qh, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(b && Qh && padded && halo, "NULL argument");
        cles::check_inst(b, i);
        CLES_REQUIRE(halo[0] >= 0 && padded[0] >= b->n[i] + 2 * (int64_t)halo[0], "padded extents smaller than n + 2*halo");
        CLES_REQUIRE(padded[1] == 1 && padded[2] == 1 && halo[1] == 0 && halo[2] == 0,
                     "a batch instance is laid out as (n + 2*halo, 1, 1) with halo (H, 0, 0)");
        const int64_t NP = padded[0] * padded[1] * padded[2];
        for (int c = 0; c < b->C; c++)
            CLES_HIP(hipMemcpyAsync(b->Q + (size_t)c * b->N + b->off[i], Qh + c * NP + halo[0],
                                    sizeof(double) * b->n[i], hipMemcpyHostToDevice, b->st()));
        CLES_HIP(hipStreamSynchronize(b->st()));
    })
}

int cles_batch_download(cles_batch *b, int64_t i, double *Qh, const int64_t padded[3], const int halo[3])
{
    CLES_TRY({
        CLES_REQUIRE(b && Qh && padded && halo, "NULL argument");
        cles::check_inst(b, i);
        CLES_REQUIRE(halo[0] >= 0 && padded[0] >= b->n[i] + 2 * (int64_t)halo[0], "padded extents smaller than n + 2*halo");
        CLES_REQUIRE(padded[1] == 1 && padded[2] == 1 && halo[1] == 0 && halo[2] == 0,
                     "a batch instance is laid out as (n + 2*halo, 1, 1) with halo (H, 0, 0)");
        const int64_t NP = padded[0] * padded[1] * padded[2];
        for (int c = 0; c < b->C; c++)
            CLES_HIP(hipMemcpyAsync(Qh + c * NP + halo[0], b->Q + (size_t)c * b->N + b->off[i],
                                    sizeof(double) * b->n[i], hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipStreamSynchronize(b->st()));
    })
}

namespace {

constexpr int GRAPH_ITERS = 8;   // iterations per replay, and between activity polls

// GRAPH_ITERS iterations of the batch captured on its stream (nothing runs
// during the capture) and instantiated into b->gexec.
extern "C++" template <typename Im>
void capture_iterations(cles_batch *b, Im &im, const cles::BatchCtl &c)
{
    b->drop_graph();
    hipError_t cap = hipStreamBeginCapture(b->st(), hipStreamCaptureModeThreadLocal);
    if (cap != hipSuccess) throw cles::Error{CLES_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(cap)};
    std::string err;
    try {
        for (int i = 0; i < GRAPH_ITERS; i++) im.iteration(c);
    } catch (const cles::Error &e) {
        err = e.msg;
    }
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(b->st(), &g);
    if (!err.empty() || end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw cles::Error{CLES_ERR_HIP, err.empty() ? std::string("hipStreamEndCapture: ") + hipGetErrorString(end) : err};
    }
    b->graph = g;
    const hipError_t inst = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    if (inst != hipSuccess) {
        b->drop_graph();
        throw cles::Error{CLES_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(inst)};
    }
}

// run! on every instance (cles_batch_run); with `noalloc` the dt history is
// staged in the per-device array of cles_noalloc.hip instead of a device
// allocation of this call.  In CLES_RUN_STEP_GRAPH mode (not with noalloc)
// whole groups of GRAPH_ITERS iterations are replays of one captured graph and
// the rest are issued from the host: the same kernels in the same order, so
// the results are those of host issue, bit for bit.
void batch_run(cles_batch *b, int64_t max_steps, int64_t nrec, double *dts, int64_t *taken, bool noalloc)
{
    {
        CLES_REQUIRE(b, "NULL argument");
        CLES_REQUIRE(nrec >= 0 && (nrec == 0 || dts), "dt history needs a host buffer");
        const size_t cnt = (size_t)b->count;
        if (noalloc && nrec > 0 && (int64_t)cnt > (int64_t)CLES_NOALLOC_DTS_MAX / nrec)
            throw cles::Error{CLES_ERR_INVALID, "count * nrec exceeds the allocation-free dt history "
                                                "(CLES_NOALLOC_DTS_MAX doubles)"};
        const bool graphs = !noalloc && b->run_mode == CLES_RUN_STEP_GRAPH;
        // the run! entry of every instance (timestep.jl:281-344): fresh
        // workspace, dt_seen = 0, active while t < tfinal and step < nmax
        std::vector<double> t(cnt), tf(cnt);
        std::vector<int64_t> stp(cnt), nmx(cnt);
        std::vector<int> status(cnt), active(cnt);
        CLES_HIP(hipMemcpyAsync(t.data(), b->t, cnt * sizeof(double), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipMemcpyAsync(tf.data(), b->tfinal, cnt * sizeof(double), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipMemcpyAsync(stp.data(), b->step, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipMemcpyAsync(nmx.data(), b->nmax, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipMemcpyAsync(status.data(), b->status, cnt * sizeof(int), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipStreamSynchronize(b->st()));
        int nact = 0;
        for (size_t i = 0; i < cnt; i++) {
            active[i] = status[i] == CLES_OK && t[i] < tf[i] && stp[i] < nmx[i];
            nact += active[i];
        }
        CLES_HIP(hipMemcpyAsync(b->active, active.data(), cnt * sizeof(int), hipMemcpyHostToDevice, b->st()));
        CLES_HIP(hipMemcpyAsync(b->step0, stp.data(), cnt * sizeof(int64_t), hipMemcpyHostToDevice, b->st()));
        CLES_HIP(hipMemsetAsync(b->dt_seen, 0, cnt * sizeof(double), b->st()));
        CLES_HIP(hipMemsetAsync(b->du, 0, (size_t)b->N * b->C * sizeof(double), b->st()));
        CLES_HIP(hipMemsetAsync(b->dQ, 0, (size_t)b->N * b->C * b->wsize(), b->st()));
        double *ddts = nullptr;
        if (nrec > 0 && noalloc) {
            CLES_HIP(hipSetDevice(b->ctx->device));
            ddts = cles::noalloc_dts(b->ctx->device, (int64_t)cnt * nrec, b->st());
        } else if (nrec > 0 && graphs) {
            // the captured kernels write the history into the batch's own
            // buffer, which outlives this call
            const size_t need = cnt * (size_t)nrec;
            if (need > b->gdts_cap) {
                b->drop_graph();
                if (b->gdts) CLES_HIP(hipFree(b->gdts));
                b->gdts = nullptr;
                b->gdts_cap = 0;
                CLES_HIP(hipMalloc(&b->gdts, need * sizeof(double)));
                b->gdts_cap = need;
            }
            ddts = b->gdts;
            CLES_HIP(hipMemsetAsync(ddts, 0, need * sizeof(double), b->st()));
        } else if (nrec > 0) {
            CLES_HIP(hipMalloc(&ddts, cnt * (size_t)nrec * sizeof(double)));
            CLES_HIP(hipMemsetAsync(ddts, 0, cnt * (size_t)nrec * sizeof(double), b->st()));
        }
        const bool own_dts = ddts && !noalloc && !graphs;
        if (graphs && nrec != b->g_nrec) b->drop_graph();
        int64_t steps = 0;
        try {
            cles::bdispatch(b, [&](auto &im) {
                const cles::BatchCtl c = im.ctl(ddts, nrec);
                const int64_t poll = GRAPH_ITERS;
                while (nact > 0 && steps < max_steps) {
                    if (graphs && max_steps - steps >= poll && steps % poll == 0) {
                        if (!b->gexec) {
                            capture_iterations(b, im, c);
                            b->g_nrec = nrec;
                        }
                        CLES_HIP(hipGraphLaunch(b->gexec, b->st()));
                        steps += poll;
                    } else {
                        im.iteration(c);
                        steps++;
                    }
                    if (steps % poll == 0 || steps == max_steps) {
                        cles::launch_batch_count(b->active, b->count, b->cnt, b->st());
                        CLES_HIP(hipMemcpyAsync(b->host_cnt, b->cnt, sizeof(int), hipMemcpyDeviceToHost, b->st()));
                        CLES_HIP(hipStreamSynchronize(b->st()));
                        nact = *b->host_cnt;
                    }
                }
                CLES_HIP(hipStreamSynchronize(b->st()));
            });
            if (ddts) {
                CLES_HIP(hipMemcpy(dts, ddts, cnt * (size_t)nrec * sizeof(double), hipMemcpyDeviceToHost));
                if (own_dts) (void)hipFree(ddts);
                ddts = nullptr;
            }
        } catch (...) {
            if (own_dts && ddts) (void)hipFree(ddts);
            throw;
        }
        if (taken) *taken = steps;
    }
}

}  // namespace

int cles_batch_run(cles_batch *b, int64_t max_steps, int64_t nrec, double *dts, int64_t *taken)
{
    CLES_TRY({ batch_run(b, max_steps, nrec, dts, taken, false); })
}

int cles_batch_run_noalloc(cles_batch *b, int64_t max_steps, int64_t nrec, double *dts, int64_t *taken)
{
    CLES_TRY({ batch_run(b, max_steps, nrec, dts, taken, true); })
}

int cles_batch_set_run_mode(cles_batch *b, int mode)
{
    CLES_TRY({
        CLES_REQUIRE(b, "NULL argument");
        CLES_REQUIRE(mode >= CLES_RUN_HOST_DT && mode <= CLES_RUN_STEP_GRAPH, "unknown run mode");
        if (mode != CLES_RUN_STEP_GRAPH) b->drop_graph();
        b->run_mode = mode;
    })
}

int cles_batch_state(cles_batch *b, double *t, int64_t *steps, int *status)
{
    CLES_TRY({
        CLES_REQUIRE(b, "NULL argument");
        const size_t cnt = (size_t)b->count;
        if (t) CLES_HIP(hipMemcpyAsync(t, b->t, cnt * sizeof(double), hipMemcpyDeviceToHost, b->st()));
        if (steps) CLES_HIP(hipMemcpyAsync(steps, b->step, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, b->st()));
        if (status) CLES_HIP(hipMemcpyAsync(status, b->status, cnt * sizeof(int), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipStreamSynchronize(b->st()));
    })
}

int cles_batch_compute_rhs(cles_batch *b, int primitives_current)
{
    CLES_TRY({
        CLES_REQUIRE(b, "NULL argument");
        cles::bdispatch(b, [&](auto &im) { im.rhs(primitives_current != 0, false, 0.0, 0.0); });
        CLES_HIP(hipStreamSynchronize(b->st()));
    })
}

int cles_batch_get_field(cles_batch *b, int64_t i, const char *name, double *host)
{
    CLES_TRY({
        CLES_REQUIRE(b && name && host, "NULL argument");
        cles::check_inst(b, i);
        std::string nm(name);
        const void *src = nullptr;
        bool is_double = false;
        int a = 0;
        if (nm == "rho") src = b->rho;
        else if (nm == "u") src = b->u;
        else if (nm == "p") src = b->p;
        else if (nm == "T") src = b->T;
        else if (nm == "c") src = b->c;
        else if (nm == "mu_art") src = b->mu_art;
        else if (nm == "beta_art") src = b->beta_art;
        else if (nm == "kappa_art") src = b->kappa_art;
        else if (std::sscanf(name, "Q[%d]", &a) == 1 && a >= 1 && a <= b->C) {
            src = b->Q + (size_t)(a - 1) * b->N;
            is_double = true;
        } else if (std::sscanf(name, "dQ[%d]", &a) == 1 && a >= 1 && a <= b->C)
            src = (const char *)b->dQ + (size_t)(a - 1) * b->N * b->wsize();
        else if (std::sscanf(name, "Y[%d]", &a) == 1 && a >= 1 && a <= b->Ns)
            src = b->Y[a - 1];
        else if (std::sscanf(name, "D_art[%d]", &a) == 1 && a >= 1 && a <= b->Ns)
            src = b->D_art[a - 1];
        if (!src) throw cles::Error{CLES_ERR_INVALID, "unknown or unallocated field '" + nm + "'"};
        const size_t n = (size_t)b->n[i];
        if (is_double || b->prec == CLES_F64) {
            CLES_HIP(hipMemcpyAsync(host, (const double *)src + b->off[i], n * sizeof(double), hipMemcpyDeviceToHost,
                                    b->st()));
            CLES_HIP(hipStreamSynchronize(b->st()));
        } else {
            std::vector<float> tmp(n);
            CLES_HIP(hipMemcpyAsync(tmp.data(), (const float *)src + b->off[i], n * sizeof(float),
                                    hipMemcpyDeviceToHost, b->st()));
            CLES_HIP(hipStreamSynchronize(b->st()));
            for (size_t k = 0; k < n; k++) host[k] = tmp[k];
        }
    })
}

int cles_batch_kinetic_energy(cles_batch *b, double *ke)
{
    CLES_TRY({
        CLES_REQUIRE(b && ke, "NULL argument");
        cles::BatchGeom g{};
        g.count = b->count;
        g.N = b->N;
        g.off = b->g_off;
        g.n = b->g_n;
        g.h = b->g_h;
        g.periodic = b->periodic;
        cles::launch_ke_batch(b->Q, b->Ns, g, b->ke, b->st());
        CLES_HIP(hipMemcpyAsync(ke, b->ke, (size_t)b->count * sizeof(double), hipMemcpyDeviceToHost, b->st()));
        CLES_HIP(hipStreamSynchronize(b->st()));
    })
}

}  // extern "C"
