// Line-tile autotuner for the solver's line operators.
//
// The chunk length of a direction's plans (points per thread; it fixes the
// chunks per line and, through them, which line kernel runs and its block
// shape) trades sequential work per thread against waves per launch: long
// chunks suit large grids, short chunks keep small grids from running a
// handful of waves.  Solver creation takes the length of each direction from
// a process-wide table, filled by cles_tune_solver (measured on the device)
// or cles_tune_load (a saved table), and from a thread-count rule for shapes
// the table does not hold.  Nothing is tuned implicitly: one table always
// gives the same plans, so results do not depend on timing or on what else
// runs on the device.
//
// The reference has no counterpart: its lines are solved serially per rank
// (src/tridiag.jl:31-57, src/banded.jl:38-80).
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cles_internal.h"

namespace cles {

namespace {

// (dim 1..3, n_x, n_y, n_z, precision, periodic, deriv q, deriv M, filter M)
using TuneKey = std::tuple<int, int64_t, int64_t, int64_t, int, int, int, int, int>;

std::mutex tune_mu;
std::map<TuneKey, int> &tune_table()
{
    static std::map<TuneKey, int> t;
    return t;
}

TuneKey key_of(int dim, const cles_solver_desc &d)
{
    const int k = dim - 1;
    const int periodic = d.bc[k][0] == CLES_BC_PERIODIC ? 1 : 0;
    return TuneKey{dim, d.n_global[0], d.n_global[1], d.n_global[2], d.precision, periodic, d.deriv.q, d.deriv.M,
                   d.filt.M};
}

constexpr int64_t TARGET_THREADS = 131072;   // 2048 waves: eight per CU

int cap_of(int dim) { return dim == 1 ? 16 : 32; }
constexpr int MAX_TILE = 32;   // the kernels' per-thread chunk capacity

}  // namespace

int default_line_chunk(int dim, int64_t n_line, int64_t lines)
{
    for (int c = cap_of(dim); c >= 8; c /= 2)
        if (lines * ((n_line + c - 1) / c) >= TARGET_THREADS) return c;
    return 8;
}

int line_chunk(int dim, const cles_solver_desc &d, bool *tuned)
{
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        auto it = tune_table().find(key_of(dim, d));
        if (it != tune_table().end()) {
            if (tuned) *tuned = true;
            return it->second;
        }
    }
    if (tuned) *tuned = false;
    int64_t N = 1;
    for (int k = 0; k < 3; k++) N *= d.n_global[k];
    const int64_t n = d.n_global[dim - 1];
    return default_line_chunk(dim, n, N / n);
}

}  // namespace cles

namespace {

// Mean milliseconds of the representative line work of one direction with
// chunk length c: one derivative launch over four fields (the gradient group)
// and two single-field filter passes (sensor smoothing), as the solver issues
// them.
double time_direction(cles_ctx *ctx, const cles_solver_desc &d, int k, int c, int reps, void *buf, int64_t N)
{
    using namespace cles;
    const int dtype = d.precision;
    const size_t es = dtype == CLES_F64 ? 8 : 4;
    const int periodic = d.bc[k][0] == CLES_BC_PERIODIC ? 1 : 0;
    const double h = periodic ? d.L_domain[k] / (double)d.n_global[k] : d.L_domain[k] / (double)(d.n_global[k] - 1);
    HostPlan dp, fp;
    for (HostPlan *hp : {&dp, &fp}) {
        hp->dim = k + 1;
        hp->dtype = dtype;
        for (int j = 0; j < 3; j++) hp->nloc[j] = d.n_global[j];
    }
    double ms = -1.0;
    try {
        build_line_plan(dp, d.deriv, (int)d.n_global[k], periodic, h, c);
        build_line_plan(fp, d.filt, (int)d.n_global[k], periodic, h, c);
        View3 v;
        for (int j = 0; j < 3; j++) v.n[j] = d.n_global[j];
        v.s[0] = 1;
        v.s[1] = v.n[0];
        v.s[2] = v.n[0] * v.n[1];
        unsigned char *b = (unsigned char *)buf;
        LineFields mf;
        mf.nf = 4;
        for (int i = 0; i < 4; i++) {
            mf.in[i] = b + (size_t)i * N * es;
            mf.out[i] = b + (size_t)(4 + i) * N * es;
        }
        EpiArgs st{};
        st.kind = EPI_STORE;
        st.out = mf.out[0];
        EpiArgs sf{};
        sf.kind = EPI_STORE;
        sf.out = b;
        hipStream_t s = ctx->stream;
        auto once = [&] {
            launch_line(dp, k + 1, v, mf.in[0], st, s, &mf);
            launch_line(fp, k + 1, v, b, sf, s);
            launch_line(fp, k + 1, v, b, sf, s);
        };
        once();
        once();
        hipEvent_t e0 = nullptr, e1 = nullptr;
        CLES_HIP(hipEventCreate(&e0));
        CLES_HIP(hipEventCreate(&e1));
        CLES_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) once();
        CLES_HIP(hipEventRecord(e1, s));
        CLES_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        CLES_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ms = (double)t / reps;
    } catch (...) {
        free_line_plan(dp);
        free_line_plan(fp);
        throw;
    }
    free_line_plan(dp);
    free_line_plan(fp);
    return ms;
}

}  // namespace

extern "C" {

int cles_tune_solver(cles_ctx *ctx, const cles_solver_desc *desc, int reps, int chunks[3], double ms_best[3])
{
    CLES_TRY({
        CLES_REQUIRE(ctx && desc, "NULL argument");
        CLES_REQUIRE(reps >= 1 && reps <= 1000, "reps must be 1..1000");
        size_t fp = 0;
        const int rc = cles_solver_footprint(desc, &fp);   // validates the descriptor
        if (rc != CLES_OK) throw cles::Error{rc, cles_last_error()};
        CLES_HIP(hipSetDevice(ctx->device));
        int64_t N = 1;
        for (int k = 0; k < 3; k++) N *= desc->n_global[k];
        const size_t es = desc->precision == CLES_F64 ? 8 : 4;
        void *buf = nullptr;
        CLES_HIP(hipMalloc(&buf, (size_t)8 * N * es));
        try {
            CLES_HIP(hipMemsetAsync(buf, 0, (size_t)8 * N * es, ctx->stream));
            for (int k = 0; k < 3; k++) {
                if (chunks) chunks[k] = 0;
                if (ms_best) ms_best[k] = 0.0;
                if (desc->n_global[k] <= 1) continue;
                int best = 0, lastP = -1;
                double tbest = 0.0;
                for (int c = cles::cap_of(k + 1); c >= 4; c /= 2) {
                    // lengths that give the same chunk count build the same plan
                    int P = 1;
                    while ((desc->n_global[k] + P - 1) / P > c) P *= 2;
                    if (P == lastP) continue;
                    lastP = P;
                    double t = 0.0;
                    try {
                        t = time_direction(ctx, *desc, k, c, reps, buf, N);
                    } catch (const cles::Error &e) {
                        if (e.code == CLES_ERR_UNSUPPORTED) continue;   // no kernel for this shape
                        throw;
                    }
                    if (best == 0 || t < tbest) {
                        best = c;
                        tbest = t;
                    }
                }
                CLES_REQUIRE(best > 0, "no chunk length is supported for this line shape");
                {
                    std::lock_guard<std::mutex> lk(cles::tune_mu);
                    cles::tune_table()[cles::key_of(k + 1, *desc)] = best;
                }
                if (chunks) chunks[k] = best;
                if (ms_best) ms_best[k] = tbest;
            }
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        CLES_HIP(hipFree(buf));
    })
}

int cles_tune_lookup(const cles_solver_desc *desc, int chunks[3], int tuned[3])
{
    CLES_TRY({
        CLES_REQUIRE(desc && chunks, "NULL argument");
        for (int k = 0; k < 3; k++) {
            bool t = false;
            chunks[k] = desc->n_global[k] > 1 ? cles::line_chunk(k + 1, *desc, &t) : 0;
            if (tuned) tuned[k] = t ? 1 : 0;
        }
    })
}

int cles_tune_set(const cles_solver_desc *desc, const int chunks[3])
{
    CLES_TRY({
        CLES_REQUIRE(desc && chunks, "NULL argument");
        for (int k = 0; k < 3; k++)
            CLES_REQUIRE(chunks[k] <= cles::MAX_TILE, "chunk lengths above 32 points are not supported");
        std::lock_guard<std::mutex> lk(cles::tune_mu);
        for (int k = 0; k < 3; k++) {
            if (desc->n_global[k] <= 1 || chunks[k] < 0) continue;
            const auto key = cles::key_of(k + 1, *desc);
            if (chunks[k] == 0)
                cles::tune_table().erase(key);
            else
                cles::tune_table()[key] = chunks[k];
        }
    })
}

int cles_tune_clear(void)
{
    CLES_TRY({
        std::lock_guard<std::mutex> lk(cles::tune_mu);
        cles::tune_table().clear();
    })
}

int cles_tune_save(const char *path)
{
    CLES_TRY({
        CLES_REQUIRE(path, "NULL argument");
        std::lock_guard<std::mutex> lk(cles::tune_mu);
        FILE *fh = std::fopen(path, "w");
        if (!fh) throw cles::Error{CLES_ERR_INVALID, std::string("cannot open ") + path};
        std::fprintf(fh, "# compactles line-tile table: dim nx ny nz precision periodic deriv_q deriv_M filt_M chunk\n");
        for (const auto &kv : cles::tune_table()) {
            const auto &k = kv.first;
            std::fprintf(fh, "%d %lld %lld %lld %d %d %d %d %d %d\n", std::get<0>(k), (long long)std::get<1>(k),
                         (long long)std::get<2>(k), (long long)std::get<3>(k), std::get<4>(k), std::get<5>(k),
                         std::get<6>(k), std::get<7>(k), std::get<8>(k), kv.second);
        }
        const bool bad = std::ferror(fh) != 0;
        if (std::fclose(fh) != 0 || bad) throw cles::Error{CLES_ERR_INVALID, std::string("error writing ") + path};
    })
}

int cles_tune_load(const char *path)
{
    CLES_TRY({
        CLES_REQUIRE(path, "NULL argument");
        FILE *fh = std::fopen(path, "r");
        if (!fh) throw cles::Error{CLES_ERR_INVALID, std::string("cannot open ") + path};
        std::vector<std::pair<cles::TuneKey, int>> rows;
        char line[512];
        int lineno = 0;
        while (std::fgets(line, sizeof line, fh)) {
            lineno++;
            if (line[0] == '#' || line[0] == '\n') continue;
            int dim, prec, per, q, M, Mf, chunk;
            long long nx, ny, nz;
            if (std::sscanf(line, "%d %lld %lld %lld %d %d %d %d %d %d", &dim, &nx, &ny, &nz, &prec, &per, &q, &M, &Mf,
                            &chunk) != 10 ||
                dim < 1 || dim > 3 || chunk < 1 || chunk > cles::MAX_TILE) {
                std::fclose(fh);
                throw cles::Error{CLES_ERR_INVALID, std::string(path) + ":" + std::to_string(lineno) +
                                                        ": malformed line-tile entry"};
            }
            rows.push_back({cles::TuneKey{dim, nx, ny, nz, prec, per, q, M, Mf}, chunk});
        }
        std::fclose(fh);
        std::lock_guard<std::mutex> lk(cles::tune_mu);
        for (const auto &r : rows) cles::tune_table()[r.first] = r.second;
    })
}

}  // extern "C"
