// extern "C" entry points: context, device memory, operator plans.
// Solver entry points live in cles_solver.cpp.
#include <cstring>
#include <string>

#include "cles_internal.h"
#include "cles_point.h"

namespace cles {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

void ctx_retain(cles_ctx *ctx) { ctx->refs++; }

void ctx_release(cles_ctx *ctx)
{
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

}  // namespace cles

extern "C" {

const char *cles_last_error(void) { return cles::g_last_error.c_str(); }

int cles_abi_version(void) { return CLES_ABI_VERSION; }

int cles_ctx_create(int device, cles_ctx **out)
{
    CLES_TRY({
        CLES_REQUIRE(out, "ctx out pointer is NULL");
        int ndev = 0;
        CLES_HIP(hipGetDeviceCount(&ndev));
        CLES_REQUIRE(device >= 0 && device < ndev, "device index out of range");
        CLES_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        CLES_HIP(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw cles::Error{CLES_ERR_UNSUPPORTED,
                              std::string("this library is built for gfx950, device is ") + prop.gcnArchName};
        cles::noalloc_setup(device);
        auto *c = new cles_ctx;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw cles::Error{CLES_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        *out = c;
    })
}

int cles_ctx_destroy(cles_ctx *ctx)
{
    CLES_TRY({
        if (!ctx) return CLES_OK;
        cles::ctx_release(ctx);   // deferred while plans or solvers still use it
    })
}

void *cles_ctx_stream(cles_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int cles_malloc(cles_ctx *ctx, size_t bytes, void **dptr)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && dptr, "NULL argument");
        CLES_HIP(hipSetDevice(ctx->device));
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e != hipSuccess) throw cles::Error{CLES_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
        *dptr = p;
    })
}

int cles_free(cles_ctx *ctx, void *dptr)
{
    CLES_TRY({
        CLES_REQUIRE(ctx, "NULL context");
        if (dptr) {
            CLES_HIP(hipStreamSynchronize(ctx->stream));
            CLES_HIP(hipFree(dptr));
        }
    })
}

int cles_memcpy_h2d(cles_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && dst && src, "NULL argument");
        CLES_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        CLES_HIP(hipStreamSynchronize(ctx->stream));
    })
}

int cles_memcpy_d2h(cles_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && dst && src, "NULL argument");
        CLES_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CLES_HIP(hipStreamSynchronize(ctx->stream));
    })
}

int cles_memset(cles_ctx *ctx, void *dptr, int value, size_t bytes)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && dptr, "NULL argument");
        CLES_HIP(hipMemsetAsync(dptr, value, bytes, ctx->stream));
        CLES_HIP(hipStreamSynchronize(ctx->stream));
    })
}

int cles_memcpy_d2d(cles_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && dst && src, "NULL argument");
        cles::launch_copy(dst, src, bytes, ctx->stream);
        CLES_HIP(hipStreamSynchronize(ctx->stream));
    })
}

int cles_synchronize(cles_ctx *ctx)
{
    CLES_TRY({
        CLES_REQUIRE(ctx, "NULL context");
        CLES_HIP(hipStreamSynchronize(ctx->stream));
    })
}

int cles_plan_create(cles_ctx *ctx, const cles_scheme_desc *scheme, int dim, const int64_t n_local[3],
                     int periodic, double h, int dtype, int lo_fold, int hi_fold, cles_plan **out)
{
    CLES_TRY({
        CLES_REQUIRE(ctx && scheme && n_local && out, "NULL argument");
        CLES_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1, 2 or 3");
        CLES_REQUIRE(dtype == CLES_F64 || dtype == CLES_F32, "dtype must be CLES_F64 or CLES_F32");
        CLES_REQUIRE(h > 0, "grid spacing h must be positive");
        for (int d = 0; d < 3; d++) CLES_REQUIRE(n_local[d] >= 1, "extents must be >= 1");
        CLES_REQUIRE(n_local[dim - 1] > 1, "plan along a collapsed dimension");
        CLES_REQUIRE(n_local[dim - 1] <= (1 << 20), "line longer than 2^20 points");
        CLES_REQUIRE(lo_fold >= -1 && lo_fold <= 1 && hi_fold >= -1 && hi_fold <= 1,
                     "lo_fold / hi_fold must be 0 (none), +1 or -1");
        // a periodic dimension has no physical edge, so the keywords do nothing
        // there (at_lo_edge / at_hi_edge, decomposition.jl:137-140)
        if (periodic) lo_fold = hi_fold = 0;
        CLES_HIP(hipSetDevice(ctx->device));
        auto *p = new cles_plan;
        p->ctx = ctx;
        p->hp.dim = dim;
        p->hp.dtype = dtype;
        for (int d = 0; d < 3; d++) p->hp.nloc[d] = n_local[d];
        try {
            cles::build_line_plan(p->hp, *scheme, (int)n_local[dim - 1], periodic, h, dim == 1 ? 16 : 32,
                                  lo_fold, hi_fold);
        } catch (...) {
            delete p;
            throw;
        }
        cles::ctx_retain(ctx);
        *out = p;
    })
}

int cles_plan_destroy(cles_plan *plan)
{
    CLES_TRY({
        if (!plan) return CLES_OK;
        (void)hipStreamSynchronize(plan->ctx->stream);
        cles::free_line_plan(plan->hp);
        cles_ctx *ctx = plan->ctx;
        delete plan;
        cles::ctx_release(ctx);
    })
}

static cles::View3 make_view(const cles_plan *plan, const int64_t padded[3], const int halo[3], int64_t *origin)
{
    cles::View3 v;
    for (int d = 0; d < 3; d++) {
        v.n[d] = plan->hp.nloc[d];
        if (padded[d] < v.n[d] + 2 * (int64_t)halo[d] || halo[d] < 0)
            throw cles::Error{CLES_ERR_INVALID, "padded extents smaller than n_local + 2*halo"};
    }
    v.s[0] = 1;
    v.s[1] = padded[0];
    v.s[2] = padded[0] * padded[1];
    *origin = halo[0] + halo[1] * v.s[1] + halo[2] * v.s[2];
    return v;
}

int cles_apply_along(cles_plan *plan, void *out, const void *f, const int64_t padded[3], const int halo[3],
                     void *stream)
{
    CLES_TRY({
        CLES_REQUIRE(plan && out && f && padded && halo, "NULL argument");
        int64_t org = 0;
        cles::View3 v = make_view(plan, padded, halo, &org);
        size_t es = plan->hp.dtype == CLES_F64 ? 8 : 4;
        cles::EpiArgs e{};
        e.kind = cles::EPI_STORE;
        e.out = (char *)out + org * es;
        hipStream_t st = stream ? (hipStream_t)stream : plan->ctx->stream;
        cles::launch_line(plan->hp, plan->hp.dim, v, (const char *)f + org * es, e, st);
    })
}

int cles_apply_along_timed(cles_plan *plan, void *out, const void *f, const int64_t padded[3],
                           const int halo[3], int iters, double *ms)
{
    CLES_TRY({
        CLES_REQUIRE(plan && out && f && padded && halo && ms && iters > 0, "bad argument");
        int64_t org = 0;
        cles::View3 v = make_view(plan, padded, halo, &org);
        size_t es = plan->hp.dtype == CLES_F64 ? 8 : 4;
        cles::EpiArgs e{};
        e.kind = cles::EPI_STORE;
        e.out = (char *)out + org * es;
        hipStream_t st = plan->ctx->stream;
        hipEvent_t a, b;
        CLES_HIP(hipEventCreate(&a));
        CLES_HIP(hipEventCreate(&b));
        CLES_HIP(hipEventRecord(a, st));
        for (int i = 0; i < iters; i++)
            cles::launch_line(plan->hp, plan->hp.dim, v, (const char *)f + org * es, e, st);
        CLES_HIP(hipEventRecord(b, st));
        CLES_HIP(hipEventSynchronize(b));
        float t = 0.f;
        CLES_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms = (double)t / iters;
    })
}

}  // extern "C"
