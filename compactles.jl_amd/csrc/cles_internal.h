// Internal declarations shared by the host plan builder, the kernels and the
// C-ABI layer.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "compactles_hip.h"

namespace cles {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);

// The allocation-free entry points' per-device array (cles_noalloc.hip):
// noalloc_setup looks its address up on `device` (idempotent); noalloc_dts
// returns it cleared over `doubles` entries on `stream`, or throws
// CLES_ERR_INVALID beyond CLES_NOALLOC_DTS_MAX.
void noalloc_setup(int device);
double *noalloc_dts(int device, int64_t doubles, hipStream_t stream);
struct Error {
    int code;
    std::string msg;
};

#define CLES_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw ::cles::Error{CLES_ERR_HIP, std::string(#call) + ": " +           \
                                                   hipGetErrorString(e_)};          \
    } while (0)

// ---------------------------------------------------------------------------
// Line-operator plan in device form.
//
// A line of n points is split into P = 2^p chunks of balanced length
// (m or m+1).  Each chunk is factorised on its own (Thomas for q = 1,
// unpivoted banded LU for q = 2, the algorithms of src/tridiag.jl:31-57 and
// src/banded.jl:38-80 applied per chunk); the coupling between chunks is the
// spike / reduced-interface system of src/tridiag.jl:1-22 with the chunks in
// the role of ranks.  The reduced block-tridiagonal (cyclic when periodic)
// system in the 2q interface unknowns per chunk is solved by block parallel
// cyclic reduction whose coefficient blocks depend only on the scheme and are
// precomputed here.  P = 1 is the reference's own single-rank algorithm.
// ---------------------------------------------------------------------------
enum RowField1 { R1_L = 0, R1_DINV, R1_C, R1_V, R1_W, R1_N };
enum RowField2 { R2_L1 = 0, R2_L2, R2_U0INV, R2_U1, R2_U2, R2_V0, R2_V1, R2_W0, R2_W1, R2_N };

struct LineDev {
    int n;          // points per line
    int q;          // LHS half-bandwidth
    int M;          // RHS half-width
    int sym;        // symmetric (filter) RHS
    int nc;         // closure rows per closed edge
    int periodic;
    int closed;     // both edges closed (one rank, non-periodic)
    int P;          // chunks per line (power of two)
    int nsteps;     // log2(P)
    int m;          // base chunk length
    int extra;      // first `extra` chunks have length m + 1
    const double *coef64;  // [0..8) ci, [8] a0, [16..80) clo[8][8], [80..144) chi[8][8]
    const float *coef32;
    const double *row64;   // per-row constants, [field * n + g]
    const float *row32;
    const double *pcr;     // per step s, chunk k: alpha, beta (B*B each, row-major);
                           // then per chunk the final inverse block
    const double *rowt64;  // chunk-major copy of the per-row constants (ragged plans):
    const float *rowt32;   //   [(field * RAG_MAX_CHUNK + t) * P + k] = row of point t of chunk k
};

constexpr int RAG_MAX_CHUNK = 32;   // points per thread in the ragged kernel
constexpr int RAG_SHORT_CHUNK = 16;   // chunk bound of the short class (lines of at most 64 x 16 points)
constexpr int COEF_CI = 0, COEF_A0 = 8, COEF_CLO = 16, COEF_CHI = 80, COEF_N = 144;

struct HostPlan {
    LineDev dev{};
    int dim = 1;
    int64_t nloc[3] = {1, 1, 1};
    int dtype = CLES_F64;
    double h = 1.0;
    void *dbuf = nullptr;   // one device allocation holding every table
    size_t dbytes = 0;
};

// Host image of one line plan's tables; `dev` holds the shape fields and the
// pointers are bound to a device copy of `image` by bind_line_tables.
struct LineTables {
    LineDev dev{};
    std::vector<unsigned char> image;
    size_t o_coef32 = 0, o_row64 = 0, o_row32 = 0, o_pcr = 0, o_rowt64 = 0, o_rowt32 = 0;
};
// chunk_major: also store the chunk-major row tables of the ragged kernel
// (RAG_MAX_CHUNK rows per chunk; max_chunk must not exceed it).
// lo_fold (0: none, +1 / -1: the solution parity sigma_g) replaces the low
// edge's closure by a parity fold on a half-offset grid (operators.jl:59-69,
// operators_banded.jl:55-64): the LHS coupling to ghost unknown j < 0 moves to
// interior column -j-1 with sign sigma_g, and the low-edge RHS rows read the
// mirror f(-j-1) sigma_f of each ghost (sigma_f = sigma_g for filters, -sigma_g
// for derivatives), carried as synthesised closure rows.
// hi_fold does the same at the high edge (ghost n + j mirrors n - 1 - j).
void build_line_tables(LineTables &out, const cles_scheme_desc &s, int n, int periodic, double h, int max_chunk,
                       int dim, bool chunk_major = false, int lo_fold = 0, int hi_fold = 0);
void bind_line_tables(LineDev &dev, const LineTables &t, unsigned char *base);

// Builds the device tables for one (scheme, n, periodic, h).  Throws Error.
void build_line_plan(HostPlan &hp, const cles_scheme_desc &s, int n, int periodic,
                     double h, int max_chunk, int lo_fold = 0, int hi_fold = 0);
void free_line_plan(HostPlan &hp);

// ---------------------------------------------------------------------------
// Array views: element (i, j, k) of the interior is at base[i*s0 + j*s1 + k*s2]
// ---------------------------------------------------------------------------
struct View3 {
    int64_t n[3];
    int64_t s[3];
};

// ---------------------------------------------------------------------------
// Line-operator epilogues: what to do with the solved value v at element
// offset o (offsets are shared by every array of one view).
// ---------------------------------------------------------------------------
enum EpiKind {
    EPI_STORE = 0,      // out[o] = v
    EPI_STORE_NEG,      // out[o] = -v
    EPI_SUB,            // out[o] -= v
    EPI_LSRK,           // dq = acc[o] - v; du = A du + dt dq; Q += B du   (FP64 state)
    EPI_ART_MUBETA,     // rs = rho max(v,0); mu = C0 rs; beta = C1 rs
    EPI_ART_KAPPA,      // kappa = C0 (rho c / max(T, eps)) max(v,0)
    EPI_ART_D,          // D = C0 c max(v,0)
    EPI_LSRK_DEV,       // EPI_LSRK with dt = *dtp, and no store where *gate == 0
    EPI_STORE_GATE,     // EPI_STORE, and no store where *gate == 0
    EPI_ART_MULTI,      // field i of the launch applies the ART_* epilogue multi[i]
    EPI_NKIND
};

struct EpiArgs {
    int kind;
    void *out;          // typed by the operator dtype (STORE/NEG/SUB/ART_*)
    void *out2;         // beta (ART_MUBETA)
    const void *a0;     // rho (ART_*), acc dQ (LSRK)
    const void *a1;     // c (ART_KAPPA, ART_D)
    const void *a2;     // T (ART_KAPPA)
    double *du;         // LSRK
    double *Q;          // LSRK
    double c0, c1;      // coefficients (art) or A, B (LSRK)
    double dt;          // LSRK
    const double *dtp;  // EPI_LSRK_DEV: dt in device memory
    const int *gate;    // EPI_LSRK_DEV, EPI_STORE_GATE: store only when *gate != 0
    unsigned ipts;      //   ... or, when ipts > 0, only where gate[o / ipts] != 0 with dt = dtp[o / ipts]
                        //   (instances of ipts points side by side)
    const EpiArgs *multi;   // EPI_ART_MULTI: one ART_* epilogue per field (device memory)
};

// Several independent fields through one line operator in one launch (grid
// dimension y = field).  Field 0 is the launch's own input and epilogue; field
// i >= 1 reads in[i] and stores through out[i] (STORE, STORE_NEG, SUB,
// STORE_GATE), or, for the LSRK kinds, offsets the accumulator, du and Q by
// eoff[i] elements, or (ART_MULTI) takes the whole epilogue multi[i].  Small grids gain waves per launch and every grid loses
// the per-launch gaps.
constexpr int LINE_MF = 12;
struct LineFields {
    int nf = 1;
    const void *in[LINE_MF] = {};
    void *out[LINE_MF] = {};
    int64_t eoff[LINE_MF] = {};
};

// Launch one line operator along `dim` (1..3) over the view (mf: more fields).
void launch_line(const HostPlan &hp, int dim, const View3 &v, const void *f,
                 const EpiArgs &epi, hipStream_t st, const LineFields *mf = nullptr);

// timestep.jl:5-21 LSRK(5,4) coefficient tables (Carpenter & Kennedy)
inline constexpr double LSRK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                       -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
inline constexpr double LSRK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                       1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                       2277821191437.0 / 14882151754819.0};
inline constexpr double LSRK_C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                       2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

// Chunk length (points per thread) of the solver's line plans along dim
// (1..3): the line-tile table entry for this descriptor when one exists
// (*tuned = true), otherwise default_line_chunk.
int line_chunk(int dim, const cles_solver_desc &d, bool *tuned = nullptr);
// The largest chunk length (x: <= 16, y/z: <= 32, >= 8) that still gives a
// launch of at least 131072 threads.
int default_line_chunk(int dim, int64_t n_line, int64_t lines);

// Ragged batch of independent lines (one line per instance, each with its own
// length and plan), laid end to end: instance i occupies [off[i], off[i]+n).
// Work items are waves: item w = (P, start, count) covers the instances
// order[start .. start+count) whose plans all have P chunks (count <= 64/P).
struct RagLines {
    const LineDev *plans;    // per instance
    const int64_t *off;      // per instance
    const int *order;        // instances grouped by chunk count
    const int *items;        // 3 ints per work item
    int nitems;
    int nshort;              // items [0, nshort) hold lines of chunks of at most RAG_SHORT_CHUNK points
    const double *dt;        // per-instance dt for the LSRK epilogue (NULL: epi.dt)
    const int *mask;         // per-instance: nonzero applies the epilogue (NULL: all)
    int q, M, sym;           // scheme shape shared by every instance
    int nf;                  // fields per launch: field f reads f + f*fstride and offsets
    int64_t fstride;         //   the epilogue's out / du / Q by f*fstride (elements)
};
void launch_line_rag(const RagLines &r, int dtype, const void *f, const EpiArgs &epi, hipStream_t st);

}  // namespace cles

struct cles_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int refs = 1;            // the owner's reference plus one per live plan / solver
};

namespace cles {
void ctx_retain(cles_ctx *ctx);
void ctx_release(cles_ctx *ctx);   // frees the context when the last reference goes
}  // namespace cles

struct cles_plan {
    cles_ctx *ctx = nullptr;
    cles::HostPlan hp;
};

#define CLES_TRY(...)                                                       \
    try {                                                                   \
        __VA_ARGS__;                                                        \
        return CLES_OK;                                                     \
    } catch (const cles::Error &e) {                                        \
        cles::set_error(e.msg);                                             \
        return e.code;                                                      \
    } catch (const std::bad_alloc &) {                                      \
        cles::set_error("host allocation failed");                          \
        return CLES_ERR_NOMEM;                                              \
    } catch (const std::exception &e) {                                     \
        cles::set_error(e.what());                                          \
        return CLES_ERR_INVALID;                                            \
    }

#define CLES_REQUIRE(cond, msg)                                             \
    do {                                                                    \
        if (!(cond)) throw cles::Error{CLES_ERR_INVALID, msg};              \
    } while (0)
