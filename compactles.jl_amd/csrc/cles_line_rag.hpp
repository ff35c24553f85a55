// Ragged batch of independent compact lines (one line per instance of a
// batch of one-dimensional problems, each with its own length and plan).
//
// One wave per work item.  The item's 64/P lines (all with P chunks of at
// most RAG_MAX_CHUNK points) are staged in LDS with coalesced loads; lane
// (line, chunk) forms its chunk's RHS from the tile, runs the chunk sweeps,
// solves the reduced interface system by block PCR over cross-lane shuffles
// and applies the spike correction (the same machinery as the x sweep); the
// results return through the tile for coalesced, per-instance epilogues.
// Every line is computed from its own data and plan only, so an instance's
// result does not depend on the other instances of the batch.
#pragma once
#include "cles_line.hpp"

namespace cles {

// a pad slot after every CL points (the chunk capacity) spreads the
// chunk-strided reads over the banks
template <int CL>
__device__ __forceinline__ int rag_slot(int x) { return x + x / CL; }

// Blocks are dispatched round-robin over the 8 XCDs; renumbering them so that
// each XCD takes one contiguous range keeps the nf blocks of one work item
// (same lines and tables, different fields) on one XCD's L2.
__device__ __forceinline__ unsigned rag_block(unsigned b, unsigned nb)
{
    const unsigned nfull = nb & ~7u;
    return b < nfull ? (b & 7u) * (nfull >> 3) + (b >> 3) : b;
}

template <typename T> __device__ __forceinline__ const T *rowtp(const LineDev &p);
template <> __device__ __forceinline__ const double *rowtp<double>(const LineDev &p) { return p.rowt64; }
template <> __device__ __forceinline__ const float *rowtp<float>(const LineDev &p) { return p.rowt32; }

// The chunk-major plan rows of chunk k (RAG_MAX_CHUNK zero-padded rows per
// chunk, so every t < RAG_MAX_CHUNK is in bounds): the P lanes of a line read
// consecutive words for each (field, t).
template <typename T>
struct RowsRag {
    const T *base;
    int P;
    __device__ __forceinline__ T operator()(int f, int t) const
    {
        return gload(base + ((size_t)f * RAG_MAX_CHUNK + t) * P);
    }
    template <typename D> __device__ __forceinline__ void fence(const D &dep)
    {
        asm volatile("" : "+v"(base) : "v"(dep));
    }
};

// CL: chunk capacity of the launch's items (RAG_SHORT_CHUNK for the short
// class, RAG_MAX_CHUNK otherwise); the plan rows keep RAG_MAX_CHUNK per chunk.
template <typename T, int Q, int M, int SYM, int EPI, int CL>
__global__ void __launch_bounds__(64) k_line_rag(RagLines r, const T *__restrict__ f, EpiArgs e)
{
    extern __shared__ double smem[];
    T *tile = (T *)smem;
    const int lane = threadIdx.x;
    const unsigned logical = rag_block(blockIdx.x, gridDim.x);
    const int item = (int)(logical / (unsigned)r.nf), fld = (int)(logical % (unsigned)r.nf);
    const int64_t fo = (int64_t)fld * r.fstride;
    f += fo;
    if constexpr (EPI == EPI_ART_MULTI) e = e.multi[fld];   // its own outputs: no field offset
    const int *it = r.items + 3 * item;
    const int P = it[0], start = it[1], count = it[2];
    const int TS = (CL + 1) * P;   // rag_slot<CL>(CL P - 1) < (CL + 1) P

    // stage the item's lines
    for (int lj = 0; lj < count; lj++) {
        const int inst = r.order[start + lj];
        const int n = r.plans[inst].n;
        const T *src = f + r.off[inst];
        T *row = tile + lj * TS;
        // a line of up to 64 CL points in one round of CL loads per lane (the
        // registers of the sweeps are not live yet)
        for (int x0 = 0; x0 < n; x0 += 64 * CL) {
            T v[CL];
#pragma unroll
            for (int i = 0; i < CL; i++) {
                const int x = x0 + lane + 64 * i;
                v[i] = x < n ? src[x] : T(0);
            }
#pragma unroll
            for (int i = 0; i < CL; i++) {
                const int x = x0 + lane + 64 * i;
                if (x < n) row[rag_slot<CL>(x)] = v[i];
            }
        }
    }
    __syncthreads();

    // lanes of empty line slots repeat line 0 and store nothing
    const int li = lane / P, k = lane & (P - 1);
    const bool live = li < count;
    const int inst = r.order[start + (live ? li : 0)];
    const LineDev p = r.plans[inst];
    const int n = p.n;
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);
    T *trow = tile + (live ? li : 0) * TS;
    T fv[CL + 2 * M];
#pragma unroll
    for (int t = -M; t < CL + M; t++) {
        T val = T(0);
        if (t < len + M) {
            int rr = s + t;
            bool ok = true;
            if (rr < 0) {
                if (p.periodic) rr += n; else ok = false;
            } else if (rr >= n) {
                if (p.periodic) rr -= n; else ok = false;
            }
            if (ok) val = trow[rag_slot<CL>(rr)];
        }
        fv[t + M] = val;
    }
    T fe[8];
#pragma unroll
    for (int kap = 0; kap < 8; kap++) fe[kap] = (p.closed && kap < n) ? trow[rag_slot<CL>(n - 1 - kap)] : T(0);
    Chunk<T, Q, M, SYM, CL> ch;
    ch.rhs(p, k, len, fv, fe);
    solve_grouped<T, Q, CL, 8>(ch.d, len, RowsRag<T>{rowtp<T>(p) + k, p.P});
    double ends[2 * Q], tp[Q], hn[Q];
    ch.ends(len, ends);
    reduced_solve_wave<Q, Q == 1>(p, k, k, li * P, ends, tp, hn);
    correct_grouped<T, Q, CL, 8>(ch.d, len, RowsRag<double>{p.rowt64 + k, p.P}, tp, hn);
    __syncthreads();   // every window read is done before the tile is overwritten
    if (live) {
#pragma unroll
        for (int t = 0; t < CL; t++)
            if (t < len) trow[rag_slot<CL>(s + t)] = ch.d[t];
    }
    __syncthreads();

    for (int lj = 0; lj < count; lj++) {
        const int ij = r.order[start + lj];
        if (r.mask && !r.mask[ij]) continue;
        const int nj = r.plans[ij].n;
        const int64_t oj = r.off[ij];
        EpiArgs ej = e;
        if (EPI != EPI_ART_MULTI && fo) {
            if (ej.out) ej.out = (T *)ej.out + fo;
            if (ej.du) ej.du += fo;
            if (ej.Q) ej.Q += fo;
        }
        if (r.dt) ej.dt = r.dt[ij];
        const T *row = tile + lj * TS;
        // at capacity 32, 16 points per lane per batch (the sweeps' registers
        // are free here; at capacity 16 that would raise the allocation)
        constexpr int NB = CL > 16 ? 16 : 8;
        for (int x0 = 0; x0 < nj; x0 += 64 * NB) {
            int64_t o[NB];
            T v[NB];
            bool ok[NB];
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const int x = x0 + lane + 64 * i;
                ok[i] = x < nj;
                o[i] = oj + (ok[i] ? x : 0);
                v[i] = ok[i] ? row[rag_slot<CL>(x)] : T(0);
            }
            epilogue_batch<T, EPI, NB>(ej, o, v, ok);
        }
    }
}

template <typename T, int Q, int M, int SYM, int EPI>
void launch_rag_variant(const RagLines &r, const T *f, const EpiArgs &e, hipStream_t st)
{
    if (r.nshort < 0 || r.nshort > r.nitems) throw Error{CLES_ERR_INVALID, "short items out of range"};
    if (r.nitems <= 0 || r.nf <= 0) return;
    if (r.nf > 1 && e.kind != EPI_STORE && e.kind != EPI_STORE_NEG && e.kind != EPI_LSRK && e.kind != EPI_ART_MULTI)
        throw Error{CLES_ERR_INVALID, "several fields per launch need a store, LSRK or multi epilogue"};
    if (e.kind == EPI_ART_MULTI && !e.multi) throw Error{CLES_ERR_INVALID, "the multi epilogue needs its table"};
    if (e.kind == EPI_LSRK && e.a0 && r.nf > 1)
        throw Error{CLES_ERR_INVALID, "the ragged LSRK epilogue takes no accumulator with several fields"};
    // the short items, then the rest, each with its own chunk capacity
    RagLines rs = r, rl = r;
    rs.nitems = r.nshort;
    rl.items = r.items + 3 * r.nshort;
    rl.nitems = r.nitems - r.nshort;
    if (rs.nitems > 0) {
        const size_t shm = (size_t)64 * (RAG_SHORT_CHUNK + 1) * sizeof(T);
        const int64_t nb = (int64_t)rs.nitems * r.nf;
        hipLaunchKernelGGL((k_line_rag<T, Q, M, SYM, EPI, RAG_SHORT_CHUNK>), dim3((unsigned)nb), dim3(64), shm, st, rs, f,
                           e);
        CLES_HIP(hipGetLastError());
    }
    if (rl.nitems > 0) {
        const size_t shm = (size_t)64 * (RAG_MAX_CHUNK + 1) * sizeof(T);
        const int64_t nb = (int64_t)rl.nitems * r.nf;
        hipLaunchKernelGGL((k_line_rag<T, Q, M, SYM, EPI, RAG_MAX_CHUNK>), dim3((unsigned)nb), dim3(64), shm, st, rl, f,
                           e);
        CLES_HIP(hipGetLastError());
    }
}

template <typename T, int Q, int M, int SYM>
void launch_rag_epi(const RagLines &r, const T *f, const EpiArgs &e, hipStream_t st)
{
    if constexpr (SYM) {
        switch (e.kind) {
        case EPI_STORE: return launch_rag_variant<T, Q, M, SYM, EPI_STORE>(r, f, e, st);
        case EPI_ART_MUBETA: return launch_rag_variant<T, Q, M, SYM, EPI_ART_MUBETA>(r, f, e, st);
        case EPI_ART_KAPPA: return launch_rag_variant<T, Q, M, SYM, EPI_ART_KAPPA>(r, f, e, st);
        case EPI_ART_D: return launch_rag_variant<T, Q, M, SYM, EPI_ART_D>(r, f, e, st);
        case EPI_ART_MULTI: return launch_rag_variant<T, Q, M, SYM, EPI_ART_MULTI>(r, f, e, st);
        default: break;
        }
    } else {
        switch (e.kind) {
        case EPI_STORE: return launch_rag_variant<T, Q, M, SYM, EPI_STORE>(r, f, e, st);
        case EPI_STORE_NEG: return launch_rag_variant<T, Q, M, SYM, EPI_STORE_NEG>(r, f, e, st);
        case EPI_LSRK: return launch_rag_variant<T, Q, M, SYM, EPI_LSRK>(r, f, e, st);
        default: break;
        }
    }
    throw Error{CLES_ERR_UNSUPPORTED, "epilogue not available for the ragged line kernel"};
}

}  // namespace cles
