// Compact line-operator kernels for gfx950 (wave64).
//
// One thread owns one chunk (<= MC points) of one line and keeps it in
// registers: it forms the compact RHS from a register window of f
// (src/operators.jl:113-174 fill), runs the chunk-local Thomas / banded LU
// sweeps (src/tridiag.jl:46-57, src/banded.jl:63-80), exchanges its 2q end
// values with the other chunks of the same line through LDS, solves the
// reduced interface system by block parallel cyclic reduction (coefficients
// precomputed by the plan), and applies the spike correction
// x = y - V tail_prev - W head_next (src/tridiag.jl:159-168,
// src/banded.jl:371-382) before handing each value to a fused epilogue.
//
//  * y/z sweeps: 64 (or fewer) consecutive x-lanes per chunk row; every load
//    and store of a row is a contiguous 512-byte segment; no transposes.
//  * x sweep: the block stages Lt whole lines in LDS with coalesced loads,
//    each thread reads its chunk window from LDS, results return through the
//    same tile for coalesced stores.
// No MFMA: the work is banded and HBM-bound.
#include <hip/hip_runtime.h>

#include "cles_internal.h"

namespace cles {

constexpr int MC = 32;     // maximum chunk length (points per thread)
constexpr int NCMAX = 4;   // closure rows per edge handled by the kernels

template <typename T> __device__ __forceinline__ const T *coefp(const LineDev &p);
template <> __device__ __forceinline__ const double *coefp<double>(const LineDev &p) { return p.coef64; }
template <> __device__ __forceinline__ const float *coefp<float>(const LineDev &p) { return p.coef32; }
template <typename T> __device__ __forceinline__ const T *rowp(const LineDev &p);
template <> __device__ __forceinline__ const double *rowp<double>(const LineDev &p) { return p.row64; }
template <> __device__ __forceinline__ const float *rowp<float>(const LineDev &p) { return p.row32; }

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
template <typename T, int EPI>
__device__ __forceinline__ void epilogue(const EpiArgs &e, int64_t o, double v)
{
    if constexpr (EPI == EPI_STORE) {
        ((T *)e.out)[o] = (T)v;
    } else if constexpr (EPI == EPI_STORE_NEG) {
        ((T *)e.out)[o] = (T)(-v);
    } else if constexpr (EPI == EPI_SUB) {
        T *out = (T *)e.out;
        out[o] = out[o] - (T)v;
    } else if constexpr (EPI == EPI_LSRK) {
        // rhs.jl:665-667 dQ -= D_z F_z, then timestep.jl:85-94 on the FP64 state
        const T *acc = (const T *)e.a0;
        double dq = (double)(acc[o] - (T)v);
        double nv = e.c0 * e.du[o] + e.dt * dq;
        e.du[o] = nv;
        e.Q[o] += e.c1 * nv;
    } else if constexpr (EPI == EPI_ART_MUBETA) {
        // artificial.jl:296-304
        const T *rho = (const T *)e.a0;
        T s = (T)v;
        T rs = rho[o] * (s > T(0) ? s : T(0));
        ((T *)e.out)[o] = (T)e.c0 * rs;
        ((T *)e.out2)[o] = (T)e.c1 * rs;
    } else if constexpr (EPI == EPI_ART_KAPPA) {
        // artificial.jl:331-339 with physics.jl:155-156
        const T *rho = (const T *)e.a0, *c = (const T *)e.a1, *Tt = (const T *)e.a2;
        T s = (T)v;
        T tt = Tt[o];
        const T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07f;
        T scale = rho[o] * c[o] / (tt > eps ? tt : eps);
        ((T *)e.out)[o] = (T)e.c0 * scale * (s > T(0) ? s : T(0));
    } else if constexpr (EPI == EPI_ART_D) {
        // artificial.jl:349-355
        const T *c = (const T *)e.a1;
        T s = (T)v;
        ((T *)e.out)[o] = (T)e.c0 * c[o] * (s > T(0) ? s : T(0));
    }
}

// ---------------------------------------------------------------------------
// Per-thread chunk machinery
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM>
struct Chunk {
    static constexpr int B = 2 * Q;
    T d[MC];

    // RHS of every row of the chunk from the window fv[t + M] = f[s + t]
    // (closure rows from fv (low edge, s = 0) and fe[kap] = f[n-1-kap]).
    __device__ __forceinline__ void rhs(const LineDev &p, int k, int len, const T (&fv)[MC + 2 * M],
                                        const T (&fe)[8])
    {
        const T *cf = coefp<T>(p);
        T ci[M];
#pragma unroll
        for (int m = 0; m < M; m++) ci[m] = cf[COEF_CI + m];
        const T a0 = cf[COEF_A0];
        const bool lo = p.closed && k == 0;
        const bool hi = p.closed && k == p.P - 1;
#pragma unroll
        for (int t = 0; t < MC; t++) {
            T acc;
            if constexpr (SYM) {
                acc = a0 * fv[t + M];
#pragma unroll
                for (int m = 1; m <= M; m++) acc += ci[m - 1] * (fv[t + M + m] + fv[t + M - m]);
            } else {
                acc = T(0);
#pragma unroll
                for (int m = 1; m <= M; m++) acc += ci[m - 1] * (fv[t + M + m] - fv[t + M - m]);
            }
            if (t < NCMAX) {
                if (lo && t < p.nc) {
                    acc = T(0);
#pragma unroll
                    for (int kap = 0; kap < 8; kap++) acc += cf[COEF_CLO + t * 8 + kap] * fv[kap + M];
                }
            }
            if (hi) {
                const int jr = len - 1 - t;
                if (jr >= 0 && jr < p.nc) {
                    acc = T(0);
#pragma unroll
                    for (int kap = 0; kap < 8; kap++) acc += cf[COEF_CHI + jr * 8 + kap] * fe[kap];
                }
            }
            d[t] = acc;
        }
    }

    // chunk-local forward/backward sweeps
    __device__ __forceinline__ void solve(const LineDev &p, int s, int len)
    {
        const int n = p.n;
        const T *rw = rowp<T>(p);
        if constexpr (Q == 1) {
            const T *rl = rw + (size_t)R1_L * n + s;
            const T *rdi = rw + (size_t)R1_DINV * n + s;
            const T *rc = rw + (size_t)R1_C * n + s;
#pragma unroll
            for (int t = 1; t < MC; t++)
                if (t < len) d[t] -= rl[t] * d[t - 1];
#pragma unroll
            for (int t = MC - 1; t >= 0; t--) {
                if (t < len) {
                    const T nxt = d[t + 1 < MC ? t + 1 : MC - 1];
                    d[t] = (t == len - 1) ? d[t] * rdi[t] : (d[t] - rc[t] * nxt) * rdi[t];
                }
            }
        } else {
            const T *l1 = rw + (size_t)R2_L1 * n + s;
            const T *l2 = rw + (size_t)R2_L2 * n + s;
            const T *u0 = rw + (size_t)R2_U0INV * n + s;
            const T *u1 = rw + (size_t)R2_U1 * n + s;
            const T *u2 = rw + (size_t)R2_U2 * n + s;
#pragma unroll
            for (int t = 0; t < MC - 1; t++) {
                if (t < len - 1) {
                    const T xt = d[t];
                    d[t + 1] -= l1[t] * xt;
                    if (t + 2 < MC)
                        if (t < len - 2) d[t + 2 < MC ? t + 2 : t] -= l2[t] * xt;
                }
            }
#pragma unroll
            for (int t = MC - 1; t >= 0; t--) {
                if (t < len) {
                    T acc = d[t];
                    if (t + 1 < MC)
                        if (t + 1 < len) acc -= u1[t] * d[t + 1 < MC ? t + 1 : t];
                    if (t + 2 < MC)
                        if (t + 2 < len) acc -= u2[t] * d[t + 2 < MC ? t + 2 : t];
                    d[t] = acc * u0[t];
                }
            }
        }
    }

    __device__ __forceinline__ void ends(int len, double (&r)[B]) const
    {
#pragma unroll
        for (int i = 0; i < Q; i++) r[i] = (double)d[i];
#pragma unroll
        for (int t = 0; t < MC; t++)
#pragma unroll
            for (int i = 0; i < Q; i++)
                if (t == len - Q + i) r[Q + i] = (double)d[t];
    }
};

// Reduced interface system: block PCR across the P chunks of each line
// (slot = chunk * nl + line-in-block), then fetch the previous chunk's tail
// and the next chunk's head.  Every thread of the block must call this.
template <int Q>
__device__ __forceinline__ void reduced_solve(const LineDev &p, double *buf, int NT, int nl, int k, int li,
                                              double (&r)[2 * Q], double (&tp)[Q], double (&hn)[Q])
{
    constexpr int B = 2 * Q;
    const int P = p.P;
    const int slot = k * nl + li;
    const double *pc = p.pcr;
    for (int st = 0, dd = 1; st < p.nsteps; st++, dd <<= 1) {
        double *bf = buf + (st & 1) * B * NT;
#pragma unroll
        for (int c = 0; c < B; c++) bf[c * NT + slot] = r[c];
        __syncthreads();
        int km = k - dd, kp = k + dd;
        bool hm = true, hp = true;
        if (p.periodic) {
            km &= (P - 1);
            kp &= (P - 1);
        } else {
            hm = km >= 0;
            hp = kp < P;
        }
        double rm[B], rp[B];
#pragma unroll
        for (int c = 0; c < B; c++) {
            rm[c] = hm ? bf[c * NT + km * nl + li] : 0.0;
            rp[c] = hp ? bf[c * NT + kp * nl + li] : 0.0;
        }
        const double *al = pc + ((size_t)(st * P + k) * 2) * B * B;
        const double *be = al + B * B;
#pragma unroll
        for (int i = 0; i < B; i++) {
            double acc = r[i];
#pragma unroll
            for (int j = 0; j < B; j++) acc += al[i * B + j] * rm[j] + be[i * B + j] * rp[j];
            r[i] = acc;
        }
    }
    const double *mi = pc + (size_t)p.nsteps * P * 2 * B * B + (size_t)k * B * B;
    double z[B];
#pragma unroll
    for (int i = 0; i < B; i++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < B; j++) acc += mi[i * B + j] * r[j];
        z[i] = acc;
    }
    double *bf = buf + (p.nsteps & 1) * B * NT;
#pragma unroll
    for (int c = 0; c < B; c++) bf[c * NT + slot] = z[c];
    __syncthreads();
    int kprev = k - 1, knext = k + 1;
    bool hasp = true, hasn = true;
    if (p.periodic) {
        kprev &= (P - 1);
        knext &= (P - 1);
    } else {
        hasp = kprev >= 0;
        hasn = knext < P;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
        tp[i] = hasp ? bf[(Q + i) * NT + kprev * nl + li] : 0.0;
        hn[i] = hasn ? bf[i * NT + knext * nl + li] : 0.0;
    }
}

template <typename T, int Q>
__device__ __forceinline__ double corrected(const LineDev &p, int g, T y, const double (&tp)[Q],
                                            const double (&hn)[Q])
{
    const int n = p.n;
    const double *rw = p.row64;
    if constexpr (Q == 1) {
        return (double)y - (rw[(size_t)R1_V * n + g] * tp[0] + rw[(size_t)R1_W * n + g] * hn[0]);
    } else {
        double acc = rw[(size_t)R2_V0 * n + g] * tp[0] + rw[(size_t)R2_W0 * n + g] * hn[0];
        acc += rw[(size_t)R2_V1 * n + g] * tp[1] + rw[(size_t)R2_W1 * n + g] * hn[1];
        return (double)y - acc;
    }
}

// ---------------------------------------------------------------------------
// y / z sweeps: lanes along x, one chunk row per load instruction.
// Line dimension stride sd, outer dimension stride so; block = Lx lanes x P chunks.
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM, int EPI>
__global__ void __launch_bounds__(512) k_line_yz(LineDev p, const T *__restrict__ f, int64_t sd, int64_t so,
                                                 int nlane, int nxb, int Lx, EpiArgs e)
{
    extern __shared__ double smem[];
    const int NT = Lx * p.P;
    const int lane = threadIdx.x % Lx;
    const int k = threadIdx.x / Lx;
    const int xb = blockIdx.x % nxb;
    const int64_t ob = blockIdx.x / nxb;
    const int xi = xb * Lx + lane;
    const bool valid = xi < nlane;
    const int64_t off = (int64_t)xi + ob * so;
    const int n = p.n;
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);

    T fv[MC + 2 * M];
#pragma unroll
    for (int t = -M; t < MC + M; t++) {
        T val = T(0);
        if (valid && t < len + M) {
            int r = s + t;
            bool ok = true;
            if (r < 0) {
                if (p.periodic) r += n; else ok = false;
            } else if (r >= n) {
                if (p.periodic) r -= n; else ok = false;
            }
            if (ok) val = f[off + (int64_t)r * sd];
        }
        fv[t + M] = val;
    }
    T fe[8];
#pragma unroll
    for (int kap = 0; kap < 8; kap++) fe[kap] = T(0);
    if (p.closed && k == p.P - 1 && valid) {
#pragma unroll
        for (int kap = 0; kap < 8; kap++)
            if (kap < n) fe[kap] = f[off + (int64_t)(n - 1 - kap) * sd];
    }

    Chunk<T, Q, M, SYM> ch;
    ch.rhs(p, k, len, fv, fe);
    ch.solve(p, s, len);
    double r[2 * Q], tp[Q], hn[Q];
    ch.ends(len, r);
    reduced_solve<Q>(p, smem, NT, Lx, k, lane, r, tp, hn);
    if (!valid) return;
#pragma unroll
    for (int t = 0; t < MC; t++) {
        if (t < len) {
            const int g = s + t;
            epilogue<T, EPI>(e, off + (int64_t)g * sd, corrected<T, Q>(p, g, ch.d[t], tp, hn));
        }
    }
}

// ---------------------------------------------------------------------------
// x sweep: Lt whole lines staged in LDS (row stride n + 1), threads own
// (line, chunk) pairs.  Lines are indexed l = j + ny*k over the y-z plane.
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM, int EPI>
__global__ void __launch_bounds__(512) k_line_x(LineDev p, const T *__restrict__ f, int64_t sy, int64_t sz,
                                                int ny, int64_t nlines, int Lt, EpiArgs e)
{
    extern __shared__ double smem[];
    const int n = p.n;
    const int TS = n + 1;
    const int NT = Lt * p.P;
    T *tile = (T *)smem;
    double *buf = smem + (((size_t)Lt * TS * sizeof(T) + 7) / 8);
    const int tid = threadIdx.x;
    const int64_t l0 = (int64_t)blockIdx.x * Lt;

    for (int li = 0; li < Lt; li++) {
        const int64_t l = l0 + li;
        if (l >= nlines) break;
        const int64_t off = (l % ny) * sy + (l / ny) * sz;
        for (int i = tid; i < n; i += NT) tile[li * TS + i] = f[off + i];
    }
    __syncthreads();

    const int li = tid % Lt;
    const int k = tid / Lt;
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);
    const T *trow = tile + li * TS;
    T fv[MC + 2 * M];
#pragma unroll
    for (int t = -M; t < MC + M; t++) {
        T val = T(0);
        if (t < len + M) {
            int r = s + t;
            bool ok = true;
            if (r < 0) {
                if (p.periodic) r += n; else ok = false;
            } else if (r >= n) {
                if (p.periodic) r -= n; else ok = false;
            }
            if (ok) val = trow[r];
        }
        fv[t + M] = val;
    }
    T fe[8];
#pragma unroll
    for (int kap = 0; kap < 8; kap++) fe[kap] = (p.closed && kap < n) ? trow[n - 1 - kap] : T(0);

    Chunk<T, Q, M, SYM> ch;
    ch.rhs(p, k, len, fv, fe);
    ch.solve(p, s, len);
    double r[2 * Q], tp[Q], hn[Q];
    ch.ends(len, r);
    reduced_solve<Q>(p, buf, NT, Lt, k, li, r, tp, hn);
    // every thread has finished reading the tile before the first reduced-solve barrier
    T *wrow = tile + li * TS;
#pragma unroll
    for (int t = 0; t < MC; t++)
        if (t < len) wrow[s + t] = (T)corrected<T, Q>(p, s + t, ch.d[t], tp, hn);
    __syncthreads();

    for (int lj = 0; lj < Lt; lj++) {
        const int64_t l = l0 + lj;
        if (l >= nlines) break;
        const int64_t off = (l % ny) * sy + (l / ny) * sz;
        for (int i = tid; i < n; i += NT) epilogue<T, EPI>(e, off + i, (double)tile[lj * TS + i]);
    }
}

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM, int EPI>
static void launch_variant(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e,
                           hipStream_t st)
{
    const LineDev &p = hp.dev;
    const int P = p.P;
    const int B = 2 * Q;
    if (dim == 1) {
        int Lt = 1;
        const size_t tile_row = (size_t)(p.n + 1) * sizeof(T);
        while (Lt * 2 * P <= 512 && (size_t)Lt * 2 * tile_row <= 48 * 1024 && Lt < 32) Lt *= 2;
        if (Lt * P > 512) throw Error{CLES_ERR_UNSUPPORTED, "line too long for the x-sweep kernel"};
        const int64_t nlines = v.n[1] * v.n[2];
        const int NT = Lt * P;
        size_t shm = (((size_t)Lt * tile_row + 7) / 8) * 8 + (size_t)2 * B * NT * sizeof(double);
        const int64_t nblk = (nlines + Lt - 1) / Lt;
        hipLaunchKernelGGL((k_line_x<T, Q, M, SYM, EPI>), dim3((unsigned)nblk), dim3(NT), shm, st, p, f, v.s[1],
                           v.s[2], (int)v.n[1], nlines, Lt, e);
    } else {
        int Lx = 64;
        while (Lx * P > 512 && Lx > 1) Lx /= 2;
        if (Lx * P > 512) throw Error{CLES_ERR_UNSUPPORTED, "line too long for the y/z-sweep kernel"};
        const int64_t nlane = v.n[0];
        while (Lx / 2 >= nlane && Lx > 1) Lx /= 2;   // narrow x extents (2-D / 1-D grids)
        const int NT = Lx * P;
        const int nxb = (int)((nlane + Lx - 1) / Lx);
        const int64_t sd = dim == 2 ? v.s[1] : v.s[2];
        const int64_t so = dim == 2 ? v.s[2] : v.s[1];
        const int64_t nouter = dim == 2 ? v.n[2] : v.n[1];
        size_t shm = (size_t)2 * B * NT * sizeof(double);
        hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI>), dim3((unsigned)(nxb * nouter)), dim3(NT), shm, st, p,
                           f, sd, so, (int)nlane, nxb, Lx, e);
    }
    CLES_HIP(hipGetLastError());
}

template <typename T, int Q, int M, int SYM>
static void launch_epi(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e, hipStream_t st)
{
    if constexpr (SYM) {
        switch (e.kind) {
        case EPI_STORE: return launch_variant<T, Q, M, SYM, EPI_STORE>(hp, dim, v, f, e, st);
        case EPI_ART_MUBETA: return launch_variant<T, Q, M, SYM, EPI_ART_MUBETA>(hp, dim, v, f, e, st);
        case EPI_ART_KAPPA: return launch_variant<T, Q, M, SYM, EPI_ART_KAPPA>(hp, dim, v, f, e, st);
        case EPI_ART_D: return launch_variant<T, Q, M, SYM, EPI_ART_D>(hp, dim, v, f, e, st);
        default: break;
        }
    } else {
        switch (e.kind) {
        case EPI_STORE: return launch_variant<T, Q, M, SYM, EPI_STORE>(hp, dim, v, f, e, st);
        case EPI_STORE_NEG: return launch_variant<T, Q, M, SYM, EPI_STORE_NEG>(hp, dim, v, f, e, st);
        case EPI_SUB: return launch_variant<T, Q, M, SYM, EPI_SUB>(hp, dim, v, f, e, st);
        case EPI_LSRK: return launch_variant<T, Q, M, SYM, EPI_LSRK>(hp, dim, v, f, e, st);
        default: break;
        }
    }
    throw Error{CLES_ERR_UNSUPPORTED, "epilogue not available for this scheme kind"};
}

template <typename T>
static void launch_typed(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e, hipStream_t st)
{
    const LineDev &p = hp.dev;
    const int q = p.q, M = p.M, sym = p.sym;
    if (q == 1 && M == 2 && !sym) return launch_epi<T, 1, 2, 0>(hp, dim, v, f, e, st);   // Lele C6
    if (q == 1 && M == 4 && sym) return launch_epi<T, 1, 4, 1>(hp, dim, v, f, e, st);    // G-V C8 filter
    if (q == 2 && M == 3 && !sym) return launch_epi<T, 2, 3, 0>(hp, dim, v, f, e, st);   // Lele C10
    if (q == 1 && M == 1 && !sym) return launch_epi<T, 1, 1, 0>(hp, dim, v, f, e, st);   // Pade C4
    // other schemes: widest instantiation of their kind (unused taps carry zero weights)
    if (q == 1 && !sym) return launch_epi<T, 1, 4, 0>(hp, dim, v, f, e, st);
    if (q == 2 && !sym) return launch_epi<T, 2, 4, 0>(hp, dim, v, f, e, st);
    if (q == 2 && sym) return launch_epi<T, 2, 4, 1>(hp, dim, v, f, e, st);
    return launch_epi<T, 1, 4, 1>(hp, dim, v, f, e, st);
}

void launch_line(const HostPlan &hp, int dim, const View3 &v, const void *f, const EpiArgs &e, hipStream_t st)
{
    if (hp.dtype == CLES_F64)
        launch_typed<double>(hp, dim, v, (const double *)f, e, st);
    else
        launch_typed<float>(hp, dim, v, (const float *)f, e, st);
}

}  // namespace cles
