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
#pragma once
#include <hip/hip_runtime.h>

#include "cles_internal.h"

// Multiply-adds in the line kernels are written as explicit fma where they
// matter and are otherwise not contracted: the compiler's choice of which
// product to fuse can differ between instantiations of the same arithmetic
// (chunk capacities, epilogues), and every kernel must round alike.
#pragma clang fp contract(off)

namespace cles {

constexpr int MC = 32;     // maximum chunk length (points per thread)
constexpr int NCMAX = 4;   // closure rows per edge handled by the kernels

// Plan tables live in global memory; a pointer read out of memory (the ragged
// kernel's per-instance plans) is otherwise a flat pointer, whose loads also
// wait on the LDS counter.
template <typename T> __device__ __forceinline__ T gload(const T *ptr)
{
    return *(const __attribute__((address_space(1))) T *)ptr;
}

template <typename T> __device__ __forceinline__ const T *coefp(const LineDev &p);
template <> __device__ __forceinline__ const double *coefp<double>(const LineDev &p) { return p.coef64; }
template <> __device__ __forceinline__ const float *coefp<float>(const LineDev &p) { return p.coef32; }
template <typename T> __device__ __forceinline__ const T *rowp(const LineDev &p);
template <> __device__ __forceinline__ const double *rowp<double>(const LineDev &p) { return p.row64; }
template <> __device__ __forceinline__ const float *rowp<float>(const LineDev &p) { return p.row32; }

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
// NT: plain stores carry the non-temporal hint (a compile-time choice: a
// run-time branch between two otherwise identical stores is merged into one
// plain store by the compiler)
template <typename T, int EPI, bool NT = false>
__device__ __forceinline__ void epilogue(const EpiArgs &e, int64_t o, double v)
{
    if constexpr (EPI == EPI_ART_MULTI) {   // e is the field's own ART_* epilogue (select_field)
        if (e.kind == EPI_ART_MUBETA) epilogue<T, EPI_ART_MUBETA, NT>(e, o, v);
        else if (e.kind == EPI_ART_KAPPA) epilogue<T, EPI_ART_KAPPA, NT>(e, o, v);
        else epilogue<T, EPI_ART_D, NT>(e, o, v);
        return;
    }
    // gated kinds: one flag and dt, or (ipts > 0: instances of ipts points side
    // by side) the flag and dt of the point's instance
    [[maybe_unused]] unsigned ii = 0;
    if constexpr (EPI == EPI_STORE_GATE || EPI == EPI_LSRK_DEV) {
        if (e.ipts) ii = (unsigned)o / e.ipts;
        if (!e.gate[ii]) return;
    }
    if constexpr (EPI == EPI_STORE || EPI == EPI_STORE_GATE || EPI == EPI_STORE_NEG) {
        const T val = EPI == EPI_STORE_NEG ? (T)(-v) : (T)v;
        if constexpr (NT)
            __builtin_nontemporal_store(val, (T *)e.out + o);
        else
            ((T *)e.out)[o] = val;
    } else if constexpr (EPI == EPI_SUB) {
        T *out = (T *)e.out;
        out[o] = out[o] - (T)v;
    } else if constexpr (EPI == EPI_LSRK || EPI == EPI_LSRK_DEV) {
        // rhs.jl:665-667 dQ -= D_z F_z, then timestep.jl:85-94 on the FP64 state
        // acc == NULL: this is the only active dimension, dQ = 0 - v
        const T *acc = (const T *)e.a0;
        const double dq = acc ? (double)(acc[o] - (T)v) : -v;
        const double dt = EPI == EPI_LSRK_DEV ? e.dtp[ii] : e.dt;
        double nv = fma(dt, dq, e.c0 * e.du[o]);
        e.du[o] = nv;
        e.Q[o] = fma(e.c1, nv, e.Q[o]);
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

// Batched epilogue: every input of NB points is loaded before any output is
// stored, so the loads of a batch are in flight together (the compiler cannot
// reorder them across stores through pointers that may alias).
template <typename T, int EPI, int NB, bool NT = false>
__device__ __forceinline__ void epilogue_batch(const EpiArgs &e, const int64_t (&o)[NB], const T (&v)[NB],
                                               const bool (&ok)[NB])
{
    if constexpr (EPI == EPI_ART_MULTI) {
        if (e.kind == EPI_ART_MUBETA) epilogue_batch<T, EPI_ART_MUBETA, NB, NT>(e, o, v, ok);
        else if (e.kind == EPI_ART_KAPPA) epilogue_batch<T, EPI_ART_KAPPA, NB, NT>(e, o, v, ok);
        else epilogue_batch<T, EPI_ART_D, NB, NT>(e, o, v, ok);
        return;
    }
    // gated kinds: one flag and dt for the launch, or (ipts > 0) those of
    // each point's instance
    bool okk[NB];
    double dtv[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        okk[i] = ok[i];
        dtv[i] = e.dt;
    }
    if constexpr (EPI == EPI_STORE_GATE || EPI == EPI_LSRK_DEV) {
        if (e.ipts) {
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const unsigned ii = ok[i] ? (unsigned)o[i] / e.ipts : 0u;
                okk[i] = ok[i] && e.gate[ii] != 0;
                if constexpr (EPI == EPI_LSRK_DEV) dtv[i] = e.dtp[ii];
            }
        } else {
            if (!*e.gate) return;
            if constexpr (EPI == EPI_LSRK_DEV) {
                const double d = *e.dtp;
#pragma unroll
                for (int i = 0; i < NB; i++) dtv[i] = d;
            }
        }
    }
    if constexpr (EPI == EPI_STORE || EPI == EPI_STORE_NEG || EPI == EPI_STORE_GATE) {
        constexpr int E1 = EPI == EPI_STORE_GATE ? (int)EPI_STORE : EPI;   // the gate is applied
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (okk[i]) epilogue<T, E1, NT>(e, o[i], (double)v[i]);
    } else if constexpr (EPI == EPI_SUB) {
        T *out = (T *)e.out;
        T a[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) a[i] = ok[i] ? out[o[i]] : T(0);
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (ok[i]) out[o[i]] = a[i] - v[i];
    } else if constexpr (EPI == EPI_LSRK || EPI == EPI_LSRK_DEV) {
        const T *acc = (const T *)e.a0;
        T a[NB];
        double du[NB], q[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            a[i] = (okk[i] && acc) ? acc[o[i]] : T(0);
            du[i] = okk[i] ? e.du[o[i]] : 0.0;
            q[i] = okk[i] ? e.Q[o[i]] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < NB; i++) {
            if (!okk[i]) continue;
            const double dq = acc ? (double)(a[i] - v[i]) : -(double)v[i];
            const double nv = fma(dtv[i], dq, e.c0 * du[i]);
            e.du[o[i]] = nv;
            e.Q[o[i]] = fma(e.c1, nv, q[i]);
        }
    } else if constexpr (EPI == EPI_ART_MUBETA) {
        const T *rho = (const T *)e.a0;
        T r[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) r[i] = ok[i] ? rho[o[i]] : T(0);
#pragma unroll
        for (int i = 0; i < NB; i++) {
            if (!ok[i]) continue;
            const T rs = r[i] * (v[i] > T(0) ? v[i] : T(0));
            ((T *)e.out)[o[i]] = (T)e.c0 * rs;
            ((T *)e.out2)[o[i]] = (T)e.c1 * rs;
        }
    } else if constexpr (EPI == EPI_ART_KAPPA) {
        const T *rho = (const T *)e.a0, *c = (const T *)e.a1, *Tt = (const T *)e.a2;
        T r[NB], cc[NB], tt[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            r[i] = ok[i] ? rho[o[i]] : T(0);
            cc[i] = ok[i] ? c[o[i]] : T(0);
            tt[i] = ok[i] ? Tt[o[i]] : T(1);
        }
        const T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07f;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            if (!ok[i]) continue;
            const T scale = r[i] * cc[i] / (tt[i] > eps ? tt[i] : eps);
            ((T *)e.out)[o[i]] = (T)e.c0 * scale * (v[i] > T(0) ? v[i] : T(0));
        }
    } else if constexpr (EPI == EPI_ART_D) {
        const T *c = (const T *)e.a1;
        T cc[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) cc[i] = ok[i] ? c[o[i]] : T(0);
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (ok[i]) ((T *)e.out)[o[i]] = (T)e.c0 * cc[i] * (v[i] > T(0) ? v[i] : T(0));
    }
}

// Field of a multi-field launch (grid y): field 0 keeps the launch's own input
// and epilogue, field i >= 1 takes its input and output from the table.
template <typename T, int EPI>
__device__ __forceinline__ void select_field(const LineFields &mf, const T *__restrict__ &f, EpiArgs &e)
{
    const int i = blockIdx.y;
    if constexpr (EPI == EPI_ART_MULTI) e = e.multi[i];
    if (i == 0) return;
    f = (const T *)mf.in[i];
    if constexpr (EPI == EPI_LSRK || EPI == EPI_LSRK_DEV) {
        const int64_t o = mf.eoff[i];
        if (e.a0) e.a0 = (const T *)e.a0 + o;
        e.du += o;
        e.Q += o;
    } else if constexpr (EPI != EPI_ART_MULTI) {
        e.out = mf.out[i];
    }
}

// ---------------------------------------------------------------------------
// Per-thread chunk machinery
// ---------------------------------------------------------------------------
// Interior RHS of one row from the window fv (centre index c):
// a0 f_c + sum_m c_m (f_{c+m} + f_{c-m}) (symmetric) or
// sum_m c_m (f_{c+m} - f_{c-m}), summed left to right, one fma per term.
template <typename T, int M, int SYM, int W>
__device__ __forceinline__ T rhs_row(const T (&fv)[W], int c, const T (&ci)[M], T a0)
{
    T acc;
    if constexpr (SYM) {
        acc = a0 * fv[c];
#pragma unroll
        for (int m = 1; m <= M; m++) acc = fma(ci[m - 1], fv[c + m] + fv[c - m], acc);
    } else {
        acc = ci[0] * (fv[c + 1] - fv[c - 1]);
#pragma unroll
        for (int m = 2; m <= M; m++) acc = fma(ci[m - 1], fv[c + m] - fv[c - m], acc);
    }
    return acc;
}

template <typename T, int Q, int M, int SYM, int CL = MC>
struct Chunk {
    static constexpr int B = 2 * Q;
    T d[CL];

    // RHS of every row of the chunk from the window fv[t + M] = f[s + t]
    // (closure rows from fv (low edge, s = 0) and fe[kap] = f[n-1-kap]).
    __device__ __forceinline__ void rhs(const LineDev &p, int k, int len, const T (&fv)[CL + 2 * M],
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
        for (int t = 0; t < CL; t++) {
            T acc = rhs_row<T, M, SYM>(fv, t + M, ci, a0);
            if (t < NCMAX) {
                if (lo && t < p.nc) {
                    const T *cr = cf + COEF_CLO + t * 8;
                    acc = cr[0] * fv[M];
#pragma unroll
                    for (int kap = 1; kap < 8; kap++) acc = fma(cr[kap], fv[kap + M], acc);
                }
            }
            if (hi) {
                const int jr = len - 1 - t;
                if (jr >= 0 && jr < p.nc) {
                    const T *cr = cf + COEF_CHI + jr * 8;
                    acc = cr[0] * fe[0];
#pragma unroll
                    for (int kap = 1; kap < 8; kap++) acc = fma(cr[kap], fe[kap], acc);
                }
            }
            d[t] = acc;
        }
    }

    // chunk-local forward/backward sweeps
    __device__ __forceinline__ void solve(const LineDev &p, int rs, int len)
    {
        const int n = p.n;
        const T *rw = rowp<T>(p);
        if constexpr (Q == 1) {
            const T *rl = rw + (size_t)R1_L * n + rs;
            const T *rdi = rw + (size_t)R1_DINV * n + rs;
            const T *rc = rw + (size_t)R1_C * n + rs;
#pragma unroll
            for (int t = 1; t < CL; t++)
                if (t < len) d[t] = fma(-rl[t], d[t - 1], d[t]);
#pragma unroll
            for (int t = CL - 1; t >= 0; t--) {
                if (t < len) {
                    const T nxt = d[t + 1 < CL ? t + 1 : CL - 1];
                    d[t] = (t == len - 1) ? d[t] * rdi[t] : fma(-rc[t], nxt, d[t]) * rdi[t];
                }
            }
        } else {
            const T *l1 = rw + (size_t)R2_L1 * n + rs;
            const T *l2 = rw + (size_t)R2_L2 * n + rs;
            const T *u0 = rw + (size_t)R2_U0INV * n + rs;
            const T *u1 = rw + (size_t)R2_U1 * n + rs;
            const T *u2 = rw + (size_t)R2_U2 * n + rs;
#pragma unroll
            for (int t = 0; t < CL - 1; t++) {
                if (t < len - 1) {
                    const T xt = d[t];
                    d[t + 1] = fma(-l1[t], xt, d[t + 1]);
                    if (t + 2 < CL)
                        if (t < len - 2) d[t + 2 < CL ? t + 2 : t] = fma(-l2[t], xt, d[t + 2 < CL ? t + 2 : t]);
                }
            }
#pragma unroll
            for (int t = CL - 1; t >= 0; t--) {
                if (t < len) {
                    T acc = d[t];
                    if (t + 1 < CL)
                        if (t + 1 < len) acc = fma(-u1[t], d[t + 1 < CL ? t + 1 : t], acc);
                    if (t + 2 < CL)
                        if (t + 2 < len) acc = fma(-u2[t], d[t + 2 < CL ? t + 2 : t], acc);
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
        for (int t = 0; t < CL; t++)
#pragma unroll
            for (int i = 0; i < Q; i++)
                if (t == len - Q + i) r[Q + i] = (double)d[t];
    }
};

// Chunk sweeps and spike correction with the row constants read G rows at a
// time: the loads of group g + 1 are issued once the arithmetic of group g - 1
// is done and before that of group g, so each group waits for one round trip
// (instead of every row waiting for its own inside the dependency chain) while
// only two groups of rows are held in registers.  Rows::fence(v) is an empty
// asm statement that makes the table pointer depend on v, which keeps the
// compiler from issuing a group's loads earlier than that.  Rows(field, t)
// must be in bounds for every t < CL (rows past the chunk's length read
// padding or a clamped row); rows t >= len are left unchanged through
// selects, so the arithmetic of rows t < len is that of Chunk::solve /
// corrected(), operation for operation.
template <typename T, int Q, int CL, int G, typename Rows>
__device__ __forceinline__ void solve_grouped(T (&d)[CL], int len, Rows R)
{
    static_assert(CL % G == 0, "group size must divide the chunk length");
    constexpr int NG = CL / G;
    if constexpr (Q == 1) {
        T cur[G], nxt[G];
#pragma unroll
        for (int i = 0; i < G; i++) cur[i] = R(R1_L, i);
#pragma unroll
        for (int g = 0; g < NG; g++) {
            if (g + 1 < NG) {
                if (g > 0) R.fence(d[g * G - 1]);
#pragma unroll
                for (int i = 0; i < G; i++) nxt[i] = R(R1_L, (g + 1) * G + i);
            }
#pragma unroll
            for (int i = 0; i < G; i++) {
                const int t = g * G + i;
                if (t >= 1) {
                    const T nv = fma(-cur[i], d[t - 1], d[t]);
                    d[t] = t < len ? nv : d[t];
                }
            }
            if (g + 1 < NG) {
#pragma unroll
                for (int i = 0; i < G; i++) cur[i] = nxt[i];
            }
        }
        T cc[G], cd[G], nc[G], nd[G];
        R.fence(d[CL - 1]);
#pragma unroll
        for (int i = 0; i < G; i++) {
            cc[i] = R(R1_C, (NG - 1) * G + i);
            cd[i] = R(R1_DINV, (NG - 1) * G + i);
        }
#pragma unroll
        for (int g = NG - 1; g >= 0; g--) {
            if (g > 0) {
                if (g < NG - 1) R.fence(d[(g + 1) * G]);
#pragma unroll
                for (int i = 0; i < G; i++) {
                    nc[i] = R(R1_C, (g - 1) * G + i);
                    nd[i] = R(R1_DINV, (g - 1) * G + i);
                }
            }
#pragma unroll
            for (int i = G - 1; i >= 0; i--) {
                const int t = g * G + i;
                const T nx = d[t + 1 < CL ? t + 1 : CL - 1];
                const T v = (t == len - 1) ? d[t] * cd[i] : fma(-cc[i], nx, d[t]) * cd[i];
                d[t] = t < len ? v : d[t];
            }
            if (g > 0) {
#pragma unroll
                for (int i = 0; i < G; i++) {
                    cc[i] = nc[i];
                    cd[i] = nd[i];
                }
            }
        }
    } else {
        T a1[G], a2[G], b1[G], b2[G];
#pragma unroll
        for (int i = 0; i < G; i++) {
            a1[i] = R(R2_L1, i);
            a2[i] = R(R2_L2, i);
        }
#pragma unroll
        for (int g = 0; g < NG; g++) {
            if (g + 1 < NG) {
                if (g > 0) R.fence(d[g * G - 1]);
#pragma unroll
                for (int i = 0; i < G; i++) {
                    b1[i] = R(R2_L1, (g + 1) * G + i);
                    b2[i] = R(R2_L2, (g + 1) * G + i);
                }
            }
#pragma unroll
            for (int i = 0; i < G; i++) {
                const int t = g * G + i;
                if (t < CL - 1) {
                    const T xt = d[t];
                    const T n1 = fma(-a1[i], xt, d[t + 1]);
                    d[t + 1] = t < len - 1 ? n1 : d[t + 1];
                    if (t + 2 < CL) {
                        const T n2 = fma(-a2[i], xt, d[t + 2 < CL ? t + 2 : t]);
                        d[t + 2 < CL ? t + 2 : t] = t < len - 2 ? n2 : d[t + 2 < CL ? t + 2 : t];
                    }
                }
            }
            if (g + 1 < NG) {
#pragma unroll
                for (int i = 0; i < G; i++) {
                    a1[i] = b1[i];
                    a2[i] = b2[i];
                }
            }
        }
        T u0[G], u1[G], u2[G], w0[G], w1[G], w2[G];
        R.fence(d[CL - 1]);
#pragma unroll
        for (int i = 0; i < G; i++) {
            u0[i] = R(R2_U0INV, (NG - 1) * G + i);
            u1[i] = R(R2_U1, (NG - 1) * G + i);
            u2[i] = R(R2_U2, (NG - 1) * G + i);
        }
#pragma unroll
        for (int g = NG - 1; g >= 0; g--) {
            if (g > 0) {
                if (g < NG - 1) R.fence(d[(g + 1) * G]);
#pragma unroll
                for (int i = 0; i < G; i++) {
                    w0[i] = R(R2_U0INV, (g - 1) * G + i);
                    w1[i] = R(R2_U1, (g - 1) * G + i);
                    w2[i] = R(R2_U2, (g - 1) * G + i);
                }
            }
#pragma unroll
            for (int i = G - 1; i >= 0; i--) {
                const int t = g * G + i;
                T acc = d[t];
                if (t + 1 < CL) {
                    const T x1 = fma(-u1[i], d[t + 1 < CL ? t + 1 : t], acc);
                    acc = t + 1 < len ? x1 : acc;
                }
                if (t + 2 < CL) {
                    const T x2 = fma(-u2[i], d[t + 2 < CL ? t + 2 : t], acc);
                    acc = t + 2 < len ? x2 : acc;
                }
                d[t] = t < len ? acc * u0[i] : d[t];
            }
            if (g > 0) {
#pragma unroll
                for (int i = 0; i < G; i++) {
                    u0[i] = w0[i];
                    u1[i] = w1[i];
                    u2[i] = w2[i];
                }
            }
        }
    }
}

// x = y - V tail_prev - W head_next for every row of the chunk, the FP64 spike
// rows read in groups as above (Rows64(field, t) in bounds for t < CL).
template <typename T, int Q, int CL, int G, typename Rows64>
__device__ __forceinline__ void correct_grouped(T (&d)[CL], int len, Rows64 R, const double (&tp)[Q],
                                                const double (&hn)[Q])
{
    constexpr int NG = CL / G;
    constexpr int NF = 2 * Q;
    const int fld[4] = {Q == 1 ? (int)R1_V : (int)R2_V0, Q == 1 ? (int)R1_W : (int)R2_W0,
                        Q == 1 ? 0 : (int)R2_V1, Q == 1 ? 0 : (int)R2_W1};
    double cur[NF][G], nxt[NF][G];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int i = 0; i < G; i++) cur[f][i] = R(fld[f], i);
#pragma unroll
    for (int g = 0; g < NG; g++) {
        if (g + 1 < NG) {
            if (g > 0) R.fence(d[g * G - 1]);
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int i = 0; i < G; i++) nxt[f][i] = R(fld[f], (g + 1) * G + i);
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int t = g * G + i;
            double acc = fma(cur[1][i], hn[0], cur[0][i] * tp[0]);
            if constexpr (Q == 2) {
                acc = fma(cur[2][i], tp[1], acc);
                acc = fma(cur[3][i], hn[1], acc);
            }
            const T c = (T)((double)d[t] - acc);
            d[t] = t < len ? c : d[t];
        }
        if (g + 1 < NG) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int i = 0; i < G; i++) cur[f][i] = nxt[f][i];
        }
    }
}

// Row-major plan rows of the chunk starting at s (row index clamped to the line)
template <typename T>
struct RowsClamped {
    const T *base;
    int n, s;
    __device__ __forceinline__ T operator()(int f, int t) const
    {
        const int g = s + t < n ? s + t : n - 1;
        return gload(base + (size_t)f * n + g);
    }
    template <typename D> __device__ __forceinline__ void fence(const D &dep)
    {
        asm volatile("" : "+v"(base) : "v"(dep));
    }
};

// Reduced interface system: block PCR across the P chunks of each line
// (slot = chunk * nl + line-in-block), then fetch the previous chunk's tail
// and the next chunk's head.  Every thread of the block must call this.
// PF: each step's coefficient blocks are loaded one step ahead, before the
// exchange through LDS (the barriers wait for LDS traffic only), so their
// latency overlaps the exchange; for per-lane chunk indices (the x sweep).
template <int Q, bool PF = false>
__device__ __forceinline__ void reduced_solve(const LineDev &p, double *buf, int NT, int nl, int k, int li,
                                              double (&r)[2 * Q], double (&tp)[Q], double (&hn)[Q])
{
    constexpr int B = 2 * Q;
    const int P = p.P;
    const int slot = k * nl + li;
    const double *pc = p.pcr;
    if constexpr (PF) {
        constexpr int C = 2 * B * B;
        // one PCR step: multiplies by the blocks in cur and loads the next
        // step's into nxt (this step's again on the last step: in bounds, unused)
        auto step = [&](int st, double (&cur)[C], double (&nxt)[C]) {
            const int dd = 1 << st;
            const int sn = st + 1 < p.nsteps ? st + 1 : st;
            const double *an = pc + ((size_t)(sn * P + k) * 2) * B * B;
#pragma unroll
            for (int c = 0; c < C; c++) nxt[c] = gload(an + c);
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
#pragma unroll
            for (int i = 0; i < B; i++) {
                double acc = r[i];
#pragma unroll
                for (int j = 0; j < B; j++) {
                    acc = fma(cur[i * B + j], rm[j], acc);
                    acc = fma(cur[B * B + i * B + j], rp[j], acc);
                }
                r[i] = acc;
            }
        };
        double ca[C], cb[C];
        if (p.nsteps > 0) {
            const double *a0 = pc + (size_t)k * 2 * B * B;
#pragma unroll
            for (int c = 0; c < C; c++) ca[c] = gload(a0 + c);
        }
        // two steps per trip, so the prefetch buffers alternate without copies
        for (int st = 0; st < p.nsteps; st += 2) {
            step(st, ca, cb);
            if (st + 1 < p.nsteps) step(st + 1, cb, ca);
        }
    } else {
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
                for (int j = 0; j < B; j++) {
                    acc = fma(al[i * B + j], rm[j], acc);
                    acc = fma(be[i * B + j], rp[j], acc);
                }
                r[i] = acc;
            }
        }
    }
    const double *mi = pc + (size_t)p.nsteps * P * 2 * B * B + (size_t)k * B * B;
    double z[B];
#pragma unroll
    for (int i = 0; i < B; i++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < B; j++) acc = fma(mi[i * B + j], r[j], acc);
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
        return (double)y - fma(rw[(size_t)R1_W * n + g], hn[0], rw[(size_t)R1_V * n + g] * tp[0]);
    } else {
        double acc = rw[(size_t)R2_V0 * n + g] * tp[0];
        acc = fma(rw[(size_t)R2_W0 * n + g], hn[0], acc);
        acc = fma(rw[(size_t)R2_V1 * n + g], tp[1], acc);
        acc = fma(rw[(size_t)R2_W1 * n + g], hn[1], acc);
        return (double)y - acc;
    }
}

// ---------------------------------------------------------------------------
// y / z sweeps: lanes along x, one chunk row per load instruction.
// Line dimension stride sd, outer dimension stride so; block = Lx lanes x P
// chunks.  UNI: Lx == 64, so each wave holds one chunk and every per-row
// constant is wave-uniform (scalar loads).
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM, int EPI, bool UNI, int CL = MC, bool NTS = false>
__global__ void __launch_bounds__(UNI && Q == 1 ? 1024 : 512, UNI ? (Q == 2 ? 4 : 1) : 3) k_line_yz(LineDev p, const T *__restrict__ f, int sd, int64_t so,
                                                  int nlane, int nxb, int Lx, EpiArgs e, LineFields mf)
{
    select_field<T, EPI>(mf, f, e);
    extern __shared__ double smem[];
    const int NT = Lx * p.P;
    const int lane = UNI ? (int)(threadIdx.x & 63) : (int)(threadIdx.x % Lx);
    int k = UNI ? (int)(threadIdx.x >> 6) : (int)(threadIdx.x / Lx);
    if constexpr (UNI) k = __builtin_amdgcn_readfirstlane(k);
    const int xb = blockIdx.x % nxb;
    const int64_t ob = blockIdx.x / nxb;
    const int xi = xb * Lx + lane;
    const bool valid = xi < nlane;
    const int64_t off = (int64_t)xi + ob * so;
    const T *fl = f + off;
    const int n = p.n;
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);

    T fv[CL + 2 * M];
    if constexpr (UNI) {
        // one chunk per wave: the branches below are uniform
#pragma unroll
        for (int t = -M; t < CL + M; t++) {
            T val = T(0);
            if (valid && t < len + M) {
                int r = s + t;
                bool ok = true;
                if (r < 0) {
                    if (p.periodic) r += n; else ok = false;
                } else if (r >= n) {
                    if (p.periodic) r -= n; else ok = false;
                }
                if (ok) val = fl[r * sd];
            }
            fv[t + M] = val;
        }
    } else {
        // lanes of one wave own different chunks, so branching per load
        // would diverge: every load is unconditional, from the wrapped row
        // clamped into the line (and the last lane of the row set for lanes
        // past it), and its value selected afterwards
        const T *fa = f + (int64_t)(valid ? xi : nlane - 1) + ob * so;
#pragma unroll
        for (int t = -M; t < CL + M; t++) {
            int r = s + t;
            bool ok = valid && t < len + M;
            if (r < 0) {
                ok = ok && p.periodic;
                r += n;
            } else if (r >= n) {
                ok = ok && p.periodic;
                r -= n;
            }
            const int rc = r < 0 ? 0 : (r > n - 1 ? n - 1 : r);
            const T val = fa[(int64_t)rc * sd];
            fv[t + M] = ok ? val : T(0);
        }
    }
    T fe[8];
#pragma unroll
    for (int kap = 0; kap < 8; kap++) fe[kap] = T(0);
    if (p.closed && k == p.P - 1 && valid) {
#pragma unroll
        for (int kap = 0; kap < 8; kap++)
            if (kap < n) fe[kap] = fl[(n - 1 - kap) * sd];
    }

    Chunk<T, Q, M, SYM, CL> ch;
    ch.rhs(p, k, len, fv, fe);
    // grouped row loads where they fit the 128 registers of a 1024-thread
    // block (chunks of up to 16 points); wave-uniform rows (UNI) are scalar loads
    constexpr bool GRP = !UNI && CL <= 16;
    constexpr int G = CL <= 8 ? CL : 4;
    if constexpr (GRP)
        solve_grouped<T, Q, CL, G>(ch.d, len, RowsClamped<T>{rowp<T>(p), n, s});
    else
        ch.solve(p, s, len);
    double r[2 * Q], tp[Q], hn[Q];
    ch.ends(len, r);
    reduced_solve<Q, !UNI && Q == 1 && sizeof(T) == 8>(p, smem, NT, Lx, k, lane, r, tp, hn);
    if (!valid) return;
    if constexpr (GRP) {
        correct_grouped<T, Q, CL, G>(ch.d, len, RowsClamped<double>{p.row64, n, s}, tp, hn);
    } else {
#pragma unroll
        for (int t = 0; t < CL; t++)
            if (t < len) ch.d[t] = (T)corrected<T, Q>(p, s + t, ch.d[t], tp, hn);
    }
    // the device-dt LSRK epilogue (per-point gate and dt beside three
    // inputs) goes four points at a time at the full capacity (no spills)
    constexpr int NB = EPI == EPI_LSRK_DEV && CL > 16 ? 4 : 8;
#pragma unroll
    for (int b = 0; b < CL; b += NB) {
        int64_t o[NB];
        T v[NB];
        bool ok[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            ok[i] = b + i < len;
            o[i] = off + (int64_t)(s + b + i) * sd;
            v[i] = ch.d[b + i];
        }
        epilogue_batch<T, EPI, NB, NTS>(e, o, v, ok);
    }
}

// ---------------------------------------------------------------------------
// x sweep: Lt whole lines staged in LDS (row stride n + 1), threads own
// (line, chunk) pairs, tid = line + Lt * chunk.  Lines are indexed
// l = j + ny*k over the y-z plane.  UNI: Lt == 64 (one chunk per wave).
// ---------------------------------------------------------------------------
template <typename T, int Q, int M, int SYM, int EPI, bool UNI, int CL = MC>
__global__ void __launch_bounds__(UNI && Q == 1 ? 1024 : 512, !UNI && CL <= 8 && Q == 1 ? 5 : 1) k_line_x(LineDev p, const T *__restrict__ f, int64_t sy, int64_t sz,
                                                 int ny, int64_t nlines, int Lt, EpiArgs e, LineFields mf)
{
    select_field<T, EPI>(mf, f, e);
    extern __shared__ double smem[];
    const int n = p.n;
    const int TS = n + 1;
    const int NT = Lt * p.P;
    T *tile = (T *)smem;
    double *buf = smem + (((size_t)Lt * TS * sizeof(T) + 7) / 8);
    const int tid = threadIdx.x;
    const int64_t l0 = (int64_t)blockIdx.x * Lt;

    if constexpr (!UNI && CL <= 8) {
        // element q = tid + NT i of the block's Lt lines is line q / n, point
        // q % n; a thread holds at most ceil(n / P) <= CL of them, and all its
        // loads are issued before the LDS writes
        const int nel = Lt * n;
        const int j0 = (int)(l0 % ny), k0 = (int)(l0 / ny);
        T tmp[CL];
        int dst[CL];
#pragma unroll
        for (int i = 0; i < CL; i++) {
            const int q = tid + NT * i;
            const int lj = q / n, x = q - lj * n;
            int j = j0 + lj, kk = k0;
            while (j >= ny) {
                j -= ny;
                kk++;
            }
            const bool ok = q < nel && l0 + lj < nlines;
            tmp[i] = ok ? f[(int64_t)j * sy + (int64_t)kk * sz + x] : T(0);
            dst[i] = q < nel ? lj * TS + x : -1;
        }
#pragma unroll
        for (int i = 0; i < CL; i++)
            if (dst[i] >= 0) tile[dst[i]] = tmp[i];
    } else {
        for (int li = 0; li < Lt; li++) {
            const int64_t l = l0 + li;
            if (l >= nlines) break;
            const T *src = f + (l % ny) * sy + (l / ny) * sz;
            for (int i = tid; i < n; i += NT) tile[li * TS + i] = src[i];
        }
    }
    __syncthreads();

    const int li = UNI ? (tid & 63) : (tid % Lt);
    int k = UNI ? (tid >> 6) : (tid / Lt);
    if constexpr (UNI) k = __builtin_amdgcn_readfirstlane(k);
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);
    const T *trow = tile + li * TS;
    T fv[CL + 2 * M];
#pragma unroll
    for (int t = -M; t < CL + M; t++) {
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

    Chunk<T, Q, M, SYM, CL> ch;
    ch.rhs(p, k, len, fv, fe);
    // grouped row loads where they fit the 128 registers of a 1024-thread
    // block (chunks of up to 16 points); wave-uniform rows (UNI) are scalar loads
    constexpr bool GRP = !UNI && CL <= 16;
    constexpr int G = CL <= 8 ? CL : 4;
    if constexpr (GRP)
        solve_grouped<T, Q, CL, G>(ch.d, len, RowsClamped<T>{rowp<T>(p), n, s});
    else
        ch.solve(p, s, len);
    double r[2 * Q], tp[Q], hn[Q];
    ch.ends(len, r);
    reduced_solve<Q, !UNI && Q == 1>(p, buf, NT, Lt, k, li, r, tp, hn);
    // every thread has finished reading the tile before the first reduced-solve barrier
    T *wrow = tile + li * TS;
    if constexpr (GRP) {
        correct_grouped<T, Q, CL, G>(ch.d, len, RowsClamped<double>{p.row64, n, s}, tp, hn);
#pragma unroll
        for (int t = 0; t < CL; t++)
            if (t < len) wrow[s + t] = ch.d[t];
    } else {
#pragma unroll
        for (int t = 0; t < CL; t++)
            if (t < len) wrow[s + t] = (T)corrected<T, Q>(p, s + t, ch.d[t], tp, hn);
    }
    __syncthreads();

    for (int lj = 0; lj < Lt; lj++) {
        const int64_t l = l0 + lj;
        if (l >= nlines) break;
        const int64_t off = (l % ny) * sy + (l / ny) * sz;
        for (int i = tid; i < n; i += NT) epilogue<T, EPI>(e, off + i, (double)tile[lj * TS + i]);
    }
}

// ---------------------------------------------------------------------------
// x sweep, wave-local: a line of n <= 64*XCL points is split into P <= 64
// chunks of <= XCL contiguous points; lane (li, k) of a wave owns chunk k of
// line li (64/P lines per wave).  The reduced interface system is solved by
// block PCR through cross-lane shuffles (no LDS tile, no barrier).
// UNIROW: periodic line with equal chunks, so every chunk has the same
// factors, spikes and PCR blocks (those of chunk 0): all per-row constants are
// wave-uniform.
// ---------------------------------------------------------------------------
constexpr int XCL = 16;

// PF: each step's coefficient blocks are loaded one step ahead (per-lane
// chunk indices kc, whose blocks are vector loads), so their latency overlaps
// the previous step's shuffles and products.
template <int Q, bool PF = false>
__device__ __forceinline__ void reduced_solve_wave(const LineDev &p, int k, int kc, int base, double (&r)[2 * Q],
                                                   double (&tp)[Q], double (&hn)[Q])
{
    constexpr int B = 2 * Q;
    const int P = p.P;
    const double *pc = p.pcr;
    if constexpr (PF) {
        constexpr int C = 2 * B * B;
        // one PCR step: multiplies by the blocks in cur and loads the next
        // step's into nxt (this step's again on the last step: in bounds, unused)
        auto step = [&](int st, double (&cur)[C], double (&nxt)[C]) {
            const int dd = 1 << st;
            const int sn = st + 1 < p.nsteps ? st + 1 : st;
            const double *an = pc + ((size_t)(sn * P + kc) * 2) * B * B;
#pragma unroll
            for (int c = 0; c < C; c++) nxt[c] = gload(an + c);
            int km = k - dd, kp = k + dd;
            bool hm = true, hp = true;
            if (p.periodic) {
                km &= (P - 1);
                kp &= (P - 1);
            } else {
                hm = km >= 0;
                hp = kp < P;
                km = hm ? km : k;
                kp = hp ? kp : k;
            }
            double rm[B], rp[B];
#pragma unroll
            for (int c = 0; c < B; c++) {
                rm[c] = __shfl(r[c], base + km, 64);
                rp[c] = __shfl(r[c], base + kp, 64);
                if (!hm) rm[c] = 0.0;
                if (!hp) rp[c] = 0.0;
            }
#pragma unroll
            for (int i = 0; i < B; i++) {
                double acc = r[i];
#pragma unroll
                for (int j = 0; j < B; j++) {
                    acc = fma(cur[i * B + j], rm[j], acc);
                    acc = fma(cur[B * B + i * B + j], rp[j], acc);
                }
                r[i] = acc;
            }
        };
        double ca[C], cb[C];
        if (p.nsteps > 0) {
            const double *a0 = pc + (size_t)kc * 2 * B * B;
#pragma unroll
            for (int c = 0; c < C; c++) ca[c] = gload(a0 + c);
        }
        // two steps per trip, so the prefetch buffers alternate without copies
        for (int st = 0; st < p.nsteps; st += 2) {
            step(st, ca, cb);
            if (st + 1 < p.nsteps) step(st + 1, cb, ca);
        }
    } else {
        for (int st = 0, dd = 1; st < p.nsteps; st++, dd <<= 1) {
            int km = k - dd, kp = k + dd;
            bool hm = true, hp = true;
            if (p.periodic) {
                km &= (P - 1);
                kp &= (P - 1);
            } else {
                hm = km >= 0;
                hp = kp < P;
                km = hm ? km : k;
                kp = hp ? kp : k;
            }
            double rm[B], rp[B];
#pragma unroll
            for (int c = 0; c < B; c++) {
                rm[c] = __shfl(r[c], base + km, 64);
                rp[c] = __shfl(r[c], base + kp, 64);
                if (!hm) rm[c] = 0.0;
                if (!hp) rp[c] = 0.0;
            }
            const double *al = pc + ((size_t)(st * P + kc) * 2) * B * B;
            const double *be = al + B * B;
#pragma unroll
            for (int i = 0; i < B; i++) {
                double acc = r[i];
#pragma unroll
                for (int j = 0; j < B; j++) {
                    acc = fma(gload(al + i * B + j), rm[j], acc);
                    acc = fma(gload(be + i * B + j), rp[j], acc);
                }
                r[i] = acc;
            }
        }
    }
    const double *mi = pc + (size_t)p.nsteps * P * 2 * B * B + (size_t)kc * B * B;
    double z[B];
#pragma unroll
    for (int i = 0; i < B; i++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < B; j++) acc = fma(gload(mi + i * B + j), r[j], acc);
        z[i] = acc;
    }
    int kprev = k - 1, knext = k + 1;
    bool hasp = true, hasn = true;
    if (p.periodic) {
        kprev &= (P - 1);
        knext &= (P - 1);
    } else {
        hasp = kprev >= 0;
        hasn = knext < P;
        kprev = hasp ? kprev : k;
        knext = hasn ? knext : k;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const double a = __shfl(z[Q + i], base + kprev, 64);
        const double b = __shfl(z[i], base + knext, 64);
        tp[i] = hasp ? a : 0.0;
        hn[i] = hasn ? b : 0.0;
    }
}

// One wave per block.  The wave's 64/P lines are staged in LDS (a pad slot
// after every XC points keeps the chunk-strided reads conflict-free), so
// global loads and stores are contiguous 512-byte rows.  XC is the chunk
// capacity (8 or 16): the sweeps run XC steps, so chunks of at most 8 points
// take the 8-step kernel.
template <int XC = XCL>
__device__ __forceinline__ int xw_slot(int x) { return x + x / XC; }

template <typename T, int Q, int M, int SYM, int EPI, bool UNIROW, int XC>
__global__ void __launch_bounds__(64, 4) k_line_xw(LineDev p, const T *__restrict__ f, int64_t sy, int64_t sz, int ny,
                                                int64_t nlines, int TS, EpiArgs e, LineFields mf)
{
    select_field<T, EPI>(mf, f, e);
    extern __shared__ double smem[];
    T *tile = (T *)smem;
    const int P = p.P;
    const int n = p.n;
    const int lane = threadIdx.x;
    const int nl = 64 / P;           // lines staged by the wave
    const int64_t l0 = (int64_t)blockIdx.x * nl;
    // stage the lines: lane-element i of the block's nl*n points is line lj,
    // point x (shifts when n is a power of two >= 64, else a division); every
    // load is issued before the LDS writes.
    const int cnt = (nl * n + 63) / 64;
    const bool pow2 = n >= 64 && (n & (n - 1)) == 0;
    const int sh = pow2 ? __builtin_ctz((unsigned)(n >> 6)) : 0;
    const int j0 = (int)(l0 % ny), k0 = (int)(l0 / ny);
    auto line_of = [&](int i, int &lj, int &x) {
        if (pow2) {
            lj = i >> sh;
            x = lane + 64 * (i & ((1 << sh) - 1));
        } else {
            const int q = lane + 64 * i;
            lj = q / n;
            x = q - lj * n;
        }
    };
    auto line_off = [&](int lj) -> int64_t {
        int j = j0 + lj, k = k0;
        while (j >= ny) {
            j -= ny;
            k++;
        }
        return (int64_t)j * sy + (int64_t)k * sz;
    };
    {
        T tmp[XC];
#pragma unroll
        for (int i = 0; i < XC; i++) {
            tmp[i] = T(0);
            int lj, x;
            line_of(i, lj, x);
            if (i < cnt && lj < nl && l0 + lj < nlines) tmp[i] = f[line_off(lj) + x];
        }
#pragma unroll
        for (int i = 0; i < XC; i++) {
            int lj, x;
            line_of(i, lj, x);
            if (i < cnt && lj < nl) tile[lj * TS + xw_slot<XC>(x)] = tmp[i];
        }
    }
    __syncthreads();
    const int k = lane & (P - 1);
    const int base = (lane / P) * P;
    const int s = k * p.m + min(k, p.extra);
    const int len = p.m + (k < p.extra ? 1 : 0);
    const int rs = UNIROW ? 0 : s;
    const int kc = UNIROW ? 0 : k;
    T res[XC];
    {
        const T *trow = tile + (lane / P) * TS;
        T fv[XC + 2 * M];
#pragma unroll
        for (int t = -M; t < XC + M; t++) {
            T val = T(0);
            if (t < len + M) {
                int r = s + t;
                bool ok = true;
                if (r < 0) {
                    if (p.periodic) r += n; else ok = false;
                } else if (r >= n) {
                    if (p.periodic) r -= n; else ok = false;
                }
                if (ok) val = trow[xw_slot<XC>(r)];
            }
            fv[t + M] = val;
        }
        T fe[8];
#pragma unroll
        for (int kap = 0; kap < 8; kap++) fe[kap] = (p.closed && kap < n) ? trow[xw_slot<XC>(n - 1 - kap)] : T(0);
        Chunk<T, Q, M, SYM, XC> ch;
        ch.rhs(p, k, len, fv, fe);
        solve_grouped<T, Q, XC, 8>(ch.d, len, RowsClamped<T>{rowp<T>(p), n, rs});
        double r[2 * Q], tp[Q], hn[Q];
        ch.ends(len, r);
        reduced_solve_wave<Q, !UNIROW && Q == 1 && sizeof(T) == 8>(p, k, kc, base, r, tp, hn);
        correct_grouped<T, Q, XC, 8>(ch.d, len, RowsClamped<double>{p.row64, n, rs}, tp, hn);
#pragma unroll
        for (int t = 0; t < XC; t++) res[t] = ch.d[t];
    }
    __syncthreads();   // all window reads done before the tile is overwritten
    {
        T *wrow = tile + (lane / P) * TS;
#pragma unroll
        for (int t = 0; t < XC; t++)
            if (t < len) wrow[xw_slot<XC>(s + t)] = res[t];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < XC; b += 8) {
        int64_t o[8];
        T v[8];
        bool ok[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            int lj, x;
            line_of(b + i, lj, x);
            ok[i] = (b + i) < cnt && lj < nl && l0 + lj < nlines;
            o[i] = ok[i] ? line_off(lj) + x : 0;
            v[i] = ok[i] ? tile[lj * TS + xw_slot<XC>(x)] : T(0);
        }
        epilogue_batch<T, EPI, 8>(e, o, v, ok);
    }
}

// x sweep for lines of exactly 16 P points in equal 16-point chunks (the
// production grids): the same algorithm as k_line_xw with the line length,
// chunk length, lines per wave (NL = 64 / P) and the element-to-line map fixed
// at compile time, line offsets computed once per wave, a mask for the
// periodic wrap, and no closure code on periodic lines.
template <typename T, int Q, int M, int SYM, int EPI, int NL, bool PER>
__global__ void __launch_bounds__(64) k_line_xwf(LineDev p, const T *__restrict__ f, int64_t sy, int64_t sz, int ny,
                                                 int64_t nlines, EpiArgs e, LineFields mf)
{
    select_field<T, EPI>(mf, f, e);
    constexpr int P = 64 / NL;            // chunks (lanes) per line
    constexpr int N = XCL * P;            // points per line
    constexpr int TS = N + N / XCL + 1;   // padded tile row
    constexpr int EPL = N / 64;           // staged elements per lane and line
    __shared__ T tile[NL * TS];
    const int lane = threadIdx.x;
    const int64_t l0 = (int64_t)blockIdx.x * NL;
    int64_t loff[NL];
    bool lval[NL];
#pragma unroll
    for (int lj = 0; lj < NL; lj++) {
        const int64_t l = l0 + lj;
        lval[lj] = l < nlines;
        loff[lj] = (l % ny) * sy + (l / ny) * sz;
    }
    {
        T tmp[NL * EPL];
#pragma unroll
        for (int lj = 0; lj < NL; lj++)
#pragma unroll
            for (int q = 0; q < EPL; q++) tmp[lj * EPL + q] = lval[lj] ? f[loff[lj] + lane + 64 * q] : T(0);
#pragma unroll
        for (int lj = 0; lj < NL; lj++)
#pragma unroll
            for (int q = 0; q < EPL; q++) tile[lj * TS + xw_slot(lane + 64 * q)] = tmp[lj * EPL + q];
    }
    __syncthreads();
    const int li = lane / P, k = lane % P;
    const int s = k * XCL;
    T *trow = tile + li * TS;
    T fv[XCL + 2 * M];
#pragma unroll
    for (int t = -M; t < XCL + M; t++) {
        int r = s + t;
        if constexpr (PER) {
            fv[t + M] = trow[xw_slot(r & (N - 1))];
        } else {
            fv[t + M] = (r >= 0 && r < N) ? trow[xw_slot(r < 0 ? 0 : (r >= N ? N - 1 : r))] : T(0);
        }
    }
    T fe[8];
#pragma unroll
    for (int kap = 0; kap < 8; kap++) fe[kap] = PER ? T(0) : trow[xw_slot(N - 1 - kap)];
    Chunk<T, Q, M, SYM, XCL> ch;
    if constexpr (PER) {
        const T *cf = coefp<T>(p);
        T ci[M];
#pragma unroll
        for (int m = 0; m < M; m++) ci[m] = cf[COEF_CI + m];
        const T a0 = cf[COEF_A0];
#pragma unroll
        for (int t = 0; t < XCL; t++) ch.d[t] = rhs_row<T, M, SYM>(fv, t + M, ci, a0);
    } else {
        ch.rhs(p, k, XCL, fv, fe);
    }
    const int rs = PER ? 0 : s;
    const int kc = PER ? 0 : k;
    ch.solve(p, rs, XCL);
    double r[2 * Q], tp[Q], hn[Q];
    ch.ends(XCL, r);
    reduced_solve_wave<Q>(p, k, kc, li * P, r, tp, hn);
    T res[XCL];
#pragma unroll
    for (int t = 0; t < XCL; t++) res[t] = (T)corrected<T, Q>(p, rs + t, ch.d[t], tp, hn);
    __syncthreads();   // all window reads done before the tile is overwritten
#pragma unroll
    for (int t = 0; t < XCL; t++) trow[xw_slot(s + t)] = res[t];
    __syncthreads();
#pragma unroll
    for (int lj = 0; lj < NL; lj++) {
#pragma unroll
        for (int q0 = 0; q0 < EPL; q0 += 8) {
            int64_t o[8];
            T v[8];
            bool ok[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int q = q0 + i;
                const bool in = q < EPL && lval[lj];
                const int x = lane + 64 * (q < EPL ? q : 0);
                ok[i] = in;
                o[i] = loff[lj] + x;
                v[i] = tile[lj * TS + xw_slot(x)];
            }
            epilogue_batch<T, EPI, 8>(e, o, v, ok);
        }
    }
}

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
inline size_t lds_limit()
{
    static size_t lim = 0;
    if (!lim) {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        lim = (size_t)v;
    }
    return lim;
}

template <typename K>
inline void allow_lds(K kern, size_t shm)
{
    if (shm > 64 * 1024)
        CLES_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
}

template <typename T, int Q, int M, int SYM, int EPI>
void launch_variant(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e,
                           hipStream_t st, const LineFields &mf)
{
    const unsigned nf = (unsigned)mf.nf;
    const LineDev &p = hp.dev;

    const int P = p.P;
    const int B = 2 * Q;
    const size_t lim = lds_limit();
    const int maxNT = Q == 1 ? 1024 : 512;
    if (dim == 1 && p.extra == 0 && p.m == XCL && P == 16 && (p.periodic || p.closed)) {
        // 256-point lines: four per wave
        const int64_t nlines = v.n[1] * v.n[2];
        const int64_t nblk = (nlines + 3) / 4;
        if (p.periodic)
            hipLaunchKernelGGL((k_line_xwf<T, Q, M, SYM, EPI, 4, true>), dim3((unsigned)nblk, nf), dim3(64), 0, st, p, f,
                               v.s[1], v.s[2], (int)v.n[1], nlines, e, mf);
        else
            hipLaunchKernelGGL((k_line_xwf<T, Q, M, SYM, EPI, 4, false>), dim3((unsigned)nblk, nf), dim3(64), 0, st, p, f,
                               v.s[1], v.s[2], (int)v.n[1], nlines, e, mf);
    } else if (dim == 1 && P <= 64 && p.m + (p.extra ? 1 : 0) <= XCL) {
        const int64_t nlines = v.n[1] * v.n[2];
        const int nl = 64 / P;
        const int64_t nblk = (nlines + nl - 1) / nl;
        const bool uni = p.periodic && p.extra == 0;
        auto go = [&](auto kern, int xc) {
            const int TS = p.n + p.n / xc + 1;
            const size_t shm = (size_t)nl * TS * sizeof(T);
            hipLaunchKernelGGL(kern, dim3((unsigned)nblk, nf), dim3(64), shm, st, p, f, v.s[1], v.s[2], (int)v.n[1],
                               nlines, TS, e, mf);
        };
        if (p.m + (p.extra ? 1 : 0) <= 8) {
            if (uni) go(k_line_xw<T, Q, M, SYM, EPI, true, 8>, 8);
            else go(k_line_xw<T, Q, M, SYM, EPI, false, 8>, 8);
        } else {
            if (uni) go(k_line_xw<T, Q, M, SYM, EPI, true, XCL>, XCL);
            else go(k_line_xw<T, Q, M, SYM, EPI, false, XCL>, XCL);
        }
    } else if (dim == 1) {
        const size_t tile_row = (size_t)(p.n + 1) * sizeof(T);
        const int64_t nlines = v.n[1] * v.n[2];
        auto shm_of = [&](int Lt) {
            return (((size_t)Lt * tile_row + 7) / 8) * 8 + (size_t)2 * B * Lt * P * sizeof(double);
        };
        if (64 * P <= maxNT && shm_of(64) <= lim && nlines >= 64) {
            const int Lt = 64, NT = 64 * P;
            const size_t shm = shm_of(Lt);
            auto kern = k_line_x<T, Q, M, SYM, EPI, true>;
            allow_lds(kern, shm);
            const int64_t nblk = (nlines + Lt - 1) / Lt;
            hipLaunchKernelGGL(kern, dim3((unsigned)nblk, nf), dim3(NT), shm, st, p, f, v.s[1], v.s[2], (int)v.n[1],
                               nlines, Lt, e, mf);
        } else {
            int Lt = 1;
            while (Lt * 2 * P <= 512 && shm_of(Lt * 2) <= 48 * 1024 && Lt < 32) Lt *= 2;
            if (Lt * P > 512 || shm_of(Lt) > lim)
                throw Error{CLES_ERR_UNSUPPORTED, "line too long for the x-sweep kernel"};
            const int NT = Lt * P;
            const size_t shm = shm_of(Lt);
            const int64_t nblk = (nlines + Lt - 1) / Lt;
            const int cl = p.m + (p.extra ? 1 : 0);
            auto go = [&](auto kern) {
                allow_lds(kern, shm);
                hipLaunchKernelGGL(kern, dim3((unsigned)nblk, nf), dim3(NT), shm, st, p, f, v.s[1], v.s[2],
                                   (int)v.n[1], nlines, Lt, e, mf);
            };
            if (cl <= 8)
                go(k_line_x<T, Q, M, SYM, EPI, false, 8>);
            else if (cl <= 16)
                go(k_line_x<T, Q, M, SYM, EPI, false, 16>);
            else
                go(k_line_x<T, Q, M, SYM, EPI, false, MC>);
        }
    } else {
        const int64_t nlane = v.n[0];
        const int64_t sd = dim == 2 ? v.s[1] : v.s[2];
        const int64_t so = dim == 2 ? v.s[2] : v.s[1];
        const int64_t nouter = dim == 2 ? v.n[2] : v.n[1];
        if ((int64_t)(p.n - 1) * sd > INT32_MAX) throw Error{CLES_ERR_UNSUPPORTED, "array too large for 32-bit line offsets"};
        if (nlane >= 48 && 64 * P <= maxNT) {
            const int Lx = 64, NT = 64 * P;
            const int nxb = (int)((nlane + Lx - 1) / Lx);
            const size_t shm = (size_t)2 * B * NT * sizeof(double);
            // The z sweep's plain stores take the non-temporal hint when the launch
            // writes more than the 256 MB Infinity Cache holds (the outputs are
            // evicted before they are read again); the x and y sweeps measured
            // faster without it.
            const bool narrow = p.m + (p.extra ? 1 : 0) <= 16;
            bool nts = false;
            if constexpr (EPI == EPI_STORE || EPI == EPI_STORE_NEG) {
                nts = dim == 3 && (double)v.n[0] * v.n[1] * v.n[2] * sizeof(T) * nf > 256.0 * 1024 * 1024;
                if (nts && narrow)
                    hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, true, 16, true>), dim3((unsigned)(nxb * nouter), nf),
                                       dim3(NT), shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
                else if (nts)
                    hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, true, MC, true>), dim3((unsigned)(nxb * nouter), nf),
                                       dim3(NT), shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
            }
            if (!nts && p.m + (p.extra ? 1 : 0) <= 8)
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, true, 8>), dim3((unsigned)(nxb * nouter), nf), dim3(NT),
                                   shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
            else if (!nts && narrow)
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, true, 16>), dim3((unsigned)(nxb * nouter), nf), dim3(NT),
                                   shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
            else if (!nts)
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, true>), dim3((unsigned)(nxb * nouter), nf), dim3(NT),
                                   shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
        } else {
            int Lx = 64;
            while (Lx * P > 512 && Lx > 1) Lx /= 2;
            while (Lx / 2 >= nlane && Lx > 1) Lx /= 2;   // narrow x extents (2-D / 1-D grids)
            // few lines (2-D grids): narrower blocks, down to 8 lanes (64-byte
            // rows), until the launch has two blocks per CU
            while (Lx > 8 && ((nlane + Lx - 1) / Lx) * nouter * (int64_t)nf < 512) Lx /= 2;
            if (Lx * P > 512) throw Error{CLES_ERR_UNSUPPORTED, "line too long for the y/z-sweep kernel"};
            const int NT = Lx * P;
            const int nxb = (int)((nlane + Lx - 1) / Lx);
            const size_t shm = (size_t)2 * B * NT * sizeof(double);
            const int cl = p.m + (p.extra ? 1 : 0);
            if (cl <= 8)
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, false, 8>), dim3((unsigned)(nxb * nouter), nf),
                                   dim3(NT), shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
            else if (cl <= 16)
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, false, 16>), dim3((unsigned)(nxb * nouter), nf),
                                   dim3(NT), shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
            else
                hipLaunchKernelGGL((k_line_yz<T, Q, M, SYM, EPI, false>), dim3((unsigned)(nxb * nouter), nf),
                                   dim3(NT), shm, st, p, f, (int)sd, so, (int)nlane, nxb, Lx, e, mf);
        }
    }
    CLES_HIP(hipGetLastError());
}

template <typename T, int Q, int M, int SYM>
void launch_epi(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e, hipStream_t st,
                const LineFields &mf)
{
    if constexpr (SYM) {
        switch (e.kind) {
        case EPI_STORE: return launch_variant<T, Q, M, SYM, EPI_STORE>(hp, dim, v, f, e, st, mf);
        case EPI_STORE_GATE: return launch_variant<T, Q, M, SYM, EPI_STORE_GATE>(hp, dim, v, f, e, st, mf);
        case EPI_ART_MUBETA: return launch_variant<T, Q, M, SYM, EPI_ART_MUBETA>(hp, dim, v, f, e, st, mf);
        case EPI_ART_KAPPA: return launch_variant<T, Q, M, SYM, EPI_ART_KAPPA>(hp, dim, v, f, e, st, mf);
        case EPI_ART_D: return launch_variant<T, Q, M, SYM, EPI_ART_D>(hp, dim, v, f, e, st, mf);
        case EPI_ART_MULTI: return launch_variant<T, Q, M, SYM, EPI_ART_MULTI>(hp, dim, v, f, e, st, mf);
        default: break;
        }
    } else {
        switch (e.kind) {
        case EPI_STORE: return launch_variant<T, Q, M, SYM, EPI_STORE>(hp, dim, v, f, e, st, mf);
        case EPI_STORE_NEG: return launch_variant<T, Q, M, SYM, EPI_STORE_NEG>(hp, dim, v, f, e, st, mf);
        case EPI_SUB: return launch_variant<T, Q, M, SYM, EPI_SUB>(hp, dim, v, f, e, st, mf);
        case EPI_LSRK: return launch_variant<T, Q, M, SYM, EPI_LSRK>(hp, dim, v, f, e, st, mf);
        case EPI_LSRK_DEV: return launch_variant<T, Q, M, SYM, EPI_LSRK_DEV>(hp, dim, v, f, e, st, mf);
        default: break;
        }
    }
    throw Error{CLES_ERR_UNSUPPORTED, "epilogue not available for this scheme kind"};
}

}  // namespace cles
