// Pointwise and stencil kernels of the per-step path (everything that is not
// a compact line solve).  Fields are unpadded column-major arrays of the
// interior, element (i,j,k) at i + nx*(j + ny*k).  W is the working type of
// the primitive / gradient / flux fields (double, or float in the mixed
// form); the conserved state Q and the LSRK register du are always double.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cles_point.h"

// Pointwise arithmetic mirrors the reference operation by operation (Julia
// does not contract a*b+c without @fastmath/@muladd), so primitives, sensor
// inputs and fluxes reproduce the CPU path to the last bit where the
// operation order is the same.
#pragma clang fp contract(off)

namespace cles {

namespace {

constexpr int TPB = 256;

inline unsigned nblocks(int64_t n, int per = TPB) { return (unsigned)((n + per - 1) / per); }

__device__ __forceinline__ double d_max(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ float d_max(float a, float b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Nasa9Mixture thermodynamics (physics.jl:290-310, 432-498).  Powers are
// formed by multiplication (within 2 ulp of Julia's integer powers).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double jl_max(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

__device__ __forceinline__ const cles_nasa9_interval &n9_interval(const cles_nasa9_species &s, double T)
{
    for (int i = 0; i < s.n_intervals - 1; i++)
        if (T <= s.iv[i].Tmax) return s.iv[i];
    return s.iv[s.n_intervals - 1];
}
__device__ __forceinline__ double n9_cp_over_R(const cles_nasa9_interval &v, double T)
{
    const double *a = v.a;
    const double T2 = T * T, T3 = T2 * T, T4 = T2 * T2;
    return a[0] / T2 + a[1] / T + a[2] + a[3] * T + a[4] * T2 + a[5] * T3 + a[6] * T4;
}
__device__ __forceinline__ double n9_h_over_R(const cles_nasa9_interval &v, double T)
{
    const double *a = v.a;
    const double T2 = T * T, T3 = T2 * T, T4 = T2 * T2, T5 = T4 * T;
    return -a[0] / T + a[1] * log(T) + a[2] * T + a[3] * T2 / 2 + a[4] * T3 / 3 + a[5] * T4 / 4 + a[6] * T5 / 5 +
           v.b1;
}
__device__ __forceinline__ double n9_cp(const N9Table &t, int k, double T)
{
    return t.R[k] * n9_cp_over_R(n9_interval(t.sp[k], T), T);
}
__device__ __forceinline__ double n9_h(const N9Table &t, int k, double T)
{
    return t.R[k] * n9_h_over_R(n9_interval(t.sp[k], T), T);
}
__device__ __forceinline__ double n9_e(const N9Table &t, int k, double T) { return n9_h(t, k, T) - t.R[k] * T; }

// physics.jl:459-482 mixture_temperature: bounded Newton on sum_k Y_k e_k(T) = e
__device__ double n9_temperature(const N9Table &t, double e, const double *Y)
{
    double e0 = 0.0, cv0 = 0.0;
    for (int k = 0; k < t.ns; k++) {
        e0 += Y[k] * n9_e(t, k, t.T_guess);
        cv0 += Y[k] * (n9_cp(t, k, t.T_guess) - t.R[k]);
    }
    double T = t.T_guess;
    if (cv0 > 0) {
        const double x = t.T_guess + (e - e0) / cv0;
        T = x > 1e9 ? 1e9 : (x < 1e-3 ? 1e-3 : x);
    }
    for (int it = 0; it < 30; it++) {
        double f = -e, cvm = 0.0;
        for (int k = 0; k < t.ns; k++) {
            f += Y[k] * n9_e(t, k, T);
            cvm += Y[k] * (n9_cp(t, k, T) - t.R[k]);
        }
        if (!(cvm > 0)) break;
        const double dl = f / cvm;
        T = jl_max(T - dl, 0.25 * T);
        if (fabs(dl) <= 1e-14 * T) break;
    }
    return T;
}

// ---------------------------------------------------------------------------
// physics.jl:604-653 _primitives! (IdealMixture), over the interior
// ---------------------------------------------------------------------------
// The primitives are read again by the gradient sweeps after the whole
// launch; when they exceed the 256 MB Infinity Cache they are written with the
// non-temporal hint (a compile-time choice, as for the fluxes).
template <bool NT, typename W>
__device__ __forceinline__ void prim_store(W *p, W v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// One-species ideal gas: the cp and Y_1 that k_primitives stores at a point
// of density rho > 0, formed by the same operations in the same order (with
// one species its sums over species start from 0 and take one term), so the
// kernels that read them get the same bits when they are not stored.
__device__ __forceinline__ void one_species_cp_y(double rho, double R1, double cv1, double &cp, double &Y)
{
    const double ri = 1 / rho;
    double Rm = rho * R1, cvm = rho * cv1;
    Rm *= ri;
    cvm *= ri;
    cp = cvm + Rm;
    Y = rho * ri;
}

// IG: an ideal-gas mixture (no stiffened gas, no NASA-9 tables); that
// instantiation compiles the other branches out, which frees registers
template <typename W, bool NT, bool IG>
__global__ void __launch_bounds__(TPB) k_primitives(PrimArgs<W> a, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    const int Ns = a.Ns;
    double rho = 0.0;
#pragma unroll
    for (int sp = 0; sp < CLES_MAXSP; sp++)
        if (sp < Ns) rho += a.Q[(int64_t)sp * N + I];
    if (!IG && a.sg) {
        // physics.jl:216-250 _primitives! (StiffenedGas): p = (gamma - 1) rho e - gamma p_inf
        const double g = a.sg_gamma, pinf = a.sg_pinf, cv = a.sg_cv, R = (g - 1) * cv;
        if (a.esens) {
            // a point with rho <= 0 keeps _primitives!' placeholder density 1
            // (physics.jl:242), which artificial.jl:321 divides by
            const double r = rho > 0 ? rho : 1.0;
            const double m1 = a.Q[(int64_t)Ns * N + I], m2 = a.Q[(int64_t)(Ns + 1) * N + I],
                         m3 = a.Q[(int64_t)(Ns + 2) * N + I];
            const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r;
            prim_store<NT>(&a.esens[I], (W)((a.Q[(int64_t)(Ns + 3) * N + I] - ke) / r));
        }
        prim_store<NT>(&a.Y[0][I], (W)(W(1)));
        prim_store<NT>(&a.cp[I], (W)(g * cv));
        if (rho > 0) {
            const double ri = 1 / rho;
            const double u = a.Q[(int64_t)Ns * N + I] * ri;
            const double v = a.Q[(int64_t)(Ns + 1) * N + I] * ri;
            const double w = a.Q[(int64_t)(Ns + 2) * N + I] * ri;
            const double e = a.Q[(int64_t)(Ns + 3) * N + I] * ri - 0.5 * (u * u + v * v + w * w);
            const double p = (g - 1) * rho * e - g * pinf;
            const double cc = g * (p + pinf) * ri;
            prim_store<NT>(&a.rho[I], (W)rho);
            prim_store<NT>(&a.u[I], (W)u);
            prim_store<NT>(&a.v[I], (W)v);
            prim_store<NT>(&a.w[I], (W)w);
            prim_store<NT>(&a.p[I], (W)p);
            prim_store<NT>(&a.T[I], (W)d_max((p + pinf) / (rho * R), 1e-300));
            prim_store<NT>(&a.c[I], (W)sqrt(cc > 0.0 ? cc : 0.0));
        } else {
            prim_store<NT>(&a.rho[I], (W)(1)); prim_store<NT>(&a.u[I], (W)(0)); prim_store<NT>(&a.v[I], (W)(0)); prim_store<NT>(&a.w[I], (W)(0));
            prim_store<NT>(&a.p[I], (W)(1)); prim_store<NT>(&a.T[I], (W)(1)); prim_store<NT>(&a.c[I], (W)(1));
        }
        return;
    }
    if (!IG && a.n9 && rho > 0) {
        // physics.jl:515-552 _primitives! (Nasa9Mixture)
        const N9Table &t = *a.n9;
        const double ri = 1 / rho;
        double Rm = 0.0, Y[CLES_MAXSP];
        for (int sp = 0; sp < Ns; sp++) {
            const double q = a.Q[(int64_t)sp * N + I];
            Y[sp] = q * ri;
            prim_store<NT>(&a.Y[sp][I], (W)Y[sp]);
            Rm += q * ri * t.R[sp];
        }
        const double u = a.Q[(int64_t)Ns * N + I] * ri;
        const double v = a.Q[(int64_t)(Ns + 1) * N + I] * ri;
        const double w = a.Q[(int64_t)(Ns + 2) * N + I] * ri;
        const double e = a.Q[(int64_t)(Ns + 3) * N + I] * ri - 0.5 * (u * u + v * v + w * w);
        const double T = n9_temperature(t, e, Y);
        double cpm = 0.0;
        for (int sp = 0; sp < Ns; sp++) cpm += Y[sp] * n9_cp(t, sp, T);
        const double cvm = cpm - Rm;
        if (a.esens) {
            const double r = d_max(rho, 1e-300);
            const double m1 = a.Q[(int64_t)Ns * N + I], m2 = a.Q[(int64_t)(Ns + 1) * N + I],
                         m3 = a.Q[(int64_t)(Ns + 2) * N + I];
            const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r;
            prim_store<NT>(&a.esens[I], (W)((a.Q[(int64_t)(Ns + 3) * N + I] - ke) / r));
        }
        const double cc = (cpm / cvm) * Rm * T;
        prim_store<NT>(&a.rho[I], (W)rho);
        prim_store<NT>(&a.u[I], (W)u);
        prim_store<NT>(&a.v[I], (W)v);
        prim_store<NT>(&a.w[I], (W)w);
        prim_store<NT>(&a.p[I], (W)(rho * Rm * T));
        prim_store<NT>(&a.T[I], (W)T);
        prim_store<NT>(&a.c[I], (W)sqrt(cc > 0.0 ? cc : 0.0));
        prim_store<NT>(&a.cp[I], (W)cpm);
    } else if (rho > 0) {
        const double ri = 1 / rho;
        double Rm = 0.0, cvm = 0.0;
#pragma unroll
        for (int sp = 0; sp < CLES_MAXSP; sp++)
            if (sp < Ns) {
                const double q = a.Q[(int64_t)sp * N + I];
                Rm += q * a.Rk[sp];
                cvm += q * a.cvk[sp];
                if (a.Y[sp]) prim_store<NT>(&a.Y[sp][I], (W)(q * ri));
            }
        Rm *= ri;
        cvm *= ri;
        const double u = a.Q[(int64_t)Ns * N + I] * ri;
        const double v = a.Q[(int64_t)(Ns + 1) * N + I] * ri;
        const double w = a.Q[(int64_t)(Ns + 2) * N + I] * ri;
        const double e = a.Q[(int64_t)(Ns + 3) * N + I] * ri - 0.5 * (u * u + v * v + w * w);
        const double T = d_max(e / cvm, 1e-300);
        const double p = rho * Rm * T;
        const double gm = 1 + Rm / cvm;
        if (a.esens) {
            // artificial.jl:321-324 with rho = max(rho, 1e-300)
            const double r = d_max(rho, 1e-300);
            const double m1 = a.Q[(int64_t)Ns * N + I], m2 = a.Q[(int64_t)(Ns + 1) * N + I],
                         m3 = a.Q[(int64_t)(Ns + 2) * N + I];
            const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r;
            prim_store<NT>(&a.esens[I], (W)((a.Q[(int64_t)(Ns + 3) * N + I] - ke) / r));
        }
        prim_store<NT>(&a.rho[I], (W)rho);
        prim_store<NT>(&a.u[I], (W)u);
        prim_store<NT>(&a.v[I], (W)v);
        prim_store<NT>(&a.w[I], (W)w);
        prim_store<NT>(&a.p[I], (W)p);
        prim_store<NT>(&a.T[I], (W)T);
        prim_store<NT>(&a.c[I], (W)sqrt(gm * Rm * T));
        if (a.cp) prim_store<NT>(&a.cp[I], (W)(cvm + Rm));
    } else {
        if (a.esens) {
            const double m1 = a.Q[(int64_t)Ns * N + I], m2 = a.Q[(int64_t)(Ns + 1) * N + I],
                         m3 = a.Q[(int64_t)(Ns + 2) * N + I];
            const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / 1.0;
            prim_store<NT>(&a.esens[I], (W)((a.Q[(int64_t)(Ns + 3) * N + I] - ke) / 1.0));
        }
        prim_store<NT>(&a.rho[I], (W)(1)); prim_store<NT>(&a.u[I], (W)(0)); prim_store<NT>(&a.v[I], (W)(0)); prim_store<NT>(&a.w[I], (W)(0));
        prim_store<NT>(&a.p[I], (W)(1)); prim_store<NT>(&a.T[I], (W)(1)); prim_store<NT>(&a.c[I], (W)(1));
        if (a.cp) prim_store<NT>(&a.cp[I], (W)(1));
#pragma unroll
        for (int sp = 0; sp < CLES_MAXSP; sp++)
            if (sp < Ns && a.Y[sp]) prim_store<NT>(&a.Y[sp][I], (W)(sp == 0 ? W(1) : W(0)));
    }
}

// ---------------------------------------------------------------------------
// artificial.jl:279-290 strain-rate magnitude |S|
// ---------------------------------------------------------------------------
template <typename W>
__global__ void __launch_bounds__(TPB) k_strain(GradPtrs<W> g, W *__restrict__ out, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    W G[9];
#pragma unroll
    for (int m = 0; m < 9; m++) G[m] = g.gu[m] ? g.gu[m][I] : W(0);
    W ss = 0;
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a2 = 0; a2 < 3; a2++) {
            const W S = W(0.5) * (G[a2 * 3 + b] + G[b * 3 + a2]);
            ss += S * S;
        }
    out[I] = sqrt(ss);
}

// ---------------------------------------------------------------------------
// artificial.jl:67-94 delta4_sum!: sum over active d of h_d^wpow |delta4_d f|,
// indices clamped at closed edges, wrapped on periodic dimensions.
// MODE 0: f is a field.  MODE 1: f = e = (E - ke)/rho from Q and rho
// (artificial.jl:319-326) evaluated at each stencil point.
// ---------------------------------------------------------------------------
template <typename W, int MODE>
__device__ __forceinline__ W d4_value(const D4Args<W> &a, int64_t J)
{
    if constexpr (MODE == 0) {
        return a.f[J];
    } else {
        const int64_t N = a.N;
        const double rho = d_max((double)a.rho[J], 1e-300);
        const double m1 = a.Q[(int64_t)a.im1 * N + J], m2 = a.Q[(int64_t)(a.im1 + 1) * N + J],
                     m3 = a.Q[(int64_t)(a.im1 + 2) * N + J];
        const double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
        return (W)((a.Q[(int64_t)(a.im1 + 3) * N + J] - ke) / rho);
    }
}

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2;
// renumbering them so that every XCD sweeps one contiguous slab keeps the
// +-2 rows and planes of the stencil in that XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned nfull = nb & ~7u;
    return b < nfull ? (b & 7u) * (nfull >> 3) + (b >> 3) : b;
}

template <typename W, int MODE>
__global__ void __launch_bounds__(TPB) k_delta4(D4Args<W> a, W *__restrict__ out)
{
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t I = (int64_t)blk * TPB + threadIdx.x;
    if (I >= a.N) return;
    const int nx = (int)a.n[0], ny = (int)a.n[1];
    const int I32 = (int)I;
    const int ix = I32 % nx, r = I32 / nx;
    const int ii[3] = {ix, r % ny, r / ny};
    const int st[3] = {1, nx, nx * ny};
    W acc_out = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!a.active[d]) continue;
        const int nd = (int)a.n[d], il = ii[d];
        double acc = 0.0;
        const double D4[5] = {1.0, -4.0, 6.0, -4.0, 1.0};
#pragma unroll
        for (int m = -2; m <= 2; m++) {
            int j = il + m;
            if (a.periodic[d]) {
                if (j < 0) j += nd;
                else if (j >= nd) j -= nd;
            } else {
                j = j < 0 ? 0 : (j > nd - 1 ? nd - 1 : j);
            }
            acc += D4[m + 2] * (double)d4_value<W, MODE>(a, I32 + (j - il) * st[d]);
        }
        acc_out += (W)(a.wd[d] * fabs(acc));
    }
    out[I] = acc_out;
}

__device__ __forceinline__ int d4_index(int v, int n, int periodic)
{
    if (periodic) return v < 0 ? v + n : (v >= n ? v - n : v);
    return v < 0 ? 0 : (v > n - 1 ? n - 1 : v);
}

// The same sum, 2.5-D: a block owns a 64 x 16 tile of (x, y) columns and
// marches zl planes in z (at most D4T_Z; fewer on small grids, so that the
// launch still has a few blocks per CU).  Each plane's tile with its two-wide
// halo sits in LDS (every slot holds the value at the wrapped or clamped
// source index, so edges and collapsed dimensions need no special case), the
// x and y stencils read LDS, and each thread keeps a five-plane z window in
// registers for its four points (with z inactive, the planes are independent
// problems side by side and the window's centre is the current plane).  HBM
// reads are the tile plus its halo once per plane; the values of the next
// plane are loaded while the current one is computed.
//
// Every value is widened to double once (into LDS, or into the z window as it
// arrives) and each stencil is v0 - 4 v1 + 6 v2 - 4 v3 + v4 summed left to
// right.  The products by -4 are exact (and by 6 for float inputs), so they
// go through fma, which then rounds exactly as the separate product and sum
// of the pointwise kernel do: the two kernels agree bitwise.
constexpr int D4T_X = 64, D4T_Y = 16, D4T_Z = 32, D4T_LX = D4T_X + 4, D4T_LY = D4T_Y + 4;

constexpr int D4T_HALO = D4T_LX * D4T_LY - D4T_X * D4T_Y;   // 336 halo slots per plane

__device__ __forceinline__ int d4_src(int v, int n, int periodic, bool active)
{
    return active ? d4_index(v, n, periodic) : 0;
}
// a tile slot at source coordinate v is read by some point of the grid only
// when v lies in [-2, n + 1] (collapsed: v = 0); other slots are zero-filled
__device__ __forceinline__ bool d4_needed(int v, int n, bool active)
{
    return active ? (v >= -2 && v <= n + 1) : v == 0;
}

// acc + c v with c v exact: fma rounds once, as the sum alone would; the
// inexact 6 v (double inputs) keeps its own rounding (no contraction)
__device__ __forceinline__ double d4_stencil(double v0, double v1, double v2, double v3, double v4, bool six_exact)
{
#pragma clang fp contract(off)
    double acc = v0;
    acc = fma(-4.0, v1, acc);
    acc = six_exact ? fma(6.0, v2, acc) : acc + 6.0 * v2;
    acc = fma(-4.0, v3, acc);
    return acc + v4;
}

// FP64: four waves per SIMD asked for (126 VGPRs, no spills; the default
// allocation takes 132 and fits only three)
template <typename W>
__global__ void __launch_bounds__(TPB, sizeof(W) == 8 ? 4 : 1) k_delta4_tile(D4Args<W> a, D4Multi<W> fm, int gx, int gy, int zl)
{
    constexpr bool SIX = sizeof(W) == 4;   // 6 v exact in double for float v
    const int fi = (int)blockIdx.y;
    W *__restrict__ out = fm.out[fi];
    double wd[3];
#pragma unroll
    for (int k = 0; k < 3; k++) wd[k] = fm.wpow[fi] == 2 ? a.wd[k] * a.wd[k] : a.wd[k];
    __shared__ double lds[D4T_LY * D4T_LX];   // input values, widened exactly
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int bx = (int)(blk % (unsigned)gx);
    const unsigned r = blk / (unsigned)gx;
    const int by = (int)(r % (unsigned)gy), bz = (int)(r / (unsigned)gy);
    const int nx = (int)a.n[0], ny = (int)a.n[1], nz = (int)a.n[2];
    const int sy = nx;
    const int64_t sz = (int64_t)nx * ny;
    const bool xa = a.active[0], ya = a.active[1], za = a.active[2];
    const int px = a.periodic[0], py = a.periodic[1], pz = a.periodic[2];
    const int tx = (int)(threadIdx.x & 63), ty = (int)(threadIdx.x >> 6);
    const int i0 = bx * D4T_X, j0 = by * D4T_Y, k0 = bz * zl;
    const int i = i0 + tx;
    const W *f = fm.f[fi];
    // Every global load below is unconditional, from an in-bounds address
    // (offset 0 of the plane for a slot nothing reads), and its value selected
    // afterwards: with no load under a branch, the waits count loads instead
    // of draining them all, so the loads of the next planes stay in flight
    // through the current plane's LDS fill and stencils.
    //
    // own points: rows j0 + 4 ty + q, q = 0..3 (four consecutive rows, so
    // their y stencils share the thread's own values and read only the two
    // rows above and below from LDS), with the z window k-2..k+2
    // widened in wz; plane k+3 is loaded during plane k into wn (stored type)
    // and widened after plane k's stencils
    double wz[4][5];
    W wn[4];
    bool own[4];
    int ooff[4];
    // LDS slots this thread fills on every plane: slots 0..3 are its own rows
    // (from the register window when the point is in the grid), slots 4..5 up
    // to two halo slots; src is the in-plane source offset, -1 for a slot that
    // no point reads (zero), and slot -1 marks a halo slot past the 336th
    int slot[6];
    int src[6];
    // z planes are wrapped with a true modulo: the window reaches 3 planes out
    auto zplane = [&](int v) { return za ? (pz ? ((v % nz) + nz) % nz : d4_index(v, nz, 0)) : d4_index(v, nz, 0); };
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int j = j0 + 4 * ty + q;
        own[q] = i < nx && j < ny;
        ooff[q] = own[q] ? j * sy + i : 0;
        slot[q] = (4 * ty + q + 2) * D4T_LX + tx + 2;
        src[q] = -1;
        if (!own[q] && d4_needed(i, nx, xa) && d4_needed(j, ny, ya))
            src[q] = d4_src(j, ny, py, ya) * sy + d4_src(i, nx, px, xa);
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int h = (int)threadIdx.x + TPB * e;
        slot[4 + e] = -1;
        src[4 + e] = -1;
        if (h < D4T_HALO) {
            int xl, yl;
            if (h < 4 * D4T_LX) {   // rows -2, -1, 16, 17 across the full width
                const int rr = h / D4T_LX;
                xl = h - rr * D4T_LX;
                yl = rr < 2 ? rr : D4T_Y + rr;
            } else {                // columns -2, -1, 64, 65 of the 16 inner rows
                const int t = h - 4 * D4T_LX;
                const int c = t / D4T_Y;
                yl = 2 + (t - c * D4T_Y);
                xl = c < 2 ? c : D4T_X + c;
            }
            slot[4 + e] = yl * D4T_LX + xl;
            if (d4_needed(i0 + xl - 2, nx, xa) && d4_needed(j0 + yl - 2, ny, ya))
                src[4 + e] = d4_src(j0 + yl - 2, ny, py, ya) * sy + d4_src(i0 + xl - 2, nx, px, xa);
        }
    }
    int soff[6];
#pragma unroll
    for (int e = 0; e < 6; e++) soff[e] = src[e] >= 0 ? src[e] : 0;
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const W *pl = f + (int64_t)zplane(k0 - 2 + m) * sz;
#pragma unroll
        for (int q = 0; q < 4; q++) wz[q][m] = (double)pl[ooff[q]];
    }
    // the slot values of plane k are loaded one plane ahead
    W cur[6], nxt[6];
    const int kend = k0 + zl < nz ? k0 + zl : nz;
    {
        const W *pl = f + (int64_t)k0 * sz;
#pragma unroll
        for (int e = 0; e < 6; e++) cur[e] = pl[soff[e]];
    }
    for (int k = k0; k < kend; k++) {
        {
            const W *pl = f + (int64_t)zplane(k + 3) * sz;
#pragma unroll
            for (int q = 0; q < 4; q++) wn[q] = pl[ooff[q]];
        }
        {
            const W *pl = f + (int64_t)(k + 1 < kend ? k + 1 : k) * sz;
#pragma unroll
            for (int e = 0; e < 6; e++) nxt[e] = pl[soff[e]];
        }
        // the values of the thread's own slots (as written to LDS)
        double rv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) rv[q] = own[q] ? wz[q][2] : (src[q] >= 0 ? (double)cur[q] : 0.0);
        __syncthreads();   // the previous plane's stencil reads are done
#pragma unroll
        for (int q = 0; q < 4; q++) lds[slot[q]] = rv[q];
#pragma unroll
        for (int e = 4; e < 6; e++)
            if (slot[e] >= 0) lds[slot[e]] = src[e] >= 0 ? (double)cur[e] : 0.0;
        __syncthreads();
        // column of rows 4 ty - 2 .. 4 ty + 5 at this thread's x: two rows
        // above and below from LDS, the four own rows from registers
        const int xl = tx + 2, yb = 4 * ty + 2;
        double yv[8];
        yv[0] = lds[(yb - 2) * D4T_LX + xl];
        yv[1] = lds[(yb - 1) * D4T_LX + xl];
#pragma unroll
        for (int q = 0; q < 4; q++) yv[2 + q] = rv[q];
        yv[6] = lds[(yb + 4) * D4T_LX + xl];
        yv[7] = lds[(yb + 5) * D4T_LX + xl];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (own[q]) {
                const int yl = yb + q;
                W acc_out = 0;
                if (xa) {
                    const double *c = lds + yl * D4T_LX + xl;
                    acc_out += (W)(wd[0] * fabs(d4_stencil(c[-2], c[-1], rv[q], c[1], c[2], SIX)));
                }
                if (ya)
                    acc_out += (W)(wd[1] * fabs(d4_stencil(yv[q], yv[q + 1], yv[q + 2], yv[q + 3], yv[q + 4], SIX)));
                if (za)
                    acc_out += (W)(wd[2] * fabs(d4_stencil(wz[q][0], wz[q][1], wz[q][2], wz[q][3], wz[q][4], SIX)));
                out[(int64_t)k * sz + ooff[q]] = acc_out;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
#pragma unroll
            for (int m = 0; m < 4; m++) wz[q][m] = wz[q][m + 1];
            wz[q][4] = (double)wn[q];
        }
#pragma unroll
        for (int e = 0; e < 6; e++) cur[e] = nxt[e];
    }
}

// ---------------------------------------------------------------------------
// artificial.jl:136-145 compression switch, 171-183 gate_beta!, 216-242 dilatation
// ---------------------------------------------------------------------------
template <typename W>
__device__ __forceinline__ W compression_switch(const GradPtrs<W> &g, int64_t I)
{
    auto G = [&](int d, int j) -> W { return g.gu[d * 3 + j] ? g.gu[d * 3 + j][I] : W(0); };
    const W div = G(0, 0) + G(1, 1) + G(2, 2);
    if (!(div < 0)) return W(0);
    const W w1 = G(1, 2) - G(2, 1);
    const W w2 = G(2, 0) - G(0, 2);
    const W w3 = G(0, 1) - G(1, 0);
    return div * div / (div * div + w1 * w1 + w2 * w2 + w3 * w3 + (W)1e-32);
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_gate_beta(GradPtrs<W> g, W *__restrict__ beta, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    beta[I] *= compression_switch(g, I);
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_dilatation_prep(GradPtrs<W> g, W *__restrict__ div, W *__restrict__ sw,
                                                         int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    auto G = [&](int d) -> W { return g.gu[d * 4] ? g.gu[d * 4][I] : W(0); };
    div[I] = G(0) + G(1) + G(2);
    sw[I] = compression_switch(g, I);
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_dilatation_beta(const W *__restrict__ rho, const W *__restrict__ sensor,
                                                         const W *__restrict__ sw, W *__restrict__ beta, double C_beta,
                                                         int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    const W s = sensor[I];
    beta[I] = (W)C_beta * rho[I] * (s > W(0) ? s : W(0)) * sw[I];
}

// ---------------------------------------------------------------------------
// rhs.jl:495-556 _assemble_fluxes! for the active dimensions
// ---------------------------------------------------------------------------
// The flux fields are read again by the divergence sweeps after the whole
// launch; when they exceed the 256 MB Infinity Cache (15 fields, 2 GB at
// 256^3 FP64) they are written with the non-temporal hint, and the velocity
// and temperature gradients, dead after this kernel, are read with it.
template <bool NT, typename W>
__device__ __forceinline__ W dead_load(const W *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename W>
__device__ __forceinline__ void flux_store(W *p, W v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// MS: the species fluxes and enthalpy diffusion are compiled in (used when
// Ns > 1); the instantiation without them takes half the registers.  SPM:
// the bound of the species loops (CLES_MAXSP, or 2 for two-species mixtures:
// the same operations on fewer registers).  IG: no NASA-9 tables.
template <typename W, bool NT, bool MS, int SPM, bool IG = false>
__device__ __forceinline__ void fluxes_point(const FluxArgs<W> &a, int64_t N, int64_t I)
{
    const int Ns = a.Ns;
    const W rho = a.rho[I];
    const W uv[3] = {a.u[I], a.v[I], a.w[I]};
    const W p = a.p[I], Tp = a.T[I];
    const W E = (W)a.Q[(int64_t)(Ns + 3) * N + I];
    const W mu0 = (W)a.mu0;
    const W mu = mu0 + (a.mu_art ? a.mu_art[I] : W(0));
    const W beta = a.beta_art ? a.beta_art[I] : W(0);
    // cp and Y_1 are NULL together (one-species ideal gas: formed from rho)
    W cpm, Y1c = W(1);
    if (a.cp) {
        cpm = a.cp[I];
    } else {
        double c, y;
        one_species_cp_y((double)rho, a.R1, a.cv1, c, y);
        cpm = (W)c;
        Y1c = (W)y;
    }
    const W kappa = mu0 * cpm / (W)a.Pr + (a.kappa_art ? a.kappa_art[I] : W(0));
    const W D0 = mu0 / ((W)a.Sc * rho);
    W G[9];
#pragma unroll
    for (int m = 0; m < 9; m++) G[m] = a.g.gu[m] ? dead_load<NT>(a.g.gu[m] + I) : W(0);
    const W divu = G[0] + G[4] + G[8];
    const W t23 = W(2.0 / 3.0);
    const W t11 = mu * (2 * G[0] - t23 * divu) + beta * divu;
    const W t22 = mu * (2 * G[4] - t23 * divu) + beta * divu;
    const W t33 = mu * (2 * G[8] - t23 * divu) + beta * divu;
    const W t12 = mu * (G[1] + G[3]);
    const W t13 = mu * (G[2] + G[6]);
    const W t23b = mu * (G[5] + G[7]);
    const W tau[3][3] = {{t11, t12, t13}, {t12, t22, t23b}, {t13, t23b, t33}};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!a.active[d]) continue;
        const W ud = uv[d];
        W *const *F = a.flux[d];
        W hdiff = 0;
        if (MS && Ns > 1) {
            // rhs.jl:544 species_enthalpy: cp_k T (IdealMixture) or h_k(T) (Nasa9Mixture)
            W hk[SPM];
#pragma unroll
            for (int sp = 0; sp < SPM; sp++)
                if (sp < Ns) hk[sp] = (!IG && a.n9) ? (W)n9_h(*a.n9, sp, (double)Tp) : (W)a.cpk[sp] * Tp;
            W Vc = 0;
#pragma unroll
            for (int sp = 0; sp < SPM; sp++)
                if (sp < Ns) Vc += (D0 + a.D_art[sp][I]) * a.gY[d][sp][I];
#pragma unroll
            for (int sp = 0; sp < SPM; sp++)
                if (sp < Ns) {
                    const W Dk = D0 + a.D_art[sp][I];
                    const W J = rho * (-Dk * a.gY[d][sp][I] + a.Y[sp][I] * Vc);
                    flux_store<NT>(F[sp] + I, rho * a.Y[sp][I] * ud + J);
                    hdiff += hk[sp] * J;
                }
        } else {
            // one species: sum_k J_k = 0 forces J_1 = 0 (the reference's
            // rho*D0*dY*(Y-1) with Y == 1 to rounding)
            flux_store<NT>(F[0] + I, rho * (a.Y[0] ? a.Y[0][I] : Y1c) * ud);
        }
        flux_store<NT>(F[Ns] + I, rho * ud * uv[0] + (d == 0 ? p : W(0)) - tau[d][0]);
        flux_store<NT>(F[Ns + 1] + I, rho * ud * uv[1] + (d == 1 ? p : W(0)) - tau[d][1]);
        flux_store<NT>(F[Ns + 2] + I, rho * ud * uv[2] + (d == 2 ? p : W(0)) - tau[d][2]);
        flux_store<NT>(F[Ns + 3] + I, (E + p) * ud - (uv[0] * tau[d][0] + uv[1] * tau[d][1] + uv[2] * tau[d][2]) -
                       kappa * dead_load<NT>(a.gT[d] + I) + hdiff);
    }
}

template <typename W, bool NT, bool MS>
__global__ void __launch_bounds__(TPB) k_fluxes(FluxArgs<W> a, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    fluxes_point<W, NT, MS, CLES_MAXSP>(a, N, I);
}

// two-species ideal-gas mixtures
template <typename W, bool NT>
__global__ void __launch_bounds__(TPB) k_fluxes_two(FluxArgs<W> a, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    fluxes_point<W, NT, true, 2, true>(a, N, I);
}

// timestep.jl:85-94 (standalone form, used when the RHS needs boundary corrections)
template <typename W>
__global__ void __launch_bounds__(TPB) k_lsrk(double *__restrict__ Q, double *__restrict__ du,
                                              const W *__restrict__ dQ, double A, double B, double dt,
                                              const double *__restrict__ dtp, const int *__restrict__ gate, int64_t NC,
                                              int64_t N, unsigned ipts)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= NC) return;
    // ipts > 0: instances of ipts points side by side, each with its own gate and dt
    const unsigned ii = ipts ? (unsigned)(I % N) / ipts : 0u;
    if (gate && !gate[ii]) return;
    if (dtp) dt = dtp[ii];
    const double v = fma(dt, (double)dQ[I], A * du[I]);
    du[I] = v;
    Q[I] += B * v;
}

// ---------------------------------------------------------------------------
// timestep.jl:125-175 max_rate: per-block max of the CFL rate and min of the
// mixture density, then a single-block final reduction (order-independent).
// ---------------------------------------------------------------------------
// timestep.jl:146-170: the rate of point I (the mixture density goes to rQ)
__device__ __forceinline__ void interior_ijk(const int64_t (&n)[3], int64_t I, int64_t (&ii)[3])
{
    ii[0] = I % n[0];
    ii[1] = (I / n[0]) % n[1];
    ii[2] = I / (n[0] * n[1]);
}
// metric.jl:114-120: the clamped scale factors at interior node ii
__device__ __forceinline__ void metric_hs(const Metric1D &g, const int64_t (&ii)[3], double (&h)[3])
{
    double s2 = 1.0, s3 = 1.0;
    if (g.kind != CLES_METRIC_CARTESIAN) {
        const double r = g.r[ii[0]];
        s2 = r;
        if (g.kind == CLES_METRIC_SPHERICAL) s3 = r * g.sth[ii[1]];
    }
    h[0] = fmax(1.0 * (g.m[0] ? g.m[0][ii[0]] : 1.0), 1e-300);
    h[1] = fmax(s2 * (g.m[1] ? g.m[1][ii[1]] : 1.0), 1e-300);
    h[2] = fmax(s3 * (g.m[2] ? g.m[2][ii[2]] : 1.0), 1e-300);
}
__device__ __forceinline__ double metric_h(const Metric1D &g, int d, const int64_t (&ii)[3])
{
    double h[3];
    metric_hs(g, ii, h);
    return h[d];
}
__device__ __forceinline__ double metric_J(const Metric1D &g, const int64_t (&ii)[3])
{
    double h[3];
    metric_hs(g, ii, h);
    return h[0] * h[1] * h[2];
}
// metric.jl:129-133: 1/r and cot(theta)/r
__device__ __forceinline__ double metric_inv_r(const Metric1D &g, const int64_t (&ii)[3])
{
    return 1 / fmax(g.r[ii[0]], 1e-300);
}
__device__ __forceinline__ double metric_cotr(const Metric1D &g, const int64_t (&ii)[3], int64_t I)
{
    return g.cotr ? g.cotr[I] : g.cth[ii[1]] / fmax(g.r[ii[0]] * g.sth[ii[1]], 1e-300);
}
__device__ __forceinline__ bool metric_on(const Metric1D &g, int d)
{
    return g.kind != CLES_METRIC_CARTESIAN || g.m[d];
}

template <typename W, int MODE>
__global__ void __launch_bounds__(TPB) k_metric_scale(MetricFields<W> fs, Metric1D g, int dim, int64_t N)
{
    for (int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x; I < N; I += (int64_t)gridDim.x * TPB) {
        int64_t ii[3];
        interior_ijk(g.n, I, ii);
        double h[3];
        metric_hs(g, ii, h);
        double sc;
        if constexpr (MODE == 0) {
            sc = 1 / h[dim];
        } else {
            const double J = h[0] * h[1] * h[2];
            sc = MODE == 2 ? 1 / J : J / h[dim];
        }
        if constexpr (MODE == 3)
            for (int f = 0; f < fs.nf; f++) fs.f[f][I] = (W)sc;
        else
            for (int f = 0; f < fs.nf; f++) fs.f[f][I] = (W)((double)fs.f[f][I] * sc);
    }
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_body_force(const double *__restrict__ Q, W *__restrict__ dQ, int Ns,
                                                   int64_t N, double g1, double g2, double g3)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    double rho = 0.0;
    for (int sp = 0; sp < Ns; sp++) rho += Q[(int64_t)sp * N + I];
    const double m1 = Q[(int64_t)Ns * N + I], m2 = Q[(int64_t)(Ns + 1) * N + I], m3 = Q[(int64_t)(Ns + 2) * N + I];
    W *d = dQ + (int64_t)Ns * N + I;
    d[0] = (W)((double)d[0] + rho * g1);
    d[N] = (W)((double)d[N] + rho * g2);
    d[2 * N] = (W)((double)d[2 * N] + rho * g3);
    d[3 * N] = (W)((double)d[3 * N] + (m1 * g1 + m2 * g2 + m3 * g3));
}

// folds.jl:118-151 _pair_index on one rank
__device__ __forceinline__ int64_t pair_partner(const PairMap &pm, int64_t (&ii)[3])
{
    int64_t m[3] = {ii[0], ii[1], ii[2]};
    if (pm.pdim >= 0) {
        const int64_t half = pm.n[pm.pdim] / 2;
        m[pm.pdim] = ii[pm.pdim] < half ? ii[pm.pdim] + half : ii[pm.pdim] - half;
    }
    if (pm.revdim >= 0) m[pm.revdim] = pm.n[pm.revdim] - 1 - ii[pm.revdim];
    return m[0] + pm.n[0] * (m[1] + pm.n[1] * m[2]);
}

// folds.jl:160-196 pair_forward!: e = (f(x) + sf f(Mx)) / 2 at the canonical
// slot, o = (f(x) - sf f(Mx)) / 2 at its partner
template <typename T>
__global__ void __launch_bounds__(TPB) k_pair_forward(const T *f, T *w, PairMap pm, double sf, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    int64_t ii[3];
    interior_ijk(pm.n, I, ii);
    if (ii[pm.sd] >= pm.n[pm.sd] / 2) return;
    const int64_t Mx = pair_partner(pm, ii);
    const double a = f[I], b = sf * (double)f[Mx];
    w[I] = (T)(0.5 * (a + b));
    w[Mx] = (T)(0.5 * (a - b));
}

// folds.jl:206-246 pair_backward! with the result's epilogue
template <typename T>
__global__ void __launch_bounds__(TPB) k_pair_backward(const T *re, const T *ro, T *dst, PairMap pm, double sf,
                                                       int epi, const int *gate, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= N) return;
    if (epi == EPI_STORE_GATE && gate && !*gate) return;
    int64_t ii[3];
    interior_ijk(pm.n, I, ii);
    if (ii[pm.sd] >= pm.n[pm.sd] / 2) return;
    const int64_t Mx = pair_partner(pm, ii);
    const double Re = re[I], Ro = ro[Mx];
    const double v[2] = {Re + Ro, sf * (Re - Ro)};
    const int64_t at[2] = {I, Mx};
    for (int t = 0; t < 2; t++) {
        T *q = dst + at[t];
        if (epi == EPI_STORE_NEG) *q = (T)(-v[t]);
        else if (epi == EPI_SUB) *q = (T)((double)*q - v[t]);
        else *q = (T)v[t];
    }
}

// metric.jl:181-219 metric_correct_gradients! (grad_u[d][j] at gu[d*3 + j])
template <typename W>
__global__ void __launch_bounds__(TPB) k_metric_grad(CurvArgs<W> a)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= a.N) return;
    int64_t ii[3];
    interior_ijk(a.g.n, I, ii);
    const double ir = metric_inv_r(a.g, ii);
    const double u = a.u[I], v = a.v[I], w = a.w[I];
    auto G = [&](int k) -> double { return a.zero[k / 3] ? 0.0 : (double)a.gu[k][I]; };
    if (a.g.kind == CLES_METRIC_CYLINDRICAL) {
        a.gu[3][I] = (W)(G(3) - v * ir);
        a.gu[4][I] = (W)(G(4) + u * ir);
        if (a.zero[1]) a.gu[5][I] = (W)0;
    } else {
        const double ctr = metric_cotr(a.g, ii, I);
        a.gu[3][I] = (W)(G(3) - v * ir);
        a.gu[4][I] = (W)(G(4) + u * ir);
        if (a.zero[1]) a.gu[5][I] = (W)0;
        a.gu[6][I] = (W)(G(6) - w * ir);
        a.gu[7][I] = (W)(G(7) - w * ctr);
        a.gu[8][I] = (W)(G(8) + (u * ir + v * ctr));
    }
}

// metric.jl:228-269 add_metric_sources! with _Pi (a, b zero-based)
template <typename W>
__global__ void __launch_bounds__(TPB) k_metric_sources(CurvArgs<W> a)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I >= a.N) return;
    int64_t ii[3];
    interior_ijk(a.g.n, I, ii);
    auto G = [&](int k) -> double { return a.gu[k] ? (double)a.gu[k][I] : 0.0; };
    const double mu = a.mu0 + (a.mu_art ? (double)a.mu_art[I] : 0.0);
    const double be = a.beta_art ? (double)a.beta_art[I] : 0.0;
    const double divu = G(0) + G(4) + G(8);
    const double rho = a.rho[I], p = a.p[I];
    const double uv[3] = {(double)a.u[I], (double)a.v[I], (double)a.w[I]};
    auto Pi = [&](int x, int y) -> double {
        const double tau = mu * (G(x * 3 + y) + G(y * 3 + x)) + (x == y ? (be - 2 * mu / 3) * divu : 0.0);
        return rho * uv[x] * uv[y] + (x == y ? p : 0.0) - tau;
    };
    const double ir = metric_inv_r(a.g, ii);
    W *m1 = a.dQ + (int64_t)a.im1 * a.N + I, *m2 = m1 + a.N, *m3 = m2 + a.N;
    if (a.g.kind == CLES_METRIC_CYLINDRICAL) {
        *m1 = (W)((double)*m1 + ir * Pi(1, 1));
        *m2 = (W)((double)*m2 - ir * Pi(1, 0));
    } else {
        const double ctr = metric_cotr(a.g, ii, I);
        const double Ptt = Pi(1, 1), Ppp = Pi(2, 2);
        *m1 = (W)((double)*m1 + ir * (Ptt + Ppp));
        *m2 = (W)((double)*m2 + (-ir * Pi(1, 0) + ctr * Ppp));
        *m3 = (W)((double)*m3 + (-ir * Pi(2, 0) - ctr * Pi(2, 1)));
    }
}

// CART: a Cartesian grid without stretching (the metric code compiled out)
template <typename W, bool CART = false>
__device__ __forceinline__ double rate_at(const RateArgs<W> &a, int64_t I, int64_t N, double &rQ)
{
    rQ = 0.0;
    for (int sp = 0; sp < a.Ns; sp++) rQ += a.Q[(int64_t)sp * N + I];
    const double rho = a.rho[I];
    const double ri = 1 / rho;
    const double c = a.c[I];
    double cp, y;
    if (CART || a.cp) {   // the Cartesian (batch) form always has cp stored
        cp = a.cp[I];
    } else {
        one_species_cp_y(rho, a.R1, a.cv1, cp, y);
        cp = (double)(W)cp;
    }
    const double uv[3] = {(double)a.u[I], (double)a.v[I], (double)a.w[I]};
    double acc = 0.0, dsum = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!a.active[d]) continue;
        double idx = a.invh[d];
        if (!CART && metric_on(a.met, d)) {   // timestep.jl:151 inv_h[d][I] / h[d]
            int64_t ii[3];
            interior_ijk(a.met.n, I, ii);
            idx = (1 / metric_h(a.met, d, ii)) / a.h[d];
        }
        acc += (fabs(uv[d]) + c) * idx;
        dsum += idx * idx;
    }
    if (!CART && a.met.kind != CLES_METRIC_CARTESIAN) {   // timestep.jl:186-197 curvature_rate
        int64_t ii[3];
        interior_ijk(a.met.n, I, ii);
        const double ir = metric_inv_r(a.met, ii);
        if (a.met.kind == CLES_METRIC_CYLINDRICAL) {
            if (!a.active[1]) acc += fabs(uv[1]) * ir;
        } else {
            double cr = 0.0;
            if (!a.active[1]) cr += fabs(uv[1]) * ir;
            if (!a.active[2]) cr += fabs(uv[2]) * (ir + fabs(metric_cotr(a.met, ii, I)));
            acc += cr;
        }
    }
    double Dmax = a.mu0 / (a.Sc * rho);
    for (int sp = 0; sp < a.Ns; sp++)
        if (a.D_art[sp]) Dmax = d_max(Dmax, a.mu0 / (a.Sc * rho) + (double)a.D_art[sp][I]);
    const double mua = a.mu_art ? (double)a.mu_art[I] : 0.0;
    const double bea = a.beta_art ? (double)a.beta_art[I] : 0.0;
    const double kaa = a.kappa_art ? (double)a.kappa_art[I] : 0.0;
    const double nu = (a.mu0 + mua + bea) * ri + (a.mu0 * cp / a.Pr + kaa) * ri / cp + Dmax;
    acc += 2 * nu * dsum;
    return acc;
}

// NaN-propagating block max / min of (rate, rmin) over TPB threads
__device__ __forceinline__ void block_maxmin(double *smax, double *smin, double rate, double rmin)
{
    smax[threadIdx.x] = rate;
    smin[threadIdx.x] = rmin;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double x = smax[threadIdx.x + s], y = smin[threadIdx.x + s];
            double X = smax[threadIdx.x], Y = smin[threadIdx.x];
            smax[threadIdx.x] = (x > X || x != x) ? x : X;
            smin[threadIdx.x] = (y < Y || y != y) ? y : Y;
        }
        __syncthreads();
    }
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_rate(RateArgs<W> a, double *__restrict__ part, int64_t N, int64_t Ni)
{
    // grid y: instances of Ni points side by side (one instance: Ni = N)
    __shared__ double smax[TPB], smin[TPB];
    const int64_t base = (int64_t)blockIdx.y * Ni;
    double rate = 0.0, rmin = INFINITY;
    for (int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x; I < Ni; I += (int64_t)gridDim.x * TPB) {
        double rQ;
        const double acc = rate_at(a, base + I, N, rQ);
        rmin = rQ < rmin ? rQ : (rQ != rQ ? rQ : rmin);
        rate = acc > rate ? acc : (acc != acc ? acc : rate);
    }
    block_maxmin(smax, smin, rate, rmin);
    if (threadIdx.x == 0) {
        const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = smax[0];
        part[2 * b + 1] = smin[0];
    }
}

// one block per instance of a batch (invh of the instance's own grid; batch
// instances are Cartesian and unstretched)
template <typename W>
__global__ void __launch_bounds__(TPB) k_rate_batch(RateArgs<W> a, BatchGeom g, const int *__restrict__ mask,
                                                    double *__restrict__ rate_out, double *__restrict__ rmin_out)
{
    __shared__ double smax[TPB], smin[TPB];
    const int64_t i = blockIdx.x;
    if (mask && !mask[i]) return;
    a.invh[0] = 1.0 / g.h[i];
    const int64_t o = g.off[i];
    const int n = g.n[i];
    double rate = 0.0, rmin = INFINITY;
    for (int x = threadIdx.x; x < n; x += TPB) {
        double rQ;
        const double acc = rate_at<W, true>(a, o + x, g.N, rQ);
        rmin = rQ < rmin ? rQ : (rQ != rQ ? rQ : rmin);
        rate = acc > rate ? acc : (acc != acc ? acc : rate);
    }
    block_maxmin(smax, smin, rate, rmin);
    if (threadIdx.x == 0) {
        rate_out[i] = smax[0];
        rmin_out[i] = smin[0];
    }
}

// one block per instance: its nb partials -> rate[i], rmin[i]
__global__ void __launch_bounds__(TPB) k_rate_final(const double *__restrict__ part, int nb, double *__restrict__ rate,
                                                    double *__restrict__ rmin)
{
    __shared__ double smax[TPB], smin[TPB];
    part += (int64_t)2 * nb * blockIdx.x;
    double mx = 0.0, mn = INFINITY;
    for (int b = threadIdx.x; b < nb; b += TPB) {
        double x = part[2 * b], y = part[2 * b + 1];
        mx = (x > mx || x != x) ? x : mx;
        mn = (y < mn || y != y) ? y : mn;
    }
    smax[threadIdx.x] = mx;
    smin[threadIdx.x] = mn;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double x = smax[threadIdx.x + s], y = smin[threadIdx.x + s];
            double X = smax[threadIdx.x], Y = smin[threadIdx.x];
            smax[threadIdx.x] = (x > X || x != x) ? x : X;
            smin[threadIdx.x] = (y < Y || y != y) ? y : Y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rate[blockIdx.x] = smax[0];
        rmin[blockIdx.x] = smin[0];
    }
}

// ---------------------------------------------------------------------------
// boundary.jl:232-243 SlipWallBC, problem.jl:196-207 DirichletBC (constant)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t plane_index(const PlaneArgs &a, int64_t t)
{
    const int d = a.dim;
    const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
    const int64_t ia = t % a.n[oa], ib = t / a.n[oa];
    int64_t idx[3];
    idx[d] = a.side == 0 ? 0 : a.n[d] - 1;
    idx[oa] = ia;
    idx[ob] = ib;
    return idx[0] + a.n[0] * (idx[1] + a.n[1] * idx[2]);
}

__global__ void __launch_bounds__(TPB) k_slipwall(PlaneArgs a, double *__restrict__ Q)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= a.count) return;
    const int64_t I = plane_index(a, t);
    const int64_t N = a.N;
    double rho = 0.0;
    for (int k = 0; k < a.Ns; k++) rho += Q[(int64_t)k * N + I];
    const int64_t mc = (int64_t)(a.Ns + a.dim) * N + I;
    const double mn = Q[mc];
    Q[(int64_t)(a.Ns + 3) * N + I] -= 0.5 * mn * mn / rho;
    Q[mc] = 0;
}

__global__ void __launch_bounds__(TPB) k_dirichlet(PlaneArgs a, double *__restrict__ Q)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= a.count) return;
    const int64_t I = plane_index(a, t);
    for (int c = 0; c < a.Ns + 4; c++) Q[(int64_t)c * a.N + I] = a.state[c];
}

// boundary.jl:245-262 NoSlipWallBC enforce!: remove the kinetic energy, zero
// the momentum, and at an isothermal wall set rho e from the EOS at Twall
// (boundary.jl:203-221 wall_internal_energy)
__global__ void __launch_bounds__(TPB) k_noslip(PlaneArgs a, double *__restrict__ Q)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= a.count) return;
    const int64_t I = plane_index(a, t);
    const int64_t N = a.N;
    double rho = 0.0;
    for (int k = 0; k < a.Ns; k++) rho += Q[(int64_t)k * N + I];
    double *m1 = Q + (int64_t)a.Ns * N + I, *m2 = m1 + N, *m3 = m2 + N, *E = m3 + N;
    const double ke = 0.5 * (*m1 * *m1 + *m2 * *m2 + *m3 * *m3) / rho;
    *E -= ke;
    *m1 = 0;
    *m2 = 0;
    *m3 = 0;
    if (a.Twall > 0) {
        double re = 0.0;
        if (a.sg) {   // boundary.jl:213-214
            re = Q[I] * a.cvk[0] * a.Twall + a.sg_pinf;
        } else if (a.n9) {
            for (int k = 0; k < a.Ns; k++) re += Q[(int64_t)k * N + I] * n9_e(*a.n9, k, a.Twall);
        } else {
            for (int k = 0; k < a.Ns; k++) re += Q[(int64_t)k * N + I] * a.cvk[k];
            re *= a.Twall;
        }
        *E = re;
    }
}

// boundary.jl:264-276 ExtrapolationBC enforce!: copy the adjacent interior plane
__global__ void __launch_bounds__(TPB) k_extrapolate(PlaneArgs a, double *__restrict__ Q)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= a.count) return;
    const int64_t I = plane_index(a, t);
    const int64_t st = a.dim == 0 ? 1 : a.dim == 1 ? a.n[0] : a.n[0] * a.n[1];
    const int64_t J = a.side == 0 ? I + st : I - st;
    for (int c = 0; c < a.Ns + 4; c++) Q[(int64_t)c * a.N + I] = Q[(int64_t)c * a.N + J];
}

// ---------------------------------------------------------------------------
// NSCBC LODI corrections of dQ on one face plane (nscbc.jl:96-152 outflow,
// :236-306 inflow).  The plane's normal and transverse pressure (inflow:
// density) derivatives come from full-field line operators run beforehand.
// Arithmetic is FP64 in the reference's operation order for every precision.
// ---------------------------------------------------------------------------
template <typename W>
__device__ __forceinline__ void nscbc_point(const NscbcArgs<W> &a, int64_t t)
{
    if (t >= a.count) return;
    const int d = a.dim;
    const int oa = d == 0 ? 1 : 0, ob = d == 2 ? 1 : 2;
    int64_t idx[3];
    idx[d] = a.side == 0 ? 0 : a.n[d] - 1;
    idx[oa] = t % a.n[oa];
    idx[ob] = t / a.n[oa];
    const int64_t I = idx[0] + a.n[0] * (idx[1] + a.n[1] * idx[2]);
    const int64_t N = a.N;
    const double rho = a.rho[I], c = a.c[I], p = a.p[I], Tp = a.T[I], cpm = a.cp[I];
    const double uv[3] = {a.vel[0][I], a.vel[1][I], a.vel[2][I]};
    const double un = uv[d];
    const double Ma = fabs(un) / c;
    if (!(Ma < 1)) return;   // supersonic point: every wave leaves
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int im = a.Ns;
    // inv_h of the face's dimensions (nscbc.jl:108,125,130,257); 1 on unit geometry
    double ih[3] = {1.0, 1.0, 1.0};
    if (a.met.kind != CLES_METRIC_CARTESIAN || a.met.m[0] || a.met.m[1] || a.met.m[2]) {
        double hs[3];
        metric_hs(a.met, idx, hs);
        for (int k = 0; k < 3; k++) ih[k] = 1 / hs[k];
    }
    // physics.jl:142-145 eos_phi for an ideal mixture; :211 for a stiffened gas
    const double Rm = p / (rho * Tp);
    const double phi = a.sg_phi > 0 ? a.sg_phi : cpm / Rm - 1;
    const double ke = 0.5 * (uv[0] * uv[0] + uv[1] * uv[1] + uv[2] * uv[2]);
    const double dpn = ih[d] * (double)a.dpn[I];
    const double dun = a.gu[d * 3 + d][I];
    if (!a.inflow) {
        const double sgn = a.side == 1 ? -1.0 : 1.0;
        const double Lcomp = a.side == 1 ? (un - c) * (dpn - rho * c * dun) : (un + c) * (dpn + rho * c * dun);
        double tin = 0.0;
        if (a.act1) {
            const double ut = uv[t1];
            tin += ut * ih[t1] * (double)a.dt1[I] + rho * c * c * (double)a.gu[t1 * 3 + t1][I] +
                   sgn * rho * c * ut * (double)a.gu[t1 * 3 + d][I];
        }
        if (a.act2) {
            const double ut = uv[t2];
            tin += ut * ih[t2] * (double)a.dt2[I] + rho * c * c * (double)a.gu[t2 * 3 + t2][I] +
                   sgn * rho * c * ut * (double)a.gu[t2 * 3 + d][I];
        }
        const double bt = a.beta_t < 0 ? Ma : a.beta_t;
        const double K = a.sigma * (1 - Ma * Ma) * c / a.Lref;
        const double dL = K * (p - a.pinf) - bt * tin - Lcomp;
        const double dd1 = dL / (2 * c * c);
        const double dd2 = dL / 2;
        const double dd3 = sgn * dL / (2 * rho * c);
        for (int k = 0; k < a.Ns; k++) {
            W *q = a.dQ + (int64_t)k * N + I;
            *q = (W)((double)*q - (double)a.Y[k][I] * dd1);
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            W *q = a.dQ + (int64_t)(im + m) * N + I;
            *q = (W)((double)*q - (uv[m] * dd1 + (m == d ? rho * dd3 : 0.0)));
        }
        W *qe = a.dQ + (int64_t)(im + 3) * N + I;
        *qe = (W)((double)*qe - (ke * dd1 + phi * dd2 + rho * un * dd3));
        return;
    }
    const double drn = ih[d] * (double)a.dt1[I];
    const double K = c / a.Lref;
    const double L1c = (un - c) * (dpn - rho * c * dun);
    const double L5c = (un + c) * (dpn + rho * c * dun);
    const double L2c = un * (c * c * drn - dpn);
    const double rel_ac = a.eta_u * rho * c * c * (1 - Ma * Ma) / a.Lref * (un - a.ut[d]);
    const bool lowface = a.side == 0;
    const double dL1 = lowface ? 0.0 : (-rel_ac - L1c);
    const double dL5 = lowface ? (rel_ac - L5c) : 0.0;
    const double c3 = c * c * c;
    const double dL2 = a.eta_T * rho * c3 * (a.Tt - Tp) / (a.Lref * Tp) - L2c;
    const double dd1 = dL2 / (c * c) + (dL5 + dL1) / (2 * c * c);
    const double dd2 = (dL5 + dL1) / 2;
    const double dd3 = (dL5 - dL1) / (2 * rho * c);
    const double dd4 = a.eta_t * K * (uv[t1] - a.ut[t1]) - un * (double)a.gu[d * 3 + t1][I];
    const double dd5 = a.eta_t * K * (uv[t2] - a.ut[t2]) - un * (double)a.gu[d * 3 + t2][I];
    const double cvm = cpm - Rm;
    double sphiY = 0.0;
    for (int k = 0; k < a.Ns; k++) {
        const double Yk = a.Y[k][I];
        const double gYk = a.gY[k] ? (double)a.gY[k][I] : 0.0;
        const double dLs = a.eta_Y * K * (Yk - a.Yt[k]) - un * gYk;
        W *q = a.dQ + (int64_t)k * N + I;
        *q = (W)((double)*q - (Yk * dd1 + rho * dLs));
        // physics.jl:148-152 eos_dphi_dY; a Nasa9Mixture takes cv_k(T) (physics.jl:504-510)
        const double cvk = a.n9 ? n9_cp(*a.n9, k, Tp) - a.Rk[k] : a.cvk[k];
        if (!(a.sg_phi > 0)) sphiY += (cvk * Rm - a.Rk[k] * cvm) / (Rm * Rm) * dLs;   // :212: 0 for a stiffened gas
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const double extra = m == d ? rho * dd3 : m == t1 ? rho * dd4 : rho * dd5;
        W *q = a.dQ + (int64_t)(im + m) * N + I;
        *q = (W)((double)*q - (uv[m] * dd1 + extra));
    }
    W *qe = a.dQ + (int64_t)(im + 3) * N + I;
    *qe = (W)((double)*qe - (ke * dd1 + phi * dd2 + p * sphiY + rho * (un * dd3 + uv[t1] * dd4 + uv[t2] * dd5)));
}

template <typename W>
__global__ void __launch_bounds__(TPB) k_nscbc(NscbcArgs<W> a)
{
    nscbc_point(a, (int64_t)blockIdx.x * TPB + threadIdx.x);
}

// Both faces of one dimension in one launch (grid y = side): their planes are
// disjoint (n >= 2 along the normal), and each thread updates dQ at its own
// face point only, so the result is that of the two launches in order.
template <typename W>
__global__ void __launch_bounds__(TPB) k_nscbc_pair(NscbcArgs<W> a0, NscbcArgs<W> a1)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (blockIdx.y == 0)
        nscbc_point(a0, t);
    else
        nscbc_point(a1, t);
}

// ---------------------------------------------------------------------------
// Diagnostics (diagnostics.jl:41-80): deterministic FP64 block sums with the
// node-centred edge half-weights, then a single-block fixed-order final sum.
// KIND 0: 0.5|m|^2/rho from Q.  KIND 1: 0.5|omega|^2 from grad_u.
// ---------------------------------------------------------------------------
template <typename W, int KIND>
__global__ void __launch_bounds__(TPB) k_integrand_sum(DiagArgs<W> a, double *__restrict__ part)
{
    __shared__ double sh[TPB];
    const int64_t N = a.N;
    double acc = 0.0;
    for (int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x; I < N; I += (int64_t)gridDim.x * TPB) {
        const int64_t nx = a.n[0], ny = a.n[1];
        const int64_t ii[3] = {I % nx, (I / nx) % ny, I / (nx * ny)};
        double wq = 1.0;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (a.closed[d] && ((ii[d] == 0 && !a.fold_lo[d]) || (ii[d] == a.n[d] - 1 && !a.fold_hi[d]))) wq *= 0.5;
        double f;
        if constexpr (KIND == 0) {
            double rho = 0.0;
            for (int sp = 0; sp < a.Ns; sp++) rho += a.Q[(int64_t)sp * N + I];
            const double m1 = a.Q[(int64_t)a.Ns * N + I], m2 = a.Q[(int64_t)(a.Ns + 1) * N + I],
                         m3 = a.Q[(int64_t)(a.Ns + 2) * N + I];
            f = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
        } else {
            auto G = [&](int d, int j) -> double { return a.g.gu[d * 3 + j] ? (double)a.g.gu[d * 3 + j][I] : 0.0; };
            const double w1 = G(1, 2) - G(2, 1), w2 = G(2, 0) - G(0, 2), w3 = G(0, 1) - G(1, 0);
            f = 0.5 * (w1 * w1 + w2 * w2 + w3 * w3);
        }
        if (metric_on(a.met, 0) || a.met.m[1] || a.met.m[2]) {   // diagnostics.jl:73-75: f / inv_J
            const double inv_J = 1 / metric_J(a.met, ii);
            acc += wq * f / inv_J;
        } else {
            acc += wq * f;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(TPB) k_sum_final(const double *__restrict__ part, int nb, double scale,
                                                   double *__restrict__ out)
{
    __shared__ double sh[TPB];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += TPB) acc += part[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

// ---------------------------------------------------------------------------
// Batches of one-dimensional instances
// ---------------------------------------------------------------------------
// artificial.jl:67-94 along x of each instance (one block per instance)
template <typename W>
__global__ void __launch_bounds__(TPB) k_delta4_batch(D4Multi<W> fm, BatchGeom g)
{
    const W *__restrict__ f = fm.f[blockIdx.y];
    W *__restrict__ out = fm.out[blockIdx.y];
    const int wpow = fm.wpow[blockIdx.y];
    const int64_t i = blockIdx.x;
    const int64_t o = g.off[i];
    const int n = g.n[i];
    const double hd = g.h[i];
    const double wd = wpow == 2 ? hd * hd : hd;
    const double D4[5] = {1.0, -4.0, 6.0, -4.0, 1.0};
    for (int x = threadIdx.x; x < n; x += TPB) {
        double acc = 0.0;
#pragma unroll
        for (int m = -2; m <= 2; m++) {
            int j = x + m;
            if (g.periodic) {
                if (j < 0) j += n;
                else if (j >= n) j -= n;
            } else {
                j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
            }
            acc += D4[m + 2] * (double)f[o + j];
        }
        W acc_out = 0;
        acc_out += (W)(wd * fabs(acc));
        out[o + x] = acc_out;
    }
}

__device__ __forceinline__ double j_max(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
__device__ __forceinline__ double j_min(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }

// timestep.jl:29-42 predicted_dt, stepcontrol.jl:184-204 check_step and the
// tfinal clip of timestep.jl:352, per instance
__global__ void __launch_bounds__(TPB) k_batch_dt(BatchCtl c)
{
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= c.count) return;
    if (!c.active[i]) {   // dt keeps its last value (a rejected candidate stays readable)
        c.stepped[i] = 0;
        return;
    }
    const double rate = c.rate[i];
    double r = rate;
    if (c.predict > 0 && c.rate_prev[i] > 0) r = j_max(r, r + c.predict * (r - c.rate_prev[i]));
    double dt = c.cfl / r;
    if (c.max_growth > 0 && c.dt_prev[i] > 0) dt = j_min(dt, c.max_growth * c.dt_prev[i]);
    int code = CLES_OK;
    const double seen = c.dt_seen[i];
    if (!isfinite(dt)) code = CLES_ERR_NONFINITE;
    else if (!(c.rhomin[i] > 0)) code = CLES_ERR_NEGATIVE_DENSITY;
    else if (!(dt > 5.391247e-44)) code = CLES_ERR_PLANCK;
    else if (!(dt >= c.dt_min)) code = CLES_ERR_DT_MIN;
    else if (c.dt_min_ratio > 0 && seen > 0 && !(dt >= c.dt_min_ratio * seen)) code = CLES_ERR_DT_COLLAPSE;
    if (code != CLES_OK) {
        c.status[i] = code;
        c.active[i] = 0;
        c.stepped[i] = 0;
        c.dt[i] = dt;   // the rejected candidate (nothing steps with it)
        return;
    }
    c.dt_seen[i] = j_max(seen, dt);
    c.dt[i] = j_min(dt, c.tfinal[i] - c.t[i]);
    c.stepped[i] = 1;
}

__global__ void __launch_bounds__(TPB) k_batch_post(BatchCtl c)
{
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= c.count) return;
    if (!c.stepped[i]) {
        c.filt[i] = 0;
        return;
    }
    const double dt = c.dt[i];
    c.t[i] += dt;
    const int64_t st = c.step[i] + 1;
    c.step[i] = st;
    c.dt_prev[i] = dt;
    c.rate_prev[i] = c.rate[i];
    if (c.dts) {
        const int64_t k = st - 1 - c.step0[i];
        if (k >= 0 && k < c.nrec) c.dts[i * c.nrec + k] = dt;
    }
    c.filt[i] = (c.filter_interval > 0 && st % c.filter_interval == 0) ? 1 : 0;
    if (!(c.t[i] < c.tfinal[i]) || st >= c.nmax[i]) c.active[i] = 0;
}

// boundary.jl:232-243 SlipWallBC and problem.jl:196-207 DirichletBC (constant)
// at both ends of every instance
__global__ void __launch_bounds__(TPB) k_batch_bcs(BatchGeom g, BatchBC b, double *__restrict__ Q)
{
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= 2 * g.count) return;
    const int64_t i = t >> 1;
    const int side = (int)(t & 1);
    if (b.mask && !b.mask[i]) return;
    const int code = side ? b.hi : b.lo;
    const int64_t I = g.off[i] + (side ? g.n[i] - 1 : 0);
    const int64_t N = g.N;
    const int C = b.Ns + 4;
    if (code == CLES_BC_DIRICHLET) {
        const double *q = b.state + (i * 2 + side) * C;
        for (int c = 0; c < C; c++) Q[(int64_t)c * N + I] = q[c];
    } else if (code == CLES_BC_SLIPWALL) {
        double rho = 0.0;
        for (int k = 0; k < b.Ns; k++) rho += Q[(int64_t)k * N + I];
        const int64_t mc = (int64_t)b.Ns * N + I;
        const double mn = Q[mc];
        Q[(int64_t)(b.Ns + 3) * N + I] -= 0.5 * mn * mn / rho;
        Q[mc] = 0;
    }
}

// number of nonzero flags (single block, fixed order)
__global__ void __launch_bounds__(TPB) k_batch_count(const int *__restrict__ flags, int64_t count,
                                                     int *__restrict__ out)
{
    __shared__ int sh[TPB];
    int acc = 0;
    for (int64_t i = threadIdx.x; i < count; i += TPB) acc += flags[i] ? 1 : 0;
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// diagnostics.jl:41-80 kinetic energy of each instance (one block per
// instance, fixed summation order)
__global__ void __launch_bounds__(TPB) k_ke_batch(const double *__restrict__ Q, int Ns, BatchGeom g,
                                                  double *__restrict__ ke)
{
    __shared__ double sh[TPB];
    const int64_t i = blockIdx.x;
    const int64_t o = g.off[i], N = g.N;
    const int n = g.n[i];
    double acc = 0.0;
    for (int x = threadIdx.x; x < n; x += TPB) {
        const int64_t I = o + x;
        const double wq = (!g.periodic && (x == 0 || x == n - 1)) ? 0.5 : 1.0;
        double rho = 0.0;
        for (int sp = 0; sp < Ns; sp++) rho += Q[(int64_t)sp * N + I];
        const double m1 = Q[(int64_t)Ns * N + I], m2 = Q[(int64_t)(Ns + 1) * N + I],
                     m3 = Q[(int64_t)(Ns + 2) * N + I];
        acc += wq * (0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) ke[i] = sh[0] * g.h[i];
}

// plain streaming copy in 16-byte words (calibration kernel for byte counters)
__global__ void __launch_bounds__(TPB) k_copy16(uint4 *__restrict__ d, const uint4 *__restrict__ s, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n16; i += (int64_t)gridDim.x * TPB) d[i] = s[i];
}

__global__ void __launch_bounds__(TPB) k_copy1(unsigned char *__restrict__ d, const unsigned char *__restrict__ s,
                                               int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) d[i] = s[i];
}

template <typename Dst, typename Src>
__global__ void __launch_bounds__(TPB) k_convert(Dst *__restrict__ d, const Src *__restrict__ s, int64_t N)
{
    const int64_t I = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (I < N) d[I] = (Dst)s[I];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define CHECK_LAUNCH() CLES_HIP(hipGetLastError())

template <typename W>
void launch_primitives(const PrimArgs<W> &a, int64_t N, hipStream_t st)
{
    // rho, u, v, w, p, T, c, cp, the mass fractions and the e-sensor input
    const bool nt = (double)N * sizeof(W) * (9 + a.Ns) > 256.0 * 1024 * 1024;
    const bool ig = !a.sg && !a.n9;
    if (nt && ig)
        hipLaunchKernelGGL((k_primitives<W, true, true>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    else if (nt)
        hipLaunchKernelGGL((k_primitives<W, true, false>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    else if (ig)
        hipLaunchKernelGGL((k_primitives<W, false, true>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    else
        hipLaunchKernelGGL((k_primitives<W, false, false>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    CHECK_LAUNCH();
}

template <typename W>
void launch_strain(const GradPtrs<W> &g, W *out, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL(k_strain<W>, dim3(nblocks(N)), dim3(TPB), 0, st, g, out, N);
    CHECK_LAUNCH();
}

template <typename W>
void launch_delta4(const D4Args<W> &a, int mode, W *out, hipStream_t st)
{
    if (mode == 0) {
        D4Multi<W> m{};
        m.nf = 1;
        m.f[0] = a.f;
        m.out[0] = out;
        m.wpow[0] = 1;   // a.wd holds the weights
        launch_delta4_multi<W>(a, m, st);
        return;
    } else
        hipLaunchKernelGGL((k_delta4<W, 1>), dim3(nblocks(a.N)), dim3(TPB), 0, st, a, out);
    CHECK_LAUNCH();
}

// a.wd: h_d (the fields' wpow apply), or the weights themselves with wpow 1
template <typename W>
void launch_delta4_multi(const D4Args<W> &a, const D4Multi<W> &m, hipStream_t st)
{
    if (m.nf < 1 || m.nf > D4_MF) throw Error{CLES_ERR_INVALID, "delta4 takes 1..4 fields per launch"};
    const int gx = (int)((a.n[0] + D4T_X - 1) / D4T_X), gy = (int)((a.n[1] + D4T_Y - 1) / D4T_Y);
    int zl = D4T_Z;
    while (zl > 1 && (int64_t)gx * gy * ((a.n[2] + zl - 1) / zl) * m.nf < 1024) zl /= 2;
    const int gz = (int)((a.n[2] + zl - 1) / zl);
    const int64_t nb = (int64_t)gx * gy * gz;
    hipLaunchKernelGGL((k_delta4_tile<W>), dim3((unsigned)nb, (unsigned)m.nf), dim3(TPB), 0, st, a, m, gx, gy, zl);
    CHECK_LAUNCH();
}

template <typename W>
void launch_gate_beta(const GradPtrs<W> &g, W *beta, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL(k_gate_beta<W>, dim3(nblocks(N)), dim3(TPB), 0, st, g, beta, N);
    CHECK_LAUNCH();
}

template <typename W>
void launch_dilatation_prep(const GradPtrs<W> &g, W *div, W *sw, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL(k_dilatation_prep<W>, dim3(nblocks(N)), dim3(TPB), 0, st, g, div, sw, N);
    CHECK_LAUNCH();
}

template <typename W>
void launch_dilatation_beta(const W *rho, const W *sensor, const W *sw, W *beta, double C_beta, int64_t N,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_dilatation_beta<W>, dim3(nblocks(N)), dim3(TPB), 0, st, rho, sensor, sw, beta, C_beta, N);
    CHECK_LAUNCH();
}

template <typename W>
void launch_fluxes(const FluxArgs<W> &a, int64_t N, hipStream_t st)
{
    const int nact = (a.active[0] != 0) + (a.active[1] != 0) + (a.active[2] != 0);
    const bool nt = (double)N * sizeof(W) * (a.Ns + 4) * nact > 256.0 * 1024 * 1024;
    // one species: the lean instantiation in FP32 (-4 % at 256^3); in FP64 the
    // full one measured 2 % faster (same results: Ns is checked at run time)
    if (a.Ns == 2 && !a.n9) {
        if (nt)
            hipLaunchKernelGGL((k_fluxes_two<W, true>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
        else
            hipLaunchKernelGGL((k_fluxes_two<W, false>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    } else if (a.Ns > 1 || sizeof(W) == 8) {
        if (nt)
            hipLaunchKernelGGL((k_fluxes<W, true, true>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
        else
            hipLaunchKernelGGL((k_fluxes<W, false, true>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    } else {
        if (nt)
            hipLaunchKernelGGL((k_fluxes<W, true, false>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
        else
            hipLaunchKernelGGL((k_fluxes<W, false, false>), dim3(nblocks(N)), dim3(TPB), 0, st, a, N);
    }
    CHECK_LAUNCH();
}

template <typename W>
void launch_lsrk(double *Q, double *du, const W *dQ, double A, double B, double dt, const double *dtp,
                 const int *gate, int64_t NC, int64_t N, unsigned ipts, hipStream_t st)
{
    hipLaunchKernelGGL(k_lsrk<W>, dim3(nblocks(NC)), dim3(TPB), 0, st, Q, du, dQ, A, B, dt, dtp, gate, NC, N,
                       ipts);
    CHECK_LAUNCH();
}

template <typename W>
void launch_rate(const RateArgs<W> &a, double *part, int nb, double *out, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL(k_rate<W>, dim3(nb), dim3(TPB), 0, st, a, part, N, N);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_rate_final, dim3(1), dim3(TPB), 0, st, part, nb, out, out + 1);
    CHECK_LAUNCH();
}

// ni instances of N / ni points side by side: rate[i], rmin[i]; part holds
// 2 nb ni doubles
template <typename W>
void launch_rate_inst(const RateArgs<W> &a, double *part, int nb, int ni, double *rate, double *rmin, int64_t N,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_rate<W>, dim3(nb, ni), dim3(TPB), 0, st, a, part, N, N / ni);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_rate_final, dim3(ni), dim3(TPB), 0, st, part, nb, rate, rmin);
    CHECK_LAUNCH();
}

template <typename W>
void launch_rate_batch(const RateArgs<W> &a, const BatchGeom &g, const int *mask, double *rate, double *rhomin,
                       hipStream_t st)
{
    if (g.count <= 0) return;
    hipLaunchKernelGGL(k_rate_batch<W>, dim3((unsigned)g.count), dim3(TPB), 0, st, a, g, mask, rate, rhomin);
    CHECK_LAUNCH();
}
template void launch_rate_batch<double>(const RateArgs<double> &, const BatchGeom &, const int *, double *,
                                        double *, hipStream_t);
template void launch_rate_batch<float>(const RateArgs<float> &, const BatchGeom &, const int *, double *, double *,
                                       hipStream_t);

template <typename W>
void launch_delta4_batch_multi(const D4Multi<W> &m, const BatchGeom &g, hipStream_t st)
{
    if (g.count <= 0) return;
    if (m.nf < 1 || m.nf > D4_MF) throw Error{CLES_ERR_INVALID, "delta4 takes 1..4 fields per launch"};
    hipLaunchKernelGGL(k_delta4_batch<W>, dim3((unsigned)g.count, (unsigned)m.nf), dim3(TPB), 0, st, m, g);
    CHECK_LAUNCH();
}
template <typename W>
void launch_delta4_batch(const W *f, W *out, const BatchGeom &g, int wpow, hipStream_t st)
{
    D4Multi<W> m{};
    m.nf = 1;
    m.f[0] = f;
    m.out[0] = out;
    m.wpow[0] = wpow;
    launch_delta4_batch_multi<W>(m, g, st);
}
template void launch_delta4_batch<double>(const double *, double *, const BatchGeom &, int, hipStream_t);
template void launch_delta4_batch<float>(const float *, float *, const BatchGeom &, int, hipStream_t);
template void launch_delta4_batch_multi<double>(const D4Multi<double> &, const BatchGeom &, hipStream_t);
template void launch_delta4_batch_multi<float>(const D4Multi<float> &, const BatchGeom &, hipStream_t);

void launch_batch_dt(const BatchCtl &c, hipStream_t st)
{
    if (c.count <= 0) return;
    hipLaunchKernelGGL(k_batch_dt, dim3(nblocks(c.count)), dim3(TPB), 0, st, c);
    CHECK_LAUNCH();
}

void launch_batch_post(const BatchCtl &c, hipStream_t st)
{
    if (c.count <= 0) return;
    hipLaunchKernelGGL(k_batch_post, dim3(nblocks(c.count)), dim3(TPB), 0, st, c);
    CHECK_LAUNCH();
}

void launch_batch_bcs(const BatchGeom &g, const BatchBC &b, double *Q, hipStream_t st)
{
    if (g.count <= 0) return;
    hipLaunchKernelGGL(k_batch_bcs, dim3(nblocks(2 * g.count)), dim3(TPB), 0, st, g, b, Q);
    CHECK_LAUNCH();
}

void launch_batch_count(const int *flags, int64_t count, int *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(TPB), 0, st, flags, count, out);
    CHECK_LAUNCH();
}

void launch_ke_batch(const double *Q, int Ns, const BatchGeom &g, double *ke, hipStream_t st)
{
    if (g.count <= 0) return;
    hipLaunchKernelGGL(k_ke_batch, dim3((unsigned)g.count), dim3(TPB), 0, st, Q, Ns, g, ke);
    CHECK_LAUNCH();
}

template <typename W>
void launch_nscbc(const NscbcArgs<W> &a, hipStream_t st)
{
    if (a.count <= 0) return;
    hipLaunchKernelGGL((k_nscbc<W>), dim3(nblocks(a.count)), dim3(TPB), 0, st, a);
    CHECK_LAUNCH();
}
template void launch_nscbc<double>(const NscbcArgs<double> &, hipStream_t);
template void launch_nscbc<float>(const NscbcArgs<float> &, hipStream_t);

template <typename W>
void launch_nscbc_pair(const NscbcArgs<W> &a0, const NscbcArgs<W> &a1, hipStream_t st)
{
    if (a0.dim != a1.dim || a0.side == a1.side || a0.n[a0.dim] < 2 || a0.count != a1.count)
        throw Error{CLES_ERR_INVALID, "an NSCBC face pair is the two faces of one dimension"};
    if (a0.count <= 0) return;
    hipLaunchKernelGGL((k_nscbc_pair<W>), dim3(nblocks(a0.count), 2), dim3(TPB), 0, st, a0, a1);
    CHECK_LAUNCH();
}
template void launch_nscbc_pair<double>(const NscbcArgs<double> &, const NscbcArgs<double> &, hipStream_t);
template void launch_nscbc_pair<float>(const NscbcArgs<float> &, const NscbcArgs<float> &, hipStream_t);

void launch_slipwall(const PlaneArgs &a, double *Q, hipStream_t st)
{
    hipLaunchKernelGGL(k_slipwall, dim3(nblocks(a.count)), dim3(TPB), 0, st, a, Q);
    CHECK_LAUNCH();
}

void launch_dirichlet(const PlaneArgs &a, double *Q, hipStream_t st)
{
    hipLaunchKernelGGL(k_dirichlet, dim3(nblocks(a.count)), dim3(TPB), 0, st, a, Q);
    CHECK_LAUNCH();
}

void launch_noslip(const PlaneArgs &a, double *Q, hipStream_t st)
{
    hipLaunchKernelGGL(k_noslip, dim3(nblocks(a.count)), dim3(TPB), 0, st, a, Q);
    CHECK_LAUNCH();
}

void launch_extrapolate(const PlaneArgs &a, double *Q, hipStream_t st)
{
    if (a.n[a.dim] < 2) throw Error{CLES_ERR_INVALID, "ExtrapolationBC needs two planes"};
    hipLaunchKernelGGL(k_extrapolate, dim3(nblocks(a.count)), dim3(TPB), 0, st, a, Q);
    CHECK_LAUNCH();
}

template <typename W>
void launch_integral(const DiagArgs<W> &a, int kind, double *part, int nb, double scale, double *out, hipStream_t st)
{
    if (kind == 0)
        hipLaunchKernelGGL((k_integrand_sum<W, 0>), dim3(nb), dim3(TPB), 0, st, a, part);
    else
        hipLaunchKernelGGL((k_integrand_sum<W, 1>), dim3(nb), dim3(TPB), 0, st, a, part);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(TPB), 0, st, part, nb, scale, out);
    CHECK_LAUNCH();
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const uintptr_t a = (uintptr_t)dst | (uintptr_t)src;
    if (a % 16 == 0) {
        const int64_t n16 = (int64_t)(bytes / 16);
        if (n16 > 0) {
            const unsigned nb = (unsigned)std::min<int64_t>((n16 + TPB - 1) / TPB, 256 * 64);
            hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(TPB), 0, st, (uint4 *)dst, (const uint4 *)src, n16);
            CHECK_LAUNCH();
        }
        const size_t done = (size_t)n16 * 16;
        if (bytes > done) {
            hipLaunchKernelGGL(k_copy1, dim3(1), dim3(TPB), 0, st, (unsigned char *)dst + done,
                               (const unsigned char *)src + done, (int64_t)(bytes - done));
            CHECK_LAUNCH();
        }
    } else if (bytes) {
        hipLaunchKernelGGL(k_copy1, dim3(nblocks((int64_t)bytes)), dim3(TPB), 0, st, (unsigned char *)dst,
                           (const unsigned char *)src, (int64_t)bytes);
        CHECK_LAUNCH();
    }
}

template <typename Dst, typename Src>
void launch_convert(Dst *d, const Src *s, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL((k_convert<Dst, Src>), dim3(nblocks(N)), dim3(TPB), 0, st, d, s, N);
    CHECK_LAUNCH();
}

#define CLES_INST(W)                                                                                       \
    template void launch_primitives<W>(const PrimArgs<W> &, int64_t, hipStream_t);                        \
    template void launch_strain<W>(const GradPtrs<W> &, W *, int64_t, hipStream_t);                        \
    template void launch_delta4<W>(const D4Args<W> &, int, W *, hipStream_t);                              \
    template void launch_delta4_multi<W>(const D4Args<W> &, const D4Multi<W> &, hipStream_t);              \
    template void launch_gate_beta<W>(const GradPtrs<W> &, W *, int64_t, hipStream_t);                     \
    template void launch_dilatation_prep<W>(const GradPtrs<W> &, W *, W *, int64_t, hipStream_t);          \
    template void launch_dilatation_beta<W>(const W *, const W *, const W *, W *, double, int64_t,         \
                                            hipStream_t);                                                  \
    template void launch_fluxes<W>(const FluxArgs<W> &, int64_t, hipStream_t);                             \
    template void launch_lsrk<W>(double *, double *, const W *, double, double, double, const double *,   \
                                 const int *, int64_t, int64_t, unsigned, hipStream_t);                    \
    template void launch_rate<W>(const RateArgs<W> &, double *, int, double *, int64_t, hipStream_t);      \
    template void launch_rate_inst<W>(const RateArgs<W> &, double *, int, int, double *, double *, int64_t,  \
                                      hipStream_t);                                                        \
    template void launch_integral<W>(const DiagArgs<W> &, int, double *, int, double, double *, hipStream_t);
CLES_INST(double)
CLES_INST(float)
template void launch_convert<double, float>(double *, const float *, int64_t, hipStream_t);
template void launch_convert<float, double>(float *, const double *, int64_t, hipStream_t);
template void launch_convert<double, double>(double *, const double *, int64_t, hipStream_t);

template <typename W>
void launch_metric_scale(const MetricFields<W> &fs, const Metric1D &g, int mode, int dim, int64_t N, hipStream_t st)
{
    if (fs.nf <= 0 || N <= 0) return;
    const unsigned nb = nblocks(N);
    if (mode == 0)
        hipLaunchKernelGGL((k_metric_scale<W, 0>), dim3(nb), dim3(TPB), 0, st, fs, g, dim, N);
    else if (mode == 1)
        hipLaunchKernelGGL((k_metric_scale<W, 1>), dim3(nb), dim3(TPB), 0, st, fs, g, dim, N);
    else if (mode == 3)
        hipLaunchKernelGGL((k_metric_scale<W, 3>), dim3(nb), dim3(TPB), 0, st, fs, g, dim, N);
    else
        hipLaunchKernelGGL((k_metric_scale<W, 2>), dim3(nb), dim3(TPB), 0, st, fs, g, dim, N);
    CHECK_LAUNCH();
}
template <typename W>
void launch_metric_grad(const CurvArgs<W> &a, hipStream_t st)
{
    if (a.N <= 0) return;
    hipLaunchKernelGGL((k_metric_grad<W>), dim3(nblocks(a.N)), dim3(TPB), 0, st, a);
    CHECK_LAUNCH();
}
template <typename W>
void launch_metric_sources(const CurvArgs<W> &a, hipStream_t st)
{
    if (a.N <= 0) return;
    hipLaunchKernelGGL((k_metric_sources<W>), dim3(nblocks(a.N)), dim3(TPB), 0, st, a);
    CHECK_LAUNCH();
}
template <typename T>
void launch_pair_forward(const T *f, T *w, const PairMap &pm, double sf, hipStream_t st)
{
    const int64_t N = pm.n[0] * pm.n[1] * pm.n[2];
    if (N <= 0) return;
    hipLaunchKernelGGL((k_pair_forward<T>), dim3(nblocks(N)), dim3(TPB), 0, st, f, w, pm, sf, N);
    CHECK_LAUNCH();
}
template <typename T>
void launch_pair_backward(const T *re, const T *ro, T *dst, const PairMap &pm, double sf, int epi, const int *gate,
                          hipStream_t st)
{
    const int64_t N = pm.n[0] * pm.n[1] * pm.n[2];
    if (N <= 0) return;
    hipLaunchKernelGGL((k_pair_backward<T>), dim3(nblocks(N)), dim3(TPB), 0, st, re, ro, dst, pm, sf, epi, gate, N);
    CHECK_LAUNCH();
}
template <typename W>
void launch_body_force(const double *Q, W *dQ, int Ns, int64_t N, const double (&g)[3], hipStream_t st)
{
    if (N <= 0) return;
    hipLaunchKernelGGL((k_body_force<W>), dim3(nblocks(N)), dim3(TPB), 0, st, Q, dQ, Ns, N, g[0], g[1], g[2]);
    CHECK_LAUNCH();
}
template void launch_body_force<double>(const double *, double *, int, int64_t, const double (&)[3], hipStream_t);
template void launch_body_force<float>(const double *, float *, int, int64_t, const double (&)[3], hipStream_t);
template void launch_pair_forward<double>(const double *, double *, const PairMap &, double, hipStream_t);
template void launch_pair_forward<float>(const float *, float *, const PairMap &, double, hipStream_t);
template void launch_pair_backward<double>(const double *, const double *, double *, const PairMap &, double, int,
                                           const int *, hipStream_t);
template void launch_pair_backward<float>(const float *, const float *, float *, const PairMap &, double, int,
                                          const int *, hipStream_t);
template void launch_metric_grad<double>(const CurvArgs<double> &, hipStream_t);
template void launch_metric_grad<float>(const CurvArgs<float> &, hipStream_t);
template void launch_metric_sources<double>(const CurvArgs<double> &, hipStream_t);
template void launch_metric_sources<float>(const CurvArgs<float> &, hipStream_t);
template void launch_metric_scale<double>(const MetricFields<double> &, const Metric1D &, int, int, int64_t,
                                          hipStream_t);
template void launch_metric_scale<float>(const MetricFields<float> &, const Metric1D &, int, int, int64_t,
                                         hipStream_t);


}  // namespace cles
