/*
 * oracle.c -- CPU restatement of CompactLES.jl's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the product
 * library never links or calls it.  Built with -O2 -ffp-contract=off so that
 * no a*b+c is contracted, matching Julia without @fastmath/@muladd.
 *
 * Layout: the reference's padded column-major arrays.  For active dimension d
 * the padded extent is n_local[d] + 2*H (H = 4), collapsed dimensions carry no
 * padding (src/decomposition.jl:115-136).  One rank, P = 1.
 *
 * Citations are path:line into stillyslalom/CompactLES.jl.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Julia's max/min on floats propagate NaN (Base.max, Base.min). */
static inline double jmax(double a, double b)
{
    if (a != a || b != b) return NAN;
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}
static inline double jmin(double a, double b)
{
    if (a != a || b != b) return NAN;
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}

/* ------------------------------------------------------------------------ */
/* Scheme presets: src/kernels.jl:61-111, src/kernels_banded.jl:54-64        */
/* ------------------------------------------------------------------------ */

static void scheme_zero(or_scheme *s) { memset(s, 0, sizeof(*s)); }

/* kernels.jl:61-67 lele_d1_6 */
void or_lele_d1_6(or_scheme *s)
{
    scheme_zero(s);
    s->q = 1; s->lhs[0] = 1.0 / 3.0; s->a0 = 0.0;
    s->M = 2; s->coeffs[0] = 7.0 / 9.0; s->coeffs[1] = 1.0 / 36.0;
    s->symmetric = 0; s->nclosure = 2;
    s->cl_lhs[0][0] = 0.0; s->cl_lhs[0][1] = 1.0; s->cl_lhs[0][2] = 2.0;
    s->cl_rhs_len[0] = 3; s->cl_rhs[0][0] = -5.0 / 2.0; s->cl_rhs[0][1] = 2.0; s->cl_rhs[0][2] = 1.0 / 2.0;
    s->cl_lhs[1][0] = 1.0 / 4.0; s->cl_lhs[1][1] = 1.0; s->cl_lhs[1][2] = 1.0 / 4.0;
    s->cl_rhs_len[1] = 3; s->cl_rhs[1][0] = -3.0 / 4.0; s->cl_rhs[1][1] = 0.0; s->cl_rhs[1][2] = 3.0 / 4.0;
}

/* kernels.jl:75-79 pade_d1_4 */
void or_pade_d1_4(or_scheme *s)
{
    scheme_zero(s);
    s->q = 1; s->lhs[0] = 1.0 / 4.0;
    s->M = 1; s->coeffs[0] = 3.0 / 4.0;
    s->symmetric = 0; s->nclosure = 1;
    s->cl_lhs[0][0] = 0.0; s->cl_lhs[0][1] = 1.0; s->cl_lhs[0][2] = 2.0;
    s->cl_rhs_len[0] = 3; s->cl_rhs[0][0] = -5.0 / 2.0; s->cl_rhs[0][1] = 2.0; s->cl_rhs[0][2] = 1.0 / 2.0;
}

/* kernels_banded.jl:54-64 lele_d1_10 */
void or_lele_d1_10(or_scheme *s)
{
    scheme_zero(s);
    s->q = 2; s->lhs[0] = 1.0 / 2.0; s->lhs[1] = 1.0 / 20.0;
    s->M = 3; s->coeffs[0] = 17.0 / 24.0; s->coeffs[1] = 101.0 / 600.0; s->coeffs[2] = 1.0 / 600.0;
    s->symmetric = 0; s->nclosure = 3;
    double r0[5] = {0, 0, 1, 2, 0}, r1[5] = {0, 0.25, 1, 0.25, 0}, r2[5] = {0, 1.0 / 3.0, 1, 1.0 / 3.0, 0};
    for (int t = 0; t < 5; t++) { s->cl_lhs[0][t] = r0[t]; s->cl_lhs[1][t] = r1[t]; s->cl_lhs[2][t] = r2[t]; }
    s->cl_rhs_len[0] = 3; s->cl_rhs[0][0] = -5.0 / 2.0; s->cl_rhs[0][1] = 2.0; s->cl_rhs[0][2] = 1.0 / 2.0;
    s->cl_rhs_len[1] = 3; s->cl_rhs[1][0] = -3.0 / 4.0; s->cl_rhs[1][1] = 0.0; s->cl_rhs[1][2] = 3.0 / 4.0;
    s->cl_rhs_len[2] = 5;
    s->cl_rhs[2][0] = -1.0 / 36.0; s->cl_rhs[2][1] = -7.0 / 9.0; s->cl_rhs[2][2] = 0.0;
    s->cl_rhs[2][3] = 7.0 / 9.0; s->cl_rhs[2][4] = 1.0 / 36.0;
}

/* kernels.jl:89-111 compact_filter: centred closure rows of order 2, 4, 6.
 * ctr(c) = [c[m+1]/2 for m = L-1:-1:1; c[1]; c[m+1]/2 for m = 1:L-1]. */
static int ctr(const double *c, int L, double *out)
{
    int k = 0;
    for (int m = L - 1; m >= 1; m--) out[k++] = c[m] / 2;
    out[k++] = c[0];
    for (int m = 1; m <= L - 1; m++) out[k++] = c[m] / 2;
    return k;
}

void or_compact_filter(or_scheme *s, double af)
{
    scheme_zero(s);
    double a0 = (93 + 70 * af) / 128;
    double a1 = (7 + 18 * af) / 16;
    double a2 = (-7 + 14 * af) / 32;
    double a3 = (1 - 2 * af) / 16;
    double a4 = (-1 + 2 * af) / 128;
    double b2[2] = {(1 + 2 * af) / 2, (1 + 2 * af) / 2};
    double b4[3] = {(5 + 6 * af) / 8, (1 + 2 * af) / 2, (-1 + 2 * af) / 8};
    double b6[4] = {(11 + 10 * af) / 16, (15 + 34 * af) / 32, (-3 + 6 * af) / 16, (1 - 2 * af) / 32};
    s->q = 1; s->lhs[0] = af; s->a0 = a0;
    s->M = 4; s->coeffs[0] = a1 / 2; s->coeffs[1] = a2 / 2; s->coeffs[2] = a3 / 2; s->coeffs[3] = a4 / 2;
    s->symmetric = 1; s->nclosure = 4;
    s->cl_lhs[0][0] = 0; s->cl_lhs[0][1] = 1; s->cl_lhs[0][2] = 0;
    s->cl_rhs_len[0] = 1; s->cl_rhs[0][0] = 1;
    for (int j = 1; j < 4; j++) { s->cl_lhs[j][0] = af; s->cl_lhs[j][1] = 1; s->cl_lhs[j][2] = af; }
    s->cl_rhs_len[1] = ctr(b2, 2, s->cl_rhs[1]);
    s->cl_rhs_len[2] = ctr(b4, 3, s->cl_rhs[2]);
    s->cl_rhs_len[3] = ctr(b6, 4, s->cl_rhs[3]);
}

/* ------------------------------------------------------------------------ */
/* Dense LU with partial pivoting for the reduced system (LAPACK getf2 /    */
/* getrs semantics, as Julia's lu!/ldiv! on a small dense matrix;           */
/* src/tridiag.jl:123,157; src/banded.jl:169,199).                          */
/* ------------------------------------------------------------------------ */
typedef struct { int n; double a[16]; int piv[4]; } dense_lu;   /* column-major */

static void lu_factor(dense_lu *F)
{
    int n = F->n;
    double *A = F->a;
#define AA(i, j) A[(i) + (j) * n]
    for (int j = 0; j < n; j++) {
        int p = j;
        double amax = fabs(AA(j, j));
        for (int i = j + 1; i < n; i++)
            if (fabs(AA(i, j)) > amax) { amax = fabs(AA(i, j)); p = i; }
        F->piv[j] = p;
        if (p != j)
            for (int k = 0; k < n; k++) { double t = AA(j, k); AA(j, k) = AA(p, k); AA(p, k) = t; }
        double r = 1.0 / AA(j, j);
        for (int i = j + 1; i < n; i++) AA(i, j) *= r;
        for (int k = j + 1; k < n; k++)
            for (int i = j + 1; i < n; i++) AA(i, k) -= AA(i, j) * AA(j, k);
    }
}

static void lu_solve(const dense_lu *F, double *b)
{
    int n = F->n;
    const double *A = F->a;
    for (int j = 0; j < n; j++)
        if (F->piv[j] != j) { double t = b[j]; b[j] = b[F->piv[j]]; b[F->piv[j]] = t; }
    for (int k = 0; k < n; k++)
        if (b[k] != 0.0)
            for (int i = k + 1; i < n; i++) b[i] -= b[k] * AA(i, k);
    for (int k = n - 1; k >= 0; k--)
        if (b[k] != 0.0) {
            b[k] = b[k] / AA(k, k);
            for (int i = 0; i < k; i++) b[i] -= b[k] * AA(i, k);
        }
#undef AA
}

/* ------------------------------------------------------------------------ */
/* Directional plan                                                          */
/* ------------------------------------------------------------------------ */
struct or_plan {
    int dim, n, q, M, sym, nc;
    int64_t nloc[3];
    int halo[3];
    int64_t pad[3];
    int periodic, lo_closed, hi_closed;
    int lo_fold, hi_fold;   /* parity folds at the edges: sigma_g (0: none) */
    double a0, ci[OR_MAXM];
    int cl_len[OR_MAXCL];
    double clo[OR_MAXCL][OR_MAXRHS], chi[OR_MAXCL][OR_MAXRHS];
    /* tridiagonal factor (tridiag.jl:24-43) */
    double *l, *dinv, *c;
    double *v, *w;
    /* banded factor (banded.jl:31-61): L[m*n + k], U[s*n + i] */
    double *BL, *BU;
    double *V, *W;          /* n x q, column t at [t*n] */
    int hasred;
    dense_lu red;
};

/* tridiag.jl:31-43 TriFactor */
static void trifactor(or_plan *p, const double *a, const double *b, const double *c)
{
    int n = p->n;
    double d = b[0];
    p->dinv[0] = 1.0 / d;
    p->l[0] = 0.0;
    for (int i = 1; i < n; i++) {
        p->l[i] = a[i] * p->dinv[i - 1];
        d = b[i] - p->l[i] * c[i - 1];
        p->dinv[i] = 1.0 / d;
    }
    memcpy(p->c, c, sizeof(double) * n);
}

/* tridiag.jl:46-57 solve_col! (stride st) */
static void tri_solve_col(const or_plan *p, double *x, int64_t st)
{
    int n = p->n;
    for (int i = 1; i < n; i++) x[i * st] -= p->l[i] * x[(i - 1) * st];
    x[(n - 1) * st] *= p->dinv[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i * st] = (x[i * st] - p->c[i] * x[(i + 1) * st]) * p->dinv[i];
}

/* banded.jl:38-61 BandFactor; Ab[(q+s)*n + i] = A[i, i+s] (0-based s+q row) */
static void bandfactor(or_plan *p, const double *Ab)
{
    int n = p->n, q = p->q;
    double *Wk = malloc(sizeof(double) * (2 * q + 1) * n);
    memcpy(Wk, Ab, sizeof(double) * (2 * q + 1) * n);
#define WK(r, i) Wk[(r) * n + (i)]     /* r = 0..2q  (Julia row r+1) */
    memset(p->BL, 0, sizeof(double) * q * n);
    for (int k = 0; k < n - 1; k++) {
        double piv = WK(q, k);
        int mmax = q < n - 1 - k ? q : n - 1 - k;
        for (int m = 1; m <= mmax; m++) {
            int i = k + m;
            double lm = WK(q - m, i) / piv;
            p->BL[(m - 1) * n + k] = lm;
            for (int t = 1; t <= mmax; t++) WK(q + t - m, i) -= lm * WK(q + t, k);
        }
    }
    memset(p->BU, 0, sizeof(double) * (q + 1) * n);
    for (int i = 0; i < n; i++)
        for (int s = 0; s <= q && s <= n - 1 - i; s++) p->BU[s * n + i] = WK(q + s, i);
#undef WK
    free(Wk);
}

/* banded.jl:63-80 solve_col! (non-transposed form, divides by U[1,i]) */
static void band_solve_col(const or_plan *p, double *x, int64_t st)
{
    int n = p->n, q = p->q;
    for (int k = 0; k < n - 1; k++) {
        double xk = x[k * st];
        for (int m = 1; m <= q && m <= n - 1 - k; m++) x[(k + m) * st] -= p->BL[(m - 1) * n + k] * xk;
    }
    for (int i = n - 1; i >= 0; i--) {
        double acc = x[i * st];
        for (int t = 1; t <= q && t <= n - 1 - i; t++) acc -= p->BU[t * n + i] * x[(i + t) * st];
        x[i * st] = acc / p->BU[i];
    }
}

static or_plan *plan_create_fold(const or_scheme *s, int dim, const int64_t n_local[3], const int halo[3],
                                 int periodic, double h, int lo_fold, int hi_fold, char *msg, int msglen);

or_plan *or_plan_create(const or_scheme *s, int dim, const int64_t n_local[3],
                        const int halo[3], int periodic, double h, char *msg, int msglen)
{
    return plan_create_fold(s, dim, n_local, halo, periodic, h, 0, 0, msg, msglen);
}

or_plan *or_plan_create_fold(const or_scheme *s, int dim, const int64_t n_local[3], const int halo[3],
                             int periodic, double h, int lo_fold, int hi_fold, char *msg, int msglen)
{
    return plan_create_fold(s, dim, n_local, halo, periodic, h, lo_fold, hi_fold, msg, msglen);
}

void or_fold_fill_halos(double *f, const int64_t n_local[3], const int halo[3], int dim, int lo, int hi,
                        int sigma)
{
    const int d = dim - 1, H = halo[d];
    const int64_t n = n_local[d];
    int64_t pad[3];
    for (int k = 0; k < 3; k++) pad[k] = n_local[k] + 2 * halo[k];
    const int64_t st = d == 0 ? 1 : d == 1 ? pad[0] : pad[0] * pad[1];
    const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
    for (int64_t b = 0; b < pad[o2]; b++)
        for (int64_t a = 0; a < pad[o1]; a++) {
            int64_t ii[3] = {0, 0, 0};
            ii[o1] = a;
            ii[o2] = b;
            double *base = f + ii[0] + pad[0] * (ii[1] + pad[1] * ii[2]);
            for (int q = 1; q <= H; q++) {
                if (lo) base[(H - q) * st] = (double)sigma * base[(H + q - 1) * st];
                if (hi) base[(H + n + q - 1) * st] = (double)sigma * base[(H + n - q) * st];
            }
        }
}

/* operators.jl:44-101 / operators_banded.jl:29-125 plan_direction with
 * lo_fold = sigma_g: the parity fold of a half-offset low edge moves the LHS
 * coupling to the ghost unknowns onto the mirrored interior columns, and the
 * RHS keeps the interior stencil reading mirror-filled halos (fold_fill). */
static or_plan *plan_create_fold(const or_scheme *s, int dim, const int64_t n_local[3], const int halo[3],
                                 int periodic, double h, int lo_fold, int hi_fold, char *msg, int msglen)
{
    int n = (int)n_local[dim - 1];
    int nc = s->nclosure, M = s->M, q = s->q;
    int nmin = 2 * nc + 1 > 2 * M + 1 ? 2 * nc + 1 : 2 * M + 1;
    if (q > 1 && 2 * q + 1 > nmin) nmin = 2 * q + 1;
    /* operators.jl:50-53, operators_banded.jl:369-374 */
    if (n < nmin) {
        if (msg) snprintf(msg, msglen, "local extent %d along dim %d too small for scheme (need >= %d)", n, dim, nmin);
        return NULL;
    }
    if (M > halo[dim - 1]) {
        if (msg) snprintf(msg, msglen, "stencil half-width %d exceeds halo width %d", M, halo[dim - 1]);
        return NULL;
    }
    or_plan *p = calloc(1, sizeof(or_plan));
    p->dim = dim; p->n = n; p->q = q; p->M = M; p->sym = s->symmetric; p->nc = nc;
    for (int d = 0; d < 3; d++) { p->nloc[d] = n_local[d]; p->halo[d] = halo[d]; p->pad[d] = n_local[d] + 2 * halo[d]; }
    p->periodic = periodic;
    p->lo_fold = periodic ? 0 : lo_fold;
    p->hi_fold = periodic ? 0 : hi_fold;
    p->lo_closed = !periodic && !p->lo_fold;   /* at_lo_edge on one rank (decomposition.jl:137-140) */
    p->hi_closed = !periodic && !p->hi_fold;

    /* operators.jl:87-92: prescale by 1/h (derivatives) or 1 (filters) */
    double hinv = 1.0 / h;
    double sgn = s->symmetric ? 1.0 : -1.0;
    double scale = s->symmetric ? 1.0 : hinv;
    p->a0 = s->a0;
    for (int m = 0; m < M; m++) p->ci[m] = s->coeffs[m] * scale;
    for (int j = 0; j < nc; j++) {
        p->cl_len[j] = s->cl_rhs_len[j];
        for (int k = 0; k < s->cl_rhs_len[j]; k++) {
            p->clo[j][k] = s->cl_rhs[j][k] * scale;
            p->chi[j][k] = s->cl_rhs[j][k] * (scale * sgn);
        }
    }

    if (q == 1) {
        /* operators.jl:59-85 */
        double alpha = s->lhs[0];
        double *a = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
        for (int i = 0; i < n; i++) { a[i] = alpha; b[i] = 1.0; c[i] = alpha; }
        if (p->lo_fold) b[0] += (double)p->lo_fold * alpha;       /* operators.jl:61-66 */
        if (p->hi_fold) b[n - 1] += (double)p->hi_fold * alpha;   /* operators.jl:67-70 */
        if (p->lo_closed)
            for (int j = 0; j < nc; j++) { a[j] = s->cl_lhs[j][0]; b[j] = s->cl_lhs[j][1]; c[j] = s->cl_lhs[j][2]; }
        if (p->hi_closed)
            for (int j = 0; j < nc; j++) {
                int r = n - 1 - j;
                a[r] = s->cl_lhs[j][2]; b[r] = s->cl_lhs[j][1]; c[r] = s->cl_lhs[j][0];
            }
        double aL = (p->lo_closed || p->lo_fold) ? 0.0 : alpha;
        double cR = (p->hi_closed || p->hi_fold) ? 0.0 : alpha;
        p->l = calloc(n, sizeof(double)); p->dinv = calloc(n, sizeof(double)); p->c = calloc(n, sizeof(double));
        p->v = calloc(n, sizeof(double)); p->w = calloc(n, sizeof(double));
        trifactor(p, a, b, c);
        /* tridiag.jl:85-124 LineSolver at P = 1 */
        if (aL != 0) { p->v[0] = aL; tri_solve_col(p, p->v, 1); }
        if (cR != 0) { p->w[n - 1] = cR; tri_solve_col(p, p->w, 1); }
        p->hasred = periodic;
        if (p->hasred) {
            double R[4] = {0, 0, 0, 0};   /* column-major 2x2 */
            R[0] += 1; R[3] += 1;
            /* cprev = 2, cnext = 1 (1-based) */
            R[0 + 1 * 2] += p->v[0];      /* R[1,2] += v1 */
            R[0 + 0 * 2] += p->w[0];      /* R[1,1] += w1 */
            R[1 + 1 * 2] += p->v[n - 1];  /* R[2,2] += vn */
            R[1 + 0 * 2] += p->w[n - 1];  /* R[2,1] += wn */
            p->red.n = 2;
            memcpy(p->red.a, R, sizeof(R));
            lu_factor(&p->red);
        }
        free(a); free(b); free(c);
    } else {
        /* operators_banded.jl:381-442 */
        double *Ab = calloc((size_t)(2 * q + 1) * n, sizeof(double));
#define AB(r, i) Ab[(r) * n + (i)]
        for (int i = 0; i < n; i++) AB(q, i) = 1.0;
        for (int sd = 1; sd <= q; sd++)
            for (int i = 0; i < n; i++) { AB(q + sd, i) = s->lhs[sd - 1]; AB(q - sd, i) = s->lhs[sd - 1]; }
        if (p->lo_fold)   /* operators_banded.jl:55-64 (1-based i, m: jcol = m - i + 1) */
            for (int i = 1; i <= q; i++)
                for (int m = 1; m <= q; m++) {
                    if (i - m >= 1) continue;
                    int jcol = m - i + 1;
                    AB(q + (jcol - i), i - 1) += (double)p->lo_fold * s->lhs[m - 1];
                }
        if (p->hi_fold)   /* operators_banded.jl:65-73: ghost column i + m > n maps onto 2n + 1 - (i + m) */
            for (int i = n - q + 1; i <= n; i++)
                for (int m = 1; m <= q; m++) {
                    if (i + m <= n) continue;
                    int jcol = 2 * n + 1 - (i + m);
                    AB(q + (jcol - i), i - 1) += (double)p->hi_fold * s->lhs[m - 1];
                }
        if (p->lo_closed)
            for (int j = 0; j < nc; j++)
                for (int sd = -q; sd <= q; sd++) AB(q + sd, j) = s->cl_lhs[j][q + sd];
        if (p->hi_closed)
            for (int j = 0; j < nc; j++) {
                int r = n - 1 - j;
                for (int sd = -q; sd <= q; sd++) AB(q + sd, r) = s->cl_lhs[j][q - sd];
            }
#undef AB
        double AL[4][4] = {{0}}, CR[4][4] = {{0}};
        int anyL = 0, anyR = 0;
        if (!p->lo_closed && !p->lo_fold)
            for (int r = 1; r <= q; r++)
                for (int t = r; t <= q; t++) { AL[r - 1][t - 1] = s->lhs[q + r - t - 1]; anyL |= AL[r - 1][t - 1] != 0; }
        if (!p->hi_closed && !p->hi_fold)
            for (int r = 1; r <= q; r++)
                for (int t = 1; t <= r; t++) { CR[r - 1][t - 1] = s->lhs[t + q - r - 1]; anyR |= CR[r - 1][t - 1] != 0; }
        p->BL = calloc((size_t)q * n, sizeof(double));
        p->BU = calloc((size_t)(q + 1) * n, sizeof(double));
        p->V = calloc((size_t)q * n, sizeof(double));
        p->W = calloc((size_t)q * n, sizeof(double));
        bandfactor(p, Ab);
        free(Ab);
        /* banded.jl:108-170 BandLineSolver at P = 1 */
        if (anyL)
            for (int t = 0; t < q; t++) {
                for (int r = 0; r < q; r++) p->V[t * n + r] = AL[r][t];
                band_solve_col(p, p->V + t * n, 1);
            }
        if (anyR)
            for (int t = 0; t < q; t++) {
                for (int r = 0; r < q; r++) p->W[t * n + n - q + r] = CR[r][t];
                band_solve_col(p, p->W + t * n, 1);
            }
        p->hasred = periodic;
        if (p->hasred) {
            int m2 = 2 * q;
            double R[16] = {0};
#define RR(i, j) R[(i) + (j) * m2]
            for (int r = 0; r < q; r++) { RR(r, r) += 1; RR(q + r, q + r) += 1; }
            int cprev = q, cnext = 0;   /* 0-based offsets, rank 0 of 1 */
            for (int t = 0; t < q; t++)
                for (int r = 0; r < q; r++) {
                    RR(r, cprev + t) += p->V[t * n + r];
                    RR(q + r, cprev + t) += p->V[t * n + n - q + r];
                    RR(r, cnext + t) += p->W[t * n + r];
                    RR(q + r, cnext + t) += p->W[t * n + n - q + r];
                }
#undef RR
            p->red.n = m2;
            memcpy(p->red.a, R, sizeof(double) * m2 * m2);
            lu_factor(&p->red);
        }
    }
    return p;
}

void or_plan_destroy(or_plan *p)
{
    if (!p) return;
    free(p->l); free(p->dinv); free(p->c); free(p->v); free(p->w);
    free(p->BL); free(p->BU); free(p->V); free(p->W);
    free(p);
}

/* Padded linear index of line coordinate i (0-based interior) and
 * orthogonal coordinates (j, k) in ascending dim order (operators.jl:105-107). */
static inline int64_t gidx(const or_plan *p, int i, int64_t j, int64_t k)
{
    int64_t a[3];
    int d = p->dim;
    if (d == 1) { a[0] = i; a[1] = j; a[2] = k; }
    else if (d == 2) { a[0] = j; a[1] = i; a[2] = k; }
    else { a[0] = j; a[1] = k; a[2] = i; }
    return (a[0] + p->halo[0]) + p->pad[0] * ((a[1] + p->halo[1]) + p->pad[1] * (a[2] + p->halo[2]));
}

/* One fill value (operators.jl:113-174 and :245-309 compute the same sums). */
static inline double fill_row(const or_plan *p, const double *f, int i, int64_t j, int64_t k)
{
    int n = p->n, nc = p->nc;
    if (p->lo_closed && i < nc) {
        double acc = 0.0;
        for (int kap = 0; kap < p->cl_len[i]; kap++) acc += p->clo[i][kap] * f[gidx(p, kap, j, k)];
        return acc;
    }
    if (p->hi_closed && i >= n - nc) {
        int jr = n - 1 - i;
        double acc = 0.0;
        for (int kap = 0; kap < p->cl_len[jr]; kap++) acc += p->chi[jr][kap] * f[gidx(p, n - 1 - kap, j, k)];
        return acc;
    }
    if (p->sym) {
        double acc = p->a0 * f[gidx(p, i, j, k)];
        for (int m = 1; m <= p->M; m++) acc += p->ci[m - 1] * (f[gidx(p, i + m, j, k)] + f[gidx(p, i - m, j, k)]);
        return acc;
    }
    double acc = 0.0;
    for (int m = 1; m <= p->M; m++) acc += p->ci[m - 1] * (f[gidx(p, i + m, j, k)] - f[gidx(p, i - m, j, k)]);
    return acc;
}

/* Solve one line held in x[0..n-1] (contiguous).  transposed = 1 mirrors the
 * row-sweep forms of src/lines_transposed.jl:22-150 (y/z sweeps); 0 mirrors
 * tridiag.jl:136-170 / banded.jl:177-214 (x sweep). */
static void solve_line(const or_plan *p, double *x, int transposed)
{
    int n = p->n, q = p->q;
    if (q == 1) {
        tri_solve_col(p, x, 1);
        if (!p->hasred) return;
        double z[2] = {x[0], x[n - 1]};
        lu_solve(&p->red, z);
        double xl = z[1], xr = z[0];   /* cprev = 2, cnext = 1 */
        for (int i = 0; i < n; i++) x[i] -= p->v[i] * xl + p->w[i] * xr;
        return;
    }
    if (!transposed) {
        band_solve_col(p, x, 1);
    } else {
        /* lines_transposed.jl:93-113 */
        for (int k = 0; k < n - 1; k++)
            for (int m = 1; m <= q && m <= n - 1 - k; m++) {
                double lm = p->BL[(m - 1) * n + k];
                if (lm == 0.0) continue;
                x[k + m] -= lm * x[k];
            }
        for (int i = n - 1; i >= 0; i--) {
            for (int t = 1; t <= q && t <= n - 1 - i; t++) x[i] -= p->BU[t * n + i] * x[i + t];
            double u0 = 1.0 / p->BU[i];
            x[i] *= u0;
        }
    }
    if (!p->hasred) return;
    double z[4];
    for (int r = 0; r < q; r++) { z[r] = x[r]; z[q + r] = x[n - q + r]; }
    lu_solve(&p->red, z);
    /* cprev = q (tail), cnext = 0 (head) */
    if (!transposed) {
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int t = 0; t < q; t++) acc += p->V[t * n + i] * z[q + t] + p->W[t * n + i] * z[t];
            x[i] -= acc;
        }
    } else {
        for (int i = 0; i < n; i++)
            for (int t = 0; t < q; t++) x[i] -= p->V[t * n + i] * z[q + t] + p->W[t * n + i] * z[t];
    }
}

/* operators.jl:199-215 apply_along! */
void or_apply_along(const or_plan *p, double *out, const double *f)
{
    int d = p->dim, n = p->n;
    int o1 = d == 1 ? 1 : 0, o2 = d == 3 ? 1 : 2;   /* orthogonal dims (0-based) */
    int64_t n1 = p->nloc[o1], n2 = p->nloc[o2];
    int64_t lines = n1 * n2;
    int transposed = d > 1;
#pragma omp parallel
    {
        double *x = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < lines; l++) {
            int64_t j = l % n1, k = l / n1;
            for (int i = 0; i < n; i++) x[i] = fill_row(p, f, i, j, k);
            solve_line(p, x, transposed);
            for (int i = 0; i < n; i++) out[gidx(p, i, j, k)] = x[i];
        }
        free(x);
    }
}

/* halo.jl:30-36 wrap_dim! on a padded field (dim 0-based) */
static void wrap_dim(double *f, const int64_t pad[3], const int64_t nloc[3], int H, int d)
{
    int64_t n = nloc[d];
    int64_t s[3] = {1, pad[0], pad[0] * pad[1]};
    int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
    for (int64_t b = 0; b < pad[o2]; b++)
        for (int64_t a = 0; a < pad[o1]; a++) {
            int64_t base = a * s[o1] + b * s[o2];
            for (int g = 0; g < H; g++) {
                f[base + g * s[d]] = f[base + (n + g) * s[d]];
                f[base + (n + H + g) * s[d]] = f[base + (H + g) * s[d]];
            }
        }
}

void or_wrap_halos(double *f, const int64_t n_local[3], const int halo[3], const int periodic[3])
{
    int64_t pad[3];
    for (int d = 0; d < 3; d++) pad[d] = n_local[d] + 2 * halo[d];
    for (int d = 0; d < 3; d++)
        if (halo[d] > 0 && periodic[d]) wrap_dim(f, pad, n_local, halo[d], d);
}

/* ------------------------------------------------------------------------ */
/* Solver                                                                    */
/* ------------------------------------------------------------------------ */
#define RKN 5
/* timestep.jl:5-21 */
static const double RKA[RKN] = {0.0,
    -567301805773.0 / 1357537059087.0,
    -2404267990393.0 / 2016746695238.0,
    -3550918686646.0 / 2091501179385.0,
    -1275806237668.0 / 842570457699.0};
static const double RKB[RKN] = {1432997174477.0 / 9575080441755.0,
    5161836677717.0 / 13612068292357.0,
    1720146321549.0 / 2090206949498.0,
    3134564353537.0 / 4481467310338.0,
    2277821191437.0 / 14882151754819.0};
static const double RKC[RKN] = {0.0,
    1432997174477.0 / 9575080441755.0,
    2526269341429.0 / 6820363962896.0,
    2006345519317.0 / 3224310063776.0,
    2802321613138.0 / 2924317926251.0};

struct or_solver {
    or_desc d;
    int Ns, C, im1, im2, im3, ie;        /* 0-based component indices */
    int active[3], periodic[3], H[3];
    int64_t n[3], pad[3], NP;
    double h[3];
    double Rk[OR_MAXSP], cvk[OR_MAXSP], cpk[OR_MAXSP];
    or_plan *dplan[3], *fplan[3];
    double *Q, *dQ, *du;
    double *rho, *u, *v, *w, *p, *T, *c, *cp, *Y[OR_MAXSP];
    double *gu[9];                       /* grad_u[d][j] at [d*3 + j] */
    double *gT[3], *gY[3][OR_MAXSP];
    double *mu_art, *beta_art, *kappa_art, *D_art[OR_MAXSP];
    double *strain_mag, *sensor, *sensor_sp, *tmp_a, *tmp_b;
    double *flux[3][OR_MAXSP + 4];
    double t, tstage, dt_prev, rate_prev;
    int64_t step;
    /* metric.jl:109-137 geometry over the padded extent, allocated when a
     * Stretch or a curvilinear metric makes it non-unit (geom = 1) */
    int geom;
    double *inv_J, *area[3], *inv_h[3], *inv_r, *cotr;
    /* coordinate-singularity folds (rhs.jl:155-279, folds.jl:56-73): which
     * ends fold, the antipodal pairing (pdim shifted by half a period, revdim
     * reversed; -1: none, then the fold is self-paired), the antipodal signs
     * of (u, v, w), the area factor's parity, and the parity-folded plans
     * (index 0: sigma_g = +1, 1: sigma_g = -1) */
    struct {
        int on, lo, hi, pdim, revdim, sv[3], area;
        or_plan *dp[2], *fp[2];
    } fold[3];
    double coord_shift[3];
    double *pairbuf, *pairout;
};

#define FOR_INTERIOR(s, I)                                                        \
    for (int64_t kk_ = 0; kk_ < (s)->n[2]; kk_++)                                 \
        for (int64_t jj_ = 0; jj_ < (s)->n[1]; jj_++)                             \
            for (int64_t ii_ = 0, I = pidx(s, 0, jj_, kk_); ii_ < (s)->n[0]; ii_++, I++)

static inline int64_t pidx(const or_solver *s, int64_t i, int64_t j, int64_t k)
{
    return (i + s->H[0]) + s->pad[0] * ((j + s->H[1]) + s->pad[1] * (k + s->H[2]));
}

static double *fld(or_solver *s) { return calloc((size_t)s->NP, sizeof(double)); }

/* rhs.jl:477-483 _scale_grad!: g *= inv_h[d] over the full array */
static void scale_grad(or_solver *s, double *g, int d)
{
    if (!s->geom) return;   /* inv_h == 1 */
    const double *ih = s->inv_h[d];
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < s->NP; I++) g[I] *= ih[I];
}

static void deriv_along_p(or_solver *s, double *out, double *f, int d, int sf);
#define deriv_along(s, out, f, d) deriv_along_p((s), (out), (f), (d), 1)
static int vel_parity(const or_solver *s, int d, int j);
static int cons_parity(const or_solver *s, int d, int c);

/* metric.jl:94-103 _phys_and_jac: the physical coordinate and mapping
 * Jacobian at padded index ip along d.  A stretched dimension takes its
 * values at the nearest interior node (halo layers beyond closed edges are
 * never read). */
static void phys_and_jac(const or_solver *s, int d, int64_t ip, const double *const *x, const double *const *jac,
                         double *xo, double *mo)
{
    const double xi = s->d.origin[d] + s->coord_shift[d] + (double)(ip - s->H[d]) * s->h[d];
    if (!jac || !jac[d]) {
        *xo = xi;
        *mo = 1.0;
        return;
    }
    int64_t i = ip - s->H[d];
    if (i < 0) i = 0;
    if (i > s->n[d] - 1) i = s->n[d] - 1;
    *xo = (x && x[d]) ? x[d][i] : xi;
    *mo = jac[d][i];
}

/* metric.jl:109-137 init_geometry! and :150-168 gcl_cotr! */
static void init_geometry(or_solver *s, const double *const *x, const double *const *jac)
{
    const double tiny = 1e-300;
    s->inv_J = fld(s);
    for (int d = 0; d < 3; d++) { s->area[d] = fld(s); s->inv_h[d] = fld(s); }
    s->inv_r = fld(s);
    s->cotr = fld(s);
    const int kind = s->d.metric;
    for (int64_t k = 0; k < s->pad[2]; k++)
        for (int64_t j = 0; j < s->pad[1]; j++)
            for (int64_t i = 0; i < s->pad[0]; i++) {
                double x1, m1, x2, m2, x3, m3;
                phys_and_jac(s, 0, i, x, jac, &x1, &m1);
                phys_and_jac(s, 1, j, x, jac, &x2, &m2);
                phys_and_jac(s, 2, k, x, jac, &x3, &m3);
                /* metric.jl:80-82 scalefactors */
                double h1 = 1.0, h2 = 1.0, h3 = 1.0;
                if (kind == 1) h2 = x1;
                else if (kind == 2) { h2 = x1; h3 = x1 * sin(x2); }
                h1 *= m1; h2 *= m2; h3 *= m3;
                h1 = fmax(h1, tiny); h2 = fmax(h2, tiny); h3 = fmax(h3, tiny);
                const double J = h1 * h2 * h3;
                const int64_t I = i + s->pad[0] * (j + s->pad[1] * k);
                s->inv_J[I] = 1 / J;
                s->area[0][I] = J / h1;
                s->area[1][I] = J / h2;
                s->area[2][I] = J / h3;
                s->inv_h[0][I] = 1 / h1;
                s->inv_h[1][I] = 1 / h2;
                s->inv_h[2][I] = 1 / h3;
                if (kind >= 1) s->inv_r[I] = 1 / fmax(x1, tiny);
                if (kind == 2) s->cotr[I] = cos(x2) / fmax(x1 * sin(x2), tiny);
            }
    if (kind == 2 && s->active[1]) {
        /* gcl_cotr!: cot(theta)/r := inv_J D_xi2(A_2) at the interior nodes */
        /* with the sign of the theta-momentum pressure flux across a pole fold */
        deriv_along_p(s, s->tmp_a, s->area[1], 1,
                      cons_parity(s, 1, s->im2) * vel_parity(s, 1, 1) * (s->fold[1].on ? s->fold[1].area : 1));
        for (int64_t k = 0; k < s->n[2]; k++)
            for (int64_t j = 0; j < s->n[1]; j++)
                for (int64_t i = 0; i < s->n[0]; i++) {
                    const int64_t I = pidx(s, i, j, k);
                    s->cotr[I] = s->inv_J[I] * s->tmp_a[I];
                }
    }
}

or_solver *or_solver_create(const or_desc *d, char *msg, int msglen)
{
    return or_solver_create_geom(d, NULL, NULL, msg, msglen);
}

or_solver *or_solver_create_ex(const or_desc *d, const double *const jac[3], char *msg, int msglen)
{
    return or_solver_create_geom(d, NULL, jac, msg, msglen);
}

/* setup with Stretch mappings (metric.jl:45-58): jac[d] holds dx/dxi at the
 * n_global[d] nodes xi = i / (n - 1) of a stretched (non-periodic) dimension;
 * its computational spacing is h = 1 / (n - 1) (rhs.jl:191-201). */
or_solver *or_solver_create_geom(const or_desc *d, const double *const x[3], const double *const jac[3], char *msg,
                                 int msglen)
{
    if (d->metric < 0 || d->metric > 2) {
        snprintf(msg, (size_t)msglen, "metric must be Cartesian, cylindrical or spherical");
        return NULL;
    }
    or_solver *s = calloc(1, sizeof(or_solver));
    s->d = *d;
    s->Ns = d->n_species;
    s->C = s->Ns + 4;
    s->im1 = s->Ns; s->im2 = s->Ns + 1; s->im3 = s->Ns + 2; s->ie = s->Ns + 3;   /* equations.jl:45-53 */
    s->NP = 1;
    for (int k = 0; k < 3; k++) {
        s->n[k] = d->n_global[k];
        s->active[k] = d->n_global[k] > 1;
        /* rhs.jl:182: periodic unless a closed BC is given on an active dim */
        s->periodic[k] = s->active[k] ? (d->bc[k][0] == OR_BC_PERIODIC) : 1;
        s->H[k] = s->active[k] ? 4 : 0;
        s->pad[k] = s->n[k] + 2 * s->H[k];
        s->NP *= s->pad[k];
        /* rhs.jl:198-204 */
        if (!s->active[k]) s->h[k] = 1.0;
        else s->h[k] = s->periodic[k] ? d->L_domain[k] / (double)d->n_global[k]
                                      : d->L_domain[k] / (double)(d->n_global[k] - 1);
        if (jac && jac[k]) {
            if (!s->active[k] || s->periodic[k]) {
                snprintf(msg, (size_t)msglen, "dimension %d: stretched dimensions must be non-periodic", k + 1);
                free(s);
                return NULL;
            }
            s->h[k] = 1.0 / (double)(d->n_global[k] - 1);
            s->geom = 1;
            /* the metric reads r (dim 1) and, spherical, theta (dim 2) */
            if (((d->metric >= 1 && k == 0) || (d->metric == 2 && k == 1)) && !(x && x[k])) {
                snprintf(msg, (size_t)msglen, "dimension %d: a stretched curvilinear coordinate needs its node "
                                              "coordinates", k + 1);
                free(s);
                return NULL;
            }
        }
    }
    if (d->metric != 0) s->geom = 1;
    /* rhs.jl:143-205: coordinate-singularity folds recognised from the BCs */
    {
        const char *err = NULL;
        for (int k = 0; k < 3 && !err; k++)
            for (int sd = 0; sd < 2; sd++) {
                const int bc = d->bc[k][sd];
                if ((bc == OR_BC_AXIS || bc == OR_BC_ORIGIN) && !(k == 0 && sd == 0))
                    err = bc == OR_BC_AXIS ? "AxisBC goes at the low end of dimension 1"
                                           : "OriginBC goes at the low end of dimension 1";
                if (bc == OR_BC_POLE && k != 1) err = "PoleBC goes on dimension 2 (theta)";
            }
        const int axis = d->bc[0][0] == OR_BC_AXIS, orig = d->bc[0][0] == OR_BC_ORIGIN;
        const int pl = d->bc[1][0] == OR_BC_POLE, ph = d->bc[1][1] == OR_BC_POLE;
        const double pi = 3.14159265358979323846;
        if (!err && pl != ph) err = "PoleBC must be applied at both ends of theta";
        if (!err && axis) {
            if (d->metric != 1) err = "AxisBC requires CylindricalMetric";
            else if (jac && jac[0]) err = "folded dimensions cannot be stretched";
            else if (fabs(d->origin[0]) >= 1e-14) err = "AxisBC requires origin[1] = 0";
            else if (s->n[1] > 1 && (s->n[1] % 2 || fabs(d->L_domain[1] - 2 * pi) > 1e-10))
                err = "resolved-theta AxisBC requires an even theta point count over 2 pi";
        }
        if (!err && orig) {
            if (d->metric != 2) err = "OriginBC requires SphericalMetric";
            else if (jac && jac[0]) err = "folded dimensions cannot be stretched";
            else if (fabs(d->origin[0]) >= 1e-14) err = "OriginBC requires origin[1] = 0";
            else if (s->n[2] > 1 && (s->n[2] % 2 || fabs(d->L_domain[2] - 2 * pi) > 1e-10))
                err = "resolved-phi OriginBC requires an even phi point count over 2 pi";
            else if (s->n[1] > 1 && fabs(2 * d->origin[1] + d->L_domain[1] - pi) > 1e-10)
                err = "OriginBC requires a theta range symmetric about pi/2";
            else if (s->n[1] > 1 && s->n[2] == 1 && s->n[1] % 2)
                err = "OriginBC with a collapsed phi needs an even theta point count";
        }
        if (!err && pl) {
            if (d->metric != 2) err = "PoleBC requires SphericalMetric";
            else if (jac && jac[1]) err = "folded dimensions cannot be stretched";
            else if (!(fabs(d->origin[1]) < 1e-14 && fabs(d->L_domain[1] - pi) <= 1e-10))
                err = "PoleBC requires the theta domain (0, pi)";
            else if (s->n[2] > 1 && (s->n[2] % 2 || fabs(d->L_domain[2] - 2 * pi) > 1e-10))
                err = "resolved-phi PoleBC requires an even phi point count over 2 pi";
        }
        if (err) {
            snprintf(msg, (size_t)msglen, "%s", err);
            free(s);
            return NULL;
        }
        /* rhs.jl:256-279 foldspec: sigvel and the area factor's parity */
        if (axis && s->active[0]) {
            s->fold[0].on = 1; s->fold[0].lo = 1;
            s->fold[0].pdim = s->active[1] ? 1 : -1; s->fold[0].revdim = -1;
            s->fold[0].sv[0] = -1; s->fold[0].sv[1] = -1; s->fold[0].sv[2] = 1;
            s->fold[0].area = -1;                                   /* A_1 = r */
        }
        if (orig && s->active[0]) {
            s->fold[0].on = 1; s->fold[0].lo = 1;
            s->fold[0].pdim = s->active[2] ? 2 : -1; s->fold[0].revdim = s->active[1] ? 1 : -1;
            s->fold[0].sv[0] = -1; s->fold[0].sv[1] = 1; s->fold[0].sv[2] = -1;
            s->fold[0].area = 1;                                    /* A_1 = r^2 sin(theta) */
        }
        if (pl && s->active[1]) {
            s->fold[1].on = 1; s->fold[1].lo = 1; s->fold[1].hi = 1;
            s->fold[1].pdim = s->active[2] ? 2 : -1; s->fold[1].revdim = -1;
            s->fold[1].sv[0] = 1; s->fold[1].sv[1] = -1; s->fold[1].sv[2] = -1;
            s->fold[1].area = -1;                                   /* A_2 = r sin(theta) */
        }
        /* rhs.jl:196-205: half-offset grids on folded dimensions */
        for (int k = 0; k < 3; k++) {
            if (!s->fold[k].on) continue;
            s->h[k] = s->fold[k].hi ? d->L_domain[k] / (double)s->n[k]
                                    : d->L_domain[k] / ((double)s->n[k] - 0.5);
            s->coord_shift[k] = s->h[k] / 2;
        }
    }
    /* boundary.jl:111-121 SwitchableBC: arms agree on periodicity, no folds */
    for (int k = 0; k < 3; k++)
        for (int sd = 0; sd < 2; sd++) {
            if (!d->switchable[k][sd]) continue;
            const int b0 = d->bc[k][sd], b1 = d->bc_after[k][sd];
            const char *err = NULL;
            if ((b0 == OR_BC_PERIODIC) != (b1 == OR_BC_PERIODIC)) err = "SwitchableBC: both conditions must agree on periodicity";
            if (b0 == OR_BC_AXIS || b0 == OR_BC_ORIGIN || b0 == OR_BC_POLE || b1 == OR_BC_AXIS || b1 == OR_BC_ORIGIN ||
                b1 == OR_BC_POLE)
                err = "SwitchableBC cannot wrap a fold condition";
            if (err) {
                snprintf(msg, (size_t)msglen, "%s", err);
                free(s);
                return NULL;
            }
        }
    /* nscbc.jl:309-316: NSCBC faces (either arm of a SwitchableBC) need a normal scale factor of one */
    for (int k = 0; k < 3; k++) {
        const int unit = d->metric == 0 || (d->metric == 1 && k != 1) || (d->metric == 2 && k == 0);
        for (int sd = 0; sd < 4; sd++) {
            if (sd >= 2 && !d->switchable[k][sd - 2]) continue;
            const int bc = sd < 2 ? d->bc[k][sd] : d->bc_after[k][sd - 2];
            if ((bc == OR_BC_NSCBC_IN || bc == OR_BC_NSCBC_OUT) && s->active[k] && !unit) {
                snprintf(msg, (size_t)msglen, "%s on dimension %d: the LODI formulation requires a face whose normal "
                         "metric scale factor is one", bc == OR_BC_NSCBC_IN ? "NSCBCInflowBC" : "NSCBCOutflowBC", k + 1);
                free(s);
                return NULL;
            }
        }
    }
    /* physics.jl:111-115 */
    for (int k = 0; k < s->Ns; k++) {
        s->Rk[k] = d->R[k];
        s->cvk[k] = d->R[k] / (d->gamma[k] - 1);
        s->cpk[k] = s->Rk[k] + s->cvk[k];
    }
    int halo[3] = {s->H[0], s->H[1], s->H[2]};
    for (int k = 0; k < 3; k++) {
        if (!s->active[k]) continue;
        s->dplan[k] = or_plan_create(&d->deriv, k + 1, s->n, halo, s->periodic[k], s->h[k], msg, msglen);
        s->fplan[k] = or_plan_create(&d->filt, k + 1, s->n, halo, s->periodic[k], s->h[k], msg, msglen);
        if (!s->dplan[k] || !s->fplan[k]) { or_solver_destroy(s); return NULL; }
    }
    for (int k = 0; k < 3; k++) {   /* rhs.jl:256-261 foldspec: plans for sigma_g = +1, -1 */
        if (!s->fold[k].on) continue;
        for (int g = 0; g < 2; g++) {
            const int sg = g == 0 ? 1 : -1;
            const int lo = s->fold[k].lo ? sg : 0, hi = s->fold[k].hi ? sg : 0;
            s->fold[k].dp[g] = plan_create_fold(&d->deriv, k + 1, s->n, halo, 0, s->h[k], lo, hi, msg, msglen);
            s->fold[k].fp[g] = plan_create_fold(&d->filt, k + 1, s->n, halo, 0, s->h[k], lo, hi, msg, msglen);
            if (!s->fold[k].dp[g] || !s->fold[k].fp[g]) { or_solver_destroy(s); return NULL; }
        }
    }
    s->Q = calloc((size_t)s->NP * s->C, sizeof(double));
    s->dQ = calloc((size_t)s->NP * s->C, sizeof(double));
    s->du = calloc((size_t)s->NP * s->C, sizeof(double));
    s->rho = fld(s); s->u = fld(s); s->v = fld(s); s->w = fld(s);
    s->p = fld(s); s->T = fld(s); s->c = fld(s); s->cp = fld(s);
    for (int k = 0; k < s->Ns; k++) { s->Y[k] = fld(s); s->D_art[k] = fld(s); }
    for (int k = 0; k < 9; k++) s->gu[k] = fld(s);
    for (int k = 0; k < 3; k++) {
        s->gT[k] = fld(s);
        for (int sp = 0; sp < s->Ns; sp++) s->gY[k][sp] = fld(s);
        for (int c = 0; c < s->C; c++) s->flux[k][c] = fld(s);
    }
    s->mu_art = fld(s); s->beta_art = fld(s); s->kappa_art = fld(s);
    s->strain_mag = fld(s); s->sensor = fld(s); s->sensor_sp = fld(s);
    s->tmp_a = fld(s); s->tmp_b = fld(s);
    s->pairbuf = fld(s); s->pairout = fld(s);
    if (s->geom) init_geometry(s, x, jac);
    return s;
}

void or_solver_destroy(or_solver *s)
{
    if (!s) return;
    for (int k = 0; k < 3; k++) { or_plan_destroy(s->dplan[k]); or_plan_destroy(s->fplan[k]); }
    for (int k = 0; k < 3; k++)
        for (int g = 0; g < 2; g++) { or_plan_destroy(s->fold[k].dp[g]); or_plan_destroy(s->fold[k].fp[g]); }
    free(s->pairbuf); free(s->pairout);
    free(s->Q); free(s->dQ); free(s->du);
    free(s->rho); free(s->u); free(s->v); free(s->w); free(s->p); free(s->T); free(s->c); free(s->cp);
    for (int k = 0; k < s->Ns; k++) { free(s->Y[k]); free(s->D_art[k]); }
    for (int k = 0; k < 9; k++) free(s->gu[k]);
    for (int k = 0; k < 3; k++) {
        free(s->gT[k]);
        for (int sp = 0; sp < s->Ns; sp++) free(s->gY[k][sp]);
        for (int c = 0; c < s->C; c++) free(s->flux[k][c]);
    }
    free(s->mu_art); free(s->beta_art); free(s->kappa_art);
    free(s->strain_mag); free(s->sensor); free(s->sensor_sp); free(s->tmp_a); free(s->tmp_b);
    free(s->inv_J); free(s->inv_r); free(s->cotr);
    for (int k = 0; k < 3; k++) { free(s->area[k]); free(s->inv_h[k]); }
    free(s);
}

void or_solver_dims(const or_solver *s, int64_t padded[3], int halo[3], int *n_cons)
{
    for (int k = 0; k < 3; k++) { padded[k] = s->pad[k]; halo[k] = s->H[k]; }
    *n_cons = s->C;
}
double *or_solver_Q(or_solver *s) { return s->Q; }
double *or_solver_dQ(or_solver *s) { return s->dQ; }
void or_solver_time(const or_solver *s, double *t, int64_t *step) { *t = s->t; *step = s->step; }
void or_solver_set_time(or_solver *s, double t, int64_t step) { s->t = t; s->step = step; }

double *or_solver_field(or_solver *s, const char *nm)
{
    int a, b;
    if (!strcmp(nm, "rho")) return s->rho;
    if (!strcmp(nm, "u")) return s->u;
    if (!strcmp(nm, "v")) return s->v;
    if (!strcmp(nm, "w")) return s->w;
    if (!strcmp(nm, "p")) return s->p;
    if (!strcmp(nm, "T")) return s->T;
    if (!strcmp(nm, "c")) return s->c;
    if (!strcmp(nm, "cp")) return s->cp;
    if (!strcmp(nm, "mu_art")) return s->mu_art;
    if (!strcmp(nm, "beta_art")) return s->beta_art;
    if (!strcmp(nm, "kappa_art")) return s->kappa_art;
    if (!strcmp(nm, "sensor")) return s->sensor;
    if (!strcmp(nm, "strain_mag")) return s->strain_mag;
    if (sscanf(nm, "grad_u[%1d%1d]", &a, &b) == 2) return s->gu[(a - 1) * 3 + (b - 1)];
    if (sscanf(nm, "grad_T[%d]", &a) == 1) return s->gT[a - 1];
    if (sscanf(nm, "flux[%1d%1d]", &a, &b) == 2) return s->flux[a - 1][b - 1];
    if (sscanf(nm, "D_art[%d]", &a) == 1) return s->D_art[a - 1];
    if (sscanf(nm, "Y[%d]", &a) == 1) return s->Y[a - 1];
    if (sscanf(nm, "grad_Y[%1d%1d]", &a, &b) == 2) return s->gY[a - 1][b - 1];
    return NULL;
}

/* problem.jl:76-90 conserved_from_prim (IdealMixture) */
void or_conserved_from_prim(const or_solver *s, const double *prim, double *q)
{
    int N = s->Ns;
    const double *Y = prim + 6;
    double Rm = 0.0, cvm = 0.0;
    for (int k = 0; k < N; k++) { Rm += Y[k] * s->Rk[k]; cvm += Y[k] * s->cvk[k]; }
    double rho = isnan(prim[0]) ? prim[4] / (Rm * prim[5]) : prim[0];
    double T = isnan(prim[5]) ? prim[4] / (rho * Rm) : prim[5];
    double ke = 0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3]);
    for (int k = 0; k < N; k++) q[k] = rho * Y[k];
    q[N] = rho * prim[1]; q[N + 1] = rho * prim[2]; q[N + 2] = rho * prim[3];
    q[N + 3] = rho * (cvm * T + ke);
}

static inline double *Qc(or_solver *s, int c) { return s->Q + (size_t)c * s->NP; }
static inline double *dQc(or_solver *s, int c) { return s->dQ + (size_t)c * s->NP; }
static inline double *duc(or_solver *s, int c) { return s->du + (size_t)c * s->NP; }

static void exchange_dim(or_solver *s, double *f, int d)
{
    if (!s->active[d]) return;
    if (s->periodic[d]) wrap_dim(f, s->pad, s->n, s->H[d], d);
    /* closed edge on one rank: neighbours are PROC_NULL, halos stay stale (halo.jl:5-9) */
}
static void exchange_halos(or_solver *s, double *f)
{
    for (int d = 0; d < 3; d++) exchange_dim(s, f, d);
}
static void exchange_state(or_solver *s)
{
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < s->C; c++) exchange_dim(s, Qc(s, c), d);
}
static void copy_interior(or_solver *s, double *dst, const double *src)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++) {
            int64_t I = pidx(s, 0, j, k);
            memcpy(dst + I, src + I, sizeof(double) * s->n[0]);
        }
}

/* physics.jl:604-653 _primitives! over the full padded arrays */
/* Nasa9Mixture (physics.jl:290-310 fits, 432-456 species functions,
 * 459-482 mixture_temperature).  Powers by multiplication. */
static const or_nasa9_interval *n9_interval(const or_nasa9_species *sp, double T)
{
    for (int i = 0; i < sp->n_intervals - 1; i++)
        if (T <= sp->iv[i].Tmax) return &sp->iv[i];
    return &sp->iv[sp->n_intervals - 1];
}
static double n9_cp_over_R(const or_nasa9_interval *v, double T)
{
    const double *a = v->a;
    const double T2 = T * T, T3 = T2 * T, T4 = T2 * T2;
    return a[0] / T2 + a[1] / T + a[2] + a[3] * T + a[4] * T2 + a[5] * T3 + a[6] * T4;
}
static double n9_h_over_R(const or_nasa9_interval *v, double T)
{
    const double *a = v->a;
    const double T2 = T * T, T3 = T2 * T, T4 = T2 * T2, T5 = T4 * T;
    return -a[0] / T + a[1] * log(T) + a[2] * T + a[3] * T2 / 2 + a[4] * T3 / 3 + a[5] * T4 / 4 + a[6] * T5 / 5 +
           v->b1;
}
static double n9_cp(const or_solver *s, int k, double T)
{
    return s->d.R[k] * n9_cp_over_R(n9_interval(&s->d.nasa9[k], T), T);
}
static double n9_h(const or_solver *s, int k, double T)
{
    return s->d.R[k] * n9_h_over_R(n9_interval(&s->d.nasa9[k], T), T);
}
static double n9_e(const or_solver *s, int k, double T) { return n9_h(s, k, T) - s->d.R[k] * T; }

static double n9_temperature(const or_solver *s, double e, const double *Y)
{
    double e0 = 0.0, cv0 = 0.0, Tg = s->d.T_guess;
    for (int k = 0; k < s->Ns; k++) {
        e0 += Y[k] * n9_e(s, k, Tg);
        cv0 += Y[k] * (n9_cp(s, k, Tg) - s->d.R[k]);
    }
    double T = Tg;
    if (cv0 > 0) {
        double x = Tg + (e - e0) / cv0;
        T = x > 1e9 ? 1e9 : (x < 1e-3 ? 1e-3 : x);
    }
    for (int it = 0; it < 30; it++) {
        double f = -e, cvm = 0.0;
        for (int k = 0; k < s->Ns; k++) {
            f += Y[k] * n9_e(s, k, T);
            cvm += Y[k] * (n9_cp(s, k, T) - s->d.R[k]);
        }
        if (!(cvm > 0)) break;
        double dl = f / cvm;
        T = jmax(T - dl, 0.25 * T);
        if (fabs(dl) <= 1e-14 * T) break;
    }
    return T;
}

/* physics.jl:515-552 _primitives! (Nasa9Mixture) at padded index I */
static void primitives_n9(or_solver *s, int64_t I, double rho)
{
    int Ns = s->Ns;
    double ri = 1 / rho, Rm = 0.0, Y[OR_MAXSP];
    for (int sp = 0; sp < Ns; sp++) {
        Y[sp] = Qc(s, sp)[I] * ri;
        s->Y[sp][I] = Y[sp];
        Rm += Qc(s, sp)[I] * ri * s->d.R[sp];
    }
    double u = Qc(s, s->im1)[I] * ri;
    double v = Qc(s, s->im2)[I] * ri;
    double w = Qc(s, s->im3)[I] * ri;
    double e = Qc(s, s->ie)[I] * ri - 0.5 * (u * u + v * v + w * w);
    double T = n9_temperature(s, e, Y);
    double cpm = 0.0;
    for (int sp = 0; sp < Ns; sp++) cpm += s->Y[sp][I] * n9_cp(s, sp, T);
    double cvm = cpm - Rm;
    double cc = (cpm / cvm) * Rm * T;
    s->rho[I] = rho; s->u[I] = u; s->v[I] = v; s->w[I] = w;
    s->p[I] = rho * Rm * T;
    s->T[I] = T;
    s->c[I] = sqrt(cc > 0.0 ? cc : 0.0);
    s->cp[I] = cpm;
}

static void primitives(or_solver *s)
{
    int Ns = s->Ns;
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < s->NP; I++) {
        double rho = 0.0;
        for (int sp = 0; sp < Ns; sp++) rho += Qc(s, sp)[I];
        if (s->d.eos_kind == 2) {
            /* physics.jl:216-250 _primitives! (StiffenedGas) */
            const double g = s->d.gamma[0], pinf = s->d.p_inf, cv = s->cvk[0], R = (g - 1) * cv;
            s->Y[0][I] = 1.0;
            s->cp[I] = g * cv;
            if (rho > 0) {
                double ri = 1 / rho;
                double u = Qc(s, s->im1)[I] * ri;
                double v = Qc(s, s->im2)[I] * ri;
                double w = Qc(s, s->im3)[I] * ri;
                double e = Qc(s, s->ie)[I] * ri - 0.5 * (u * u + v * v + w * w);
                double p = (g - 1) * rho * e - g * pinf;
                s->rho[I] = rho; s->u[I] = u; s->v[I] = v; s->w[I] = w;
                s->p[I] = p;
                s->T[I] = jmax((p + pinf) / (rho * R), 1e-300);
                s->c[I] = sqrt(jmax(g * (p + pinf) * ri, 0.0));
            } else {
                s->rho[I] = 1; s->u[I] = 0; s->v[I] = 0; s->w[I] = 0;
                s->p[I] = 1; s->T[I] = 1; s->c[I] = 1;
            }
        } else if (rho > 0 && s->d.eos_kind == 1) {
            primitives_n9(s, I, rho);
        } else if (rho > 0) {
            double ri = 1 / rho;
            double Rm = 0.0, cvm = 0.0;
            for (int sp = 0; sp < Ns; sp++) {
                double q = Qc(s, sp)[I];
                Rm += q * s->Rk[sp];
                cvm += q * s->cvk[sp];
                s->Y[sp][I] = q * ri;
            }
            Rm *= ri; cvm *= ri;
            double u = Qc(s, s->im1)[I] * ri;
            double v = Qc(s, s->im2)[I] * ri;
            double w = Qc(s, s->im3)[I] * ri;
            double e = Qc(s, s->ie)[I] * ri - 0.5 * (u * u + v * v + w * w);
            double T = jmax(e / cvm, 1e-300);
            double p = rho * Rm * T;
            double gm = 1 + Rm / cvm;
            s->rho[I] = rho; s->u[I] = u; s->v[I] = v; s->w[I] = w;
            s->p[I] = p; s->T[I] = T;
            s->c[I] = sqrt(gm * Rm * T);
            s->cp[I] = cvm + Rm;
        } else {
            s->rho[I] = 1; s->u[I] = 0; s->v[I] = 0; s->w[I] = 0;
            s->p[I] = 1; s->T[I] = 1; s->c[I] = 1; s->cp[I] = 1;
            for (int sp = 0; sp < Ns; sp++) s->Y[sp][I] = sp == 0 ? 1.0 : 0.0;
        }
    }
}

/* folds.jl:79-105 fold_fill!: halo layers beyond a folded end of dimension
 * d mirror the interior layers with sign sf (half-offset: ghost q <-> interior q) */
static void fold_fill(or_solver *s, double *f, int d, int sf)
{
    const int H = s->H[d];
    const int64_t n = s->n[d];
    const int64_t st = d == 0 ? 1 : d == 1 ? s->pad[0] : s->pad[0] * s->pad[1];
    const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
    for (int64_t b = 0; b < s->pad[o2]; b++)
        for (int64_t a = 0; a < s->pad[o1]; a++) {
            int64_t ii[3] = {0, 0, 0};
            ii[o1] = a;
            ii[o2] = b;
            double *base = f + ii[0] + s->pad[0] * (ii[1] + s->pad[1] * ii[2]);
            for (int q = 1; q <= H; q++) {
                if (s->fold[d].lo) base[(H - q) * st] = (double)sf * base[(H + q - 1) * st];
                if (s->fold[d].hi) base[(H + n + q - 1) * st] = (double)sf * base[(H + n - q) * st];
            }
        }
}

/* folds.jl:118-151 _pair_index: the antipodal partner of interior point
 * (i, j, k) (pdim shifted by half its extent, revdim reversed) */
static void pair_ijk(const or_solver *s, int d, const int64_t *ii, int64_t *out)
{
    out[0] = ii[0]; out[1] = ii[1]; out[2] = ii[2];
    const int pd = s->fold[d].pdim, rd = s->fold[d].revdim;
    if (pd >= 0) {
        const int64_t half = s->n[pd] / 2;
        out[pd] = ii[pd] < half ? ii[pd] + half : ii[pd] - half;
    }
    if (rd >= 0) out[rd] = s->n[rd] - 1 - ii[rd];
}
static int fold_paired(const or_solver *s, int d) { return s->fold[d].pdim >= 0 || s->fold[d].revdim >= 0; }
static int pair_sd(const or_solver *s, int d) { return s->fold[d].pdim >= 0 ? s->fold[d].pdim : s->fold[d].revdim; }

/* folds.jl:160-196 pair_forward! (one rank): e = (f(x) + sf f(Mx)) / 2 at the
 * canonical slot x (lower half of the mapping dimension), o = (f(x) - sf
 * f(Mx)) / 2 at the partner slot */
static void pair_forward(or_solver *s, int d, double *w, const double *f, int sf)
{
    const int sd = pair_sd(s, d);
    const int64_t half = s->n[sd] / 2;
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                const int64_t ii[3] = {i, j, k};
                if (ii[sd] >= half) continue;
                int64_t mm[3];
                pair_ijk(s, d, ii, mm);
                const int64_t I = pidx(s, i, j, k), Mx = pidx(s, mm[0], mm[1], mm[2]);
                const double a = f[I], b = (double)sf * f[Mx];
                w[I] = 0.5 * (a + b);
                w[Mx] = 0.5 * (a - b);
            }
}

/* folds.jl:206-246 pair_backward! (one rank): out(x) = Re + Ro,
 * out(Mx) = sf (Re - Ro) */
static void pair_backward(or_solver *s, int d, double *out, int sf)
{
    const int sd = pair_sd(s, d);
    const int64_t half = s->n[sd] / 2;
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                const int64_t ii[3] = {i, j, k};
                if (ii[sd] >= half) continue;
                int64_t mm[3];
                pair_ijk(s, d, ii, mm);
                const int64_t I = pidx(s, i, j, k), Mx = pidx(s, mm[0], mm[1], mm[2]);
                const double Re = out[I], Ro = out[Mx];
                out[I] = Re + Ro;
                out[Mx] = (double)sf * (Re - Ro);
            }
}

/* folds.jl:276-330 fold_apply! (one rank).  Self-paired: mirror fill and one
 * parity plan.  Paired: the e combo folds with sf and the o combo always as
 * odd; both plans run and the mapping dimension's upper half takes the odd
 * result, then the inverse butterfly. */
static void fold_apply(or_solver *s, double *out, double *f, int d, int sf, int isfilter)
{
#define FPLAN(sig) (isfilter ? s->fold[d].fp[(sig) > 0 ? 0 : 1] : s->fold[d].dp[-(sig) > 0 ? 0 : 1])
    if (!fold_paired(s, d)) {
        fold_fill(s, f, d, sf);
        or_apply_along(FPLAN(sf), out, f);
        return;
    }
    double *w = s->pairbuf;
    pair_forward(s, d, w, f, sf);
    fold_fill(s, w, d, sf);
    or_apply_along(FPLAN(sf), out, w);
    fold_fill(s, w, d, -1);
    or_apply_along(FPLAN(-1), s->pairout, w);
    const int sd = pair_sd(s, d);
    const int64_t half = s->n[sd] / 2;
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                const int64_t ii[3] = {i, j, k};
                if (ii[sd] >= half) out[pidx(s, i, j, k)] = s->pairout[pidx(s, i, j, k)];
            }
    pair_backward(s, d, out, sf);
#undef FPLAN
}

/* rhs.jl:452-475 deriv_along! / filt_along!: sf is the field's antipodal sign
 * for the fold of dimension d (ignored where there is none) */
static void deriv_along_p(or_solver *s, double *out, double *f, int d, int sf)
{
    if (s->fold[d].on) {
        fold_apply(s, out, f, d, sf, 0);
        return;
    }
    or_apply_along(s->dplan[d], out, f);
    /* _scale_grad! (rhs.jl:477-483) follows as scale_grad() for the gradients;
     * it is the identity on unstretched dimensions. */
}
static void filt_along_p(or_solver *s, double *out, double *f, int d, int sf)
{
    if (s->fold[d].on) {
        fold_apply(s, out, f, d, sf, 1);
        return;
    }
    or_apply_along(s->fplan[d], out, f);
}
#define filt_along(s, out, f, d) filt_along_p((s), (out), (f), (d), 1)

/* equations.jl:56-66: parity of velocity component j and of conserved c
 * across the fold of dimension d (axis: u_r, u_theta odd) */
static int vel_parity(const or_solver *s, int d, int j)
{
    return s->fold[d].on ? s->fold[d].sv[j] : 1;
}
static int cons_parity(const or_solver *s, int d, int c)
{
    if (c == s->im1) return vel_parity(s, d, 0);
    if (c == s->im2) return vel_parity(s, d, 1);
    if (c == s->im3) return vel_parity(s, d, 2);
    return 1;
}

/* metric.jl:177-219: curvature corrections added to the scaled physical
 * velocity gradients (grad_u[d][j] at gu[d*3 + j]) */
static void metric_correct_gradients(or_solver *s)
{
    if (s->d.metric == 0) return;
    double **g = s->gu;
    FOR_INTERIOR(s, I) {
        const double ir = s->inv_r[I];
        if (s->d.metric == 1) {
            g[3][I] -= s->v[I] * ir;
            g[4][I] += s->u[I] * ir;
        } else {
            const double ctr = s->cotr[I];
            const double ur = s->u[I], ut = s->v[I], up = s->w[I];
            g[3][I] -= ut * ir;
            g[4][I] += ur * ir;
            g[6][I] -= up * ir;
            g[7][I] -= up * ctr;
            g[8][I] += ur * ir + ut * ctr;
        }
    }
}

/* metric.jl:228-237 _Pi: the momentum flux tensor from the stored
 * primitives and gradients (a, b zero-based) */
static double Pi_ab(const or_solver *s, int64_t I, int a, int b)
{
    double *const *g = s->gu;
    const double mu = s->d.mu0 + s->mu_art[I];
    const double be = s->beta_art[I];
    const double divu = g[0][I] + g[4][I] + g[8][I];
    const double tau = mu * (g[a * 3 + b][I] + g[b * 3 + a][I]) + (a == b ? (be - 2 * mu / 3) * divu : 0.0);
    const double uv[3] = {s->u[I], s->v[I], s->w[I]};
    return s->rho[I] * uv[a] * uv[b] + (a == b ? s->p[I] : 0.0) - tau;
}

/* metric.jl:239-269 add_metric_sources! */
static void add_metric_sources(or_solver *s)
{
    if (s->d.metric == 0) return;
    double *m1 = dQc(s, s->im1), *m2 = dQc(s, s->im2), *m3 = dQc(s, s->im3);
    FOR_INTERIOR(s, I) {
        const double ir = s->inv_r[I];
        if (s->d.metric == 1) {
            m1[I] += ir * Pi_ab(s, I, 1, 1);
            m2[I] -= ir * Pi_ab(s, I, 1, 0);
        } else {
            const double ctr = s->cotr[I];
            const double Ptt = Pi_ab(s, I, 1, 1), Ppp = Pi_ab(s, I, 2, 2);
            m1[I] += ir * (Ptt + Ppp);
            m2[I] += -ir * Pi_ab(s, I, 1, 0) + ctr * Ppp;
            m3[I] += -ir * Pi_ab(s, I, 2, 0) - ctr * Pi_ab(s, I, 2, 1);
        }
    }
}

/* rhs.jl:591-606 */
static void primitives_and_gradients(or_solver *s, int primitives_current)
{
    if (!primitives_current) { exchange_state(s); primitives(s); }
    double *vel[3] = {s->u, s->v, s->w};
    for (int jj = 0; jj < 3; jj++)
        for (int d = 0; d < 3; d++) {
            if (s->active[d]) {
                deriv_along_p(s, s->gu[d * 3 + jj], vel[jj], d, vel_parity(s, d, jj));
                scale_grad(s, s->gu[d * 3 + jj], d);
            } else {
                memset(s->gu[d * 3 + jj], 0, sizeof(double) * s->NP);
            }
        }
    metric_correct_gradients(s);
}

/* artificial.jl:67-94 delta4_sum! */
static void delta4_sum(or_solver *s, double *out, const double *f, int wpow, int accumulate)
{
    static const double D4[5] = {1.0, -4.0, 6.0, -4.0, 1.0};
    if (!accumulate) memset(out, 0, sizeof(double) * s->NP);
    int64_t st[3] = {1, s->pad[0], s->pad[0] * s->pad[1]};
    for (int d = 0; d < 3; d++) {
        if (!s->active[d]) continue;
        double wd = wpow == 2 ? s->h[d] * s->h[d] : s->h[d];
        int64_t nd = s->n[d];
        int64_t lomin = s->periodic[d] ? -2 : 0;   /* 0-based: 1 -> 0, -1 -> -2 */
        int64_t himax = s->periodic[d] ? nd + 1 : nd - 1;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < s->n[2]; k++)
            for (int64_t j = 0; j < s->n[1]; j++)
                for (int64_t i = 0; i < s->n[0]; i++) {
                    int64_t I = pidx(s, i, j, k);
                    int64_t il = d == 0 ? i : d == 1 ? j : k;
                    double acc = 0.0;
                    for (int m = -2; m <= 2; m++) {
                        int64_t ilm = il + m;
                        if (ilm < lomin) ilm = lomin;
                        if (ilm > himax) ilm = himax;
                        acc += D4[m + 2] * f[I + (ilm - il) * st[d]];
                    }
                    out[I] += wd * fabs(acc);
                }
    }
}

/* artificial.jl:98-111 smooth! */
static void smooth(or_solver *s, double *f)
{
    for (int d = 0; d < 3; d++) {
        if (!s->active[d]) continue;
        exchange_dim(s, f, d);
        filt_along(s, s->tmp_a, f, d);
        copy_interior(s, f, s->tmp_a);
    }
}

#define DUCROS_EPS 1e-32
/* artificial.jl:136-145 */
static inline double compression_switch(const or_solver *s, int64_t I)
{
    double div = s->gu[0][I] + s->gu[4][I] + s->gu[8][I];
    if (!(div < 0)) return 0.0;
    double w1 = s->gu[1 * 3 + 2][I] - s->gu[2 * 3 + 1][I];
    double w2 = s->gu[2 * 3 + 0][I] - s->gu[0 * 3 + 2][I];
    double w3 = s->gu[0 * 3 + 1][I] - s->gu[1 * 3 + 0][I];
    return div * div / (div * div + w1 * w1 + w2 * w2 + w3 * w3 + DUCROS_EPS);
}

/* artificial.jl:261-359 compute_artificial! */
static void compute_artificial(or_solver *s)
{
    const or_desc *d = &s->d;
    if (!d->art_enabled) return;
    double C_mu = d->C_mu, C_beta = d->C_beta, C_kappa = d->C_kappa, C_D = d->C_D;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                int64_t I = pidx(s, i, j, k);
                double ss = 0.0;
                for (int b = 0; b < 3; b++)
                    for (int a = 0; a < 3; a++) {
                        double Sab = 0.5 * (s->gu[a * 3 + b][I] + s->gu[b * 3 + a][I]);
                        ss += Sab * Sab;
                    }
                s->strain_mag[I] = sqrt(ss);
            }
    exchange_halos(s, s->strain_mag);
    delta4_sum(s, s->sensor, s->strain_mag, 2, 0);
    smooth(s, s->sensor);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                int64_t I = pidx(s, i, j, k);
                double rs = s->rho[I] * jmax(s->sensor[I], 0.0);
                s->mu_art[I] = C_mu * rs;
                s->beta_art[I] = C_beta * rs;
            }
    if (d->beta_sensor == 1) {
        /* artificial.jl:171-183 gate_beta! */
        FOR_INTERIOR(s, I) s->beta_art[I] *= compression_switch(s, I);
    } else if (d->beta_sensor == 2) {
        /* artificial.jl:216-242 dilatation_beta! */
        FOR_INTERIOR(s, I) {
            s->tmp_a[I] = s->gu[0][I] + s->gu[4][I] + s->gu[8][I];
            s->tmp_b[I] = compression_switch(s, I);
        }
        exchange_halos(s, s->tmp_a);
        delta4_sum(s, s->sensor, s->tmp_a, 2, 0);
        smooth(s, s->sensor);
        FOR_INTERIOR(s, I) s->beta_art[I] = C_beta * s->rho[I] * jmax(s->sensor[I], 0.0) * s->tmp_b[I];
    }
    /* κ* sensor over the full padded arrays (artificial.jl:316-339) */
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < s->NP; I++) {
        double rho = jmax(s->rho[I], 1e-300);
        double m1 = Qc(s, s->im1)[I], m2 = Qc(s, s->im2)[I], m3 = Qc(s, s->im3)[I];
        double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
        s->tmp_a[I] = (Qc(s, s->ie)[I] - ke) / rho;
    }
    exchange_halos(s, s->tmp_a);
    delta4_sum(s, s->sensor, s->tmp_a, 1, 0);
    smooth(s, s->sensor);
    const double eps = 2.220446049250313e-16;
    FOR_INTERIOR(s, I) {
        double scale = s->rho[I] * s->c[I] / jmax(s->T[I], eps);   /* physics.jl:155-156 */
        s->kappa_art[I] = C_kappa * scale * jmax(s->sensor[I], 0.0);
    }
    if (s->Ns > 1) {
        for (int sp = 0; sp < s->Ns; sp++) {
            delta4_sum(s, s->sensor_sp, s->Y[sp], 1, 0);
            smooth(s, s->sensor_sp);
            FOR_INTERIOR(s, I) s->D_art[sp][I] = C_D * s->c[I] * jmax(s->sensor_sp[I], 0.0);
        }
    }
}

/* rhs.jl:495-556 _assemble_fluxes! */
static void assemble_fluxes(or_solver *s)
{
    const or_desc *d = &s->d;
    double mu0 = d->mu0, Pr = d->Pr, Sc = d->Sc;
    int Ns = s->Ns;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s->n[2]; k++)
        for (int64_t j = 0; j < s->n[1]; j++)
            for (int64_t i = 0; i < s->n[0]; i++) {
                int64_t I = pidx(s, i, j, k);
                double rho = s->rho[I];
                double uv[3] = {s->u[I], s->v[I], s->w[I]};
                double p = s->p[I], Tp = s->T[I];
                double E = Qc(s, s->ie)[I];
                double mu = mu0 + s->mu_art[I];
                double beta = s->beta_art[I];
                double kappa = mu0 * s->cp[I] / Pr + s->kappa_art[I];
                double D0 = mu0 / (Sc * rho);
#define G(a, b) s->gu[((a) - 1) * 3 + ((b) - 1)][I]
                double divu = G(1, 1) + G(2, 2) + G(3, 3);
                double t11 = mu * (2 * G(1, 1) - (2.0 / 3.0) * divu) + beta * divu;
                double t22 = mu * (2 * G(2, 2) - (2.0 / 3.0) * divu) + beta * divu;
                double t33 = mu * (2 * G(3, 3) - (2.0 / 3.0) * divu) + beta * divu;
                double t12 = mu * (G(1, 2) + G(2, 1));
                double t13 = mu * (G(1, 3) + G(3, 1));
                double t23 = mu * (G(2, 3) + G(3, 2));
#undef G
                double tau[3][3] = {{t11, t12, t13}, {t12, t22, t23}, {t13, t23, t33}};
                for (int dd = 0; dd < 3; dd++) {
                    double ud = uv[dd];
                    double Vc = 0.0;
                    for (int sp = 0; sp < Ns; sp++) Vc += (D0 + s->D_art[sp][I]) * s->gY[dd][sp][I];
                    double hdiff = 0.0;
                    for (int sp = 0; sp < Ns; sp++) {
                        double Dk = D0 + s->D_art[sp][I];
                        double Jkd = rho * (-Dk * s->gY[dd][sp][I] + s->Y[sp][I] * Vc);
                        s->flux[dd][sp][I] = rho * s->Y[sp][I] * ud + Jkd;
                        /* species_enthalpy: cp_k T (physics.jl:137) or h_k(T) (physics.jl:446-450) */
                        const double hk = s->d.eos_kind == 1 ? n9_h(s, sp, Tp) : s->cpk[sp] * Tp;
                        hdiff += hk * Jkd;
                    }
                    s->flux[dd][s->im1][I] = rho * ud * uv[0] + (dd == 0 ? p : 0.0) - tau[dd][0];
                    s->flux[dd][s->im2][I] = rho * ud * uv[1] + (dd == 1 ? p : 0.0) - tau[dd][1];
                    s->flux[dd][s->im3][I] = rho * ud * uv[2] + (dd == 2 ? p : 0.0) - tau[dd][2];
                    s->flux[dd][s->ie][I] = (E + p) * ud - (uv[0] * tau[dd][0] + uv[1] * tau[dd][1] + uv[2] * tau[dd][2])
                                            - kappa * s->gT[dd][I] + hdiff;
                }
            }
}

/* physics.jl:142-152 */
static inline double eos_phi(double rho, double p, double T, double cp)
{
    double Rm = p / (rho * T);
    return cp / Rm - 1;
}
static inline double eos_dphi_dY(const or_solver *s, int k, double rho, double p, double T, double cp)
{
    double Rm = p / (rho * T);
    double cvm = cp - Rm;
    /* physics.jl:504-510: a Nasa9Mixture takes cv_k at the local temperature */
    double cvk = s->d.eos_kind == 1 ? n9_cp(s, k, T) - s->Rk[k] : s->cvk[k];
    return (cvk * Rm - s->Rk[k] * cvm) / (Rm * Rm);
}

/* wall plane of (dim d 0-based, side 0 low / 1 high): boundary.jl:150-170 */
#define FOR_PLANE(s, d, side, I)                                                     \
    for (int64_t b_ = 0; b_ < (s)->n[(d) == 2 ? 1 : 2]; b_++)                        \
        for (int64_t a_ = 0; a_ < (s)->n[(d) == 0 ? 1 : 0]; a_++)                    \
            for (int64_t I = plane_idx(s, d, side, a_, b_), once_ = 1; once_; once_ = 0)

static inline int64_t plane_idx(const or_solver *s, int d, int side, int64_t a, int64_t b)
{
    int64_t i0 = side == 0 ? 0 : s->n[d] - 1;
    if (d == 0) return pidx(s, i0, a, b);
    if (d == 1) return pidx(s, a, i0, b);
    return pidx(s, a, b, i0);
}

/* nscbc.jl:59-154 NSCBCOutflowBC correct_rhs! */
static void nscbc_outflow(or_solver *s, int d, int side)
{
    const or_nscbc *bc = &s->d.nscbc[d][side];
    deriv_along(s, s->tmp_a, s->p, d);
    int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    int act1 = s->active[t1], act2 = s->active[t2];
    if (act1) deriv_along(s, s->tmp_b, s->p, t1);
    if (act2) deriv_along(s, s->sensor_sp, s->p, t2);
    double Lref = bc->Lref > 0 ? bc->Lref : s->d.L_domain[d];
    double *vel[3] = {s->u, s->v, s->w};
    int m[3] = {s->im1, s->im2, s->im3};
    double *Gdd = s->gu[d * 3 + d];
    double sgn = side == 1 ? -1.0 : 1.0;
    FOR_PLANE(s, d, side, I) {
        double rho = s->rho[I], c = s->c[I], p = s->p[I];
        double un = vel[d][I];
        double Ma = fabs(un) / c;
        if (!(Ma < 1)) continue;
        double dpn = (s->geom ? s->inv_h[d][I] : 1.0) * s->tmp_a[I];
        double dun = Gdd[I];
        double Lcomp = side == 1 ? (un - c) * (dpn - rho * c * dun) : (un + c) * (dpn + rho * c * dun);
        double tin = 0.0;
        if (act1) {
            double ut = vel[t1][I];
            tin += ut * (s->geom ? s->inv_h[t1][I] : 1.0) * s->tmp_b[I] + rho * c * c * s->gu[t1 * 3 + t1][I] + sgn * rho * c * ut * s->gu[t1 * 3 + d][I];
        }
        if (act2) {
            double ut = vel[t2][I];
            tin += ut * (s->geom ? s->inv_h[t2][I] : 1.0) * s->sensor_sp[I] + rho * c * c * s->gu[t2 * 3 + t2][I] + sgn * rho * c * ut * s->gu[t2 * 3 + d][I];
        }
        double bt = bc->beta_t < 0 ? Ma : bc->beta_t;
        double K = bc->sigma * (1 - Ma * Ma) * c / Lref;
        double dL = K * (p - bc->pinf) - bt * tin - Lcomp;
        double dd1 = dL / (2 * c * c);
        double dd2 = dL / 2;
        double dd3 = sgn * dL / (2 * rho * c);
        double phi = s->d.eos_kind == 2 ? 1 / (s->d.gamma[0] - 1) : eos_phi(rho, p, s->T[I], s->cp[I]);
        double u1 = s->u[I], u2 = s->v[I], u3 = s->w[I];
        double ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
        for (int k = 0; k < s->Ns; k++) dQc(s, k)[I] -= s->Y[k][I] * dd1;
        double uv[3] = {u1, u2, u3};
        for (int a = 0; a < 3; a++) dQc(s, m[a])[I] -= uv[a] * dd1 + (a == d ? rho * dd3 : 0.0);
        dQc(s, s->ie)[I] -= ke * dd1 + phi * dd2 + rho * un * dd3;
    }
}

/* nscbc.jl:220-297 NSCBCInflowBC correct_rhs! (constant targets) */
static void nscbc_inflow(or_solver *s, int d, int side)
{
    const or_nscbc *bc = &s->d.nscbc[d][side];
    deriv_along(s, s->tmp_a, s->p, d);
    deriv_along(s, s->tmp_b, s->rho, d);
    double Lref = bc->Lref > 0 ? bc->Lref : s->d.L_domain[d];
    double *vel[3] = {s->u, s->v, s->w};
    int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    int m[3] = {s->im1, s->im2, s->im3};
    double *Gdd = s->gu[d * 3 + d];
    int lowface = side == 0;
    FOR_PLANE(s, d, side, I) {
        double rho = s->rho[I], c = s->c[I], p = s->p[I], Tp = s->T[I];
        double un = vel[d][I];
        double Ma = fabs(un) / c;
        if (!(Ma < 1)) continue;
        double ih = s->geom ? s->inv_h[d][I] : 1.0;
        double dpn = ih * s->tmp_a[I];
        double drn = ih * s->tmp_b[I];
        double dun = Gdd[I];
        double K = c / Lref;
        double L1c = (un - c) * (dpn - rho * c * dun);
        double L5c = (un + c) * (dpn + rho * c * dun);
        double L2c = un * (c * c * drn - dpn);
        double rel_ac = bc->eta_u * rho * c * c * (1 - Ma * Ma) / Lref * (un - bc->u[d]);
        double dL1 = lowface ? 0.0 : (-rel_ac - L1c);
        double dL5 = lowface ? (rel_ac - L5c) : 0.0;
        double c3 = c * c * c;   /* c^3 (Julia literal power) */
        double dL2 = bc->eta_T * rho * c3 * (bc->T - Tp) / (Lref * Tp) - L2c;
        double dd1 = dL2 / (c * c) + (dL5 + dL1) / (2 * c * c);
        double dd2 = (dL5 + dL1) / 2;
        double dd3 = (dL5 - dL1) / (2 * rho * c);
        double dd4 = bc->eta_t * K * (vel[t1][I] - bc->u[t1]) - un * s->gu[d * 3 + t1][I];
        double dd5 = bc->eta_t * K * (vel[t2][I] - bc->u[t2]) - un * s->gu[d * 3 + t2][I];
        double cpm = s->cp[I];
        double phi = s->d.eos_kind == 2 ? 1 / (s->d.gamma[0] - 1) : eos_phi(rho, p, Tp, cpm);
        double u1 = s->u[I], u2 = s->v[I], u3 = s->w[I];
        double ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
        double uv[3] = {u1, u2, u3};
        double sphiY = 0.0;
        for (int k = 0; k < s->Ns; k++) {
            double dLs = bc->eta_Y * K * (s->Y[k][I] - bc->Y[k]) - un * s->gY[d][k][I];
            dQc(s, k)[I] -= s->Y[k][I] * dd1 + rho * dLs;
            if (s->d.eos_kind != 2) sphiY += eos_dphi_dY(s, k, rho, p, Tp, cpm) * dLs;   /* physics.jl:212 */
        }
        for (int a = 0; a < 3; a++) {
            double extra = a == d ? rho * dd3 : a == t1 ? rho * dd4 : rho * dd5;
            dQc(s, m[a])[I] -= uv[a] * dd1 + extra;
        }
        dQc(s, s->ie)[I] -= ke * dd1 + phi * dd2 + p * sphiY + rho * (un * dd3 + uv[t1] * dd4 + uv[t2] * dd5);
    }
}

/* rhs.jl:624-694 compute_rhs! (Cartesian, unit geometry) */
void or_compute_rhs(or_solver *s, int primitives_current)
{
    primitives_and_gradients(s, primitives_current);
    compute_artificial(s);
    for (int d = 0; d < 3; d++) {
        if (!s->active[d]) continue;
        deriv_along(s, s->gT[d], s->T, d);
        scale_grad(s, s->gT[d], d);
        for (int sp = 0; sp < s->Ns; sp++) {
            deriv_along(s, s->gY[d][sp], s->Y[sp], d);
            scale_grad(s, s->gY[d][sp], d);
        }
    }
    assemble_fluxes(s);
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < s->C; c++) exchange_dim(s, s->flux[d][c], d);
    for (int c = 0; c < s->C; c++) {
        double *dq = dQc(s, c);
        FOR_INTERIOR(s, I) dq[I] = 0;
        for (int d = 0; d < 3; d++) {
            if (!s->active[d]) continue;
            if (!s->geom) {   /* rhs.jl:649-663: unit geometry, A_d = inv_J = 1 */
                deriv_along(s, s->tmp_a, s->flux[d][c], d);
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < s->n[2]; k++)
                    for (int64_t j = 0; j < s->n[1]; j++)
                        for (int64_t i = 0; i < s->n[0]; i++) {
                            int64_t I = pidx(s, i, j, k);
                            dq[I] -= s->tmp_a[I];
                        }
                continue;
            }
            /* rhs.jl:672-684: tmp_b = A_d F_d (A_d = J / h_d), dQ -= inv_J D_d tmp_b */
            const double *F = s->flux[d][c], *Ad = s->area[d];
#pragma omp parallel for schedule(static)
            for (int64_t I = 0; I < s->NP; I++) s->tmp_b[I] = Ad[I] * F[I];
            /* equations.jl:62-66 flux_parities: conserved x normal-velocity x area parity */
            deriv_along_p(s, s->tmp_a, s->tmp_b, d,
                          cons_parity(s, d, c) * vel_parity(s, d, d) * (s->fold[d].on ? s->fold[d].area : 1));
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < s->n[2]; k++)
                for (int64_t j = 0; j < s->n[1]; j++)
                    for (int64_t i = 0; i < s->n[0]; i++) {
                        int64_t I = pidx(s, i, j, k);
                        dq[I] -= s->inv_J[I] * s->tmp_a[I];
                    }
        }
    }
    add_metric_sources(s);   /* rhs.jl:687 */
    for (int d = 0; d < 3; d++)
        for (int side = 0; side < 2; side++) {
            if (!s->active[d]) continue;
            int bc = s->d.bc[d][side];
            if (bc == OR_BC_NSCBC_OUT) nscbc_outflow(s, d, side);
            else if (bc == OR_BC_NSCBC_IN) nscbc_inflow(s, d, side);
        }
    /* rhs.jl:692 add_sources!: sources.jl:44-65 ConstantBodyForce, in order */
    for (int i = 0; i < s->d.n_body_forces; i++) {
        const double g1 = s->d.body_force[i][0], g2 = s->d.body_force[i][1], g3 = s->d.body_force[i][2];
        double *d1 = dQc(s, s->im1), *d2 = dQc(s, s->im2), *d3 = dQc(s, s->im3), *de = dQc(s, s->ie);
        FOR_INTERIOR(s, I) {
            double rho = 0.0;
            for (int sp = 0; sp < s->Ns; sp++) rho += Qc(s, sp)[I];
            d1[I] += rho * g1;
            d2[I] += rho * g2;
            d3[I] += rho * g3;
            de[I] += Qc(s, s->im1)[I] * g1 + Qc(s, s->im2)[I] * g2 + Qc(s, s->im3)[I] * g3;
        }
    }
}

/* boundary.jl:232-243 SlipWallBC, problem.jl:196-207 DirichletBC (constant) */
void or_apply_bcs(or_solver *s)
{
    for (int d = 0; d < 3; d++)
        for (int side = 0; side < 2; side++) {
            if (!s->active[d]) continue;
            int bc = s->d.bc[d][side];
            if (bc == OR_BC_SLIPWALL) {
                int mc = s->im1 + d;
                FOR_PLANE(s, d, side, I) {
                    double rho = 0.0;
                    for (int k = 0; k < s->Ns; k++) rho += Qc(s, k)[I];
                    double mn = Qc(s, mc)[I];
                    Qc(s, s->ie)[I] -= 0.5 * mn * mn / rho;
                    Qc(s, mc)[I] = 0;
                }
            } else if (bc == OR_BC_DIRICHLET) {
                FOR_PLANE(s, d, side, I) {
                    for (int c = 0; c < s->C; c++) Qc(s, c)[I] = s->d.bc_state[d][side][c];
                }
            } else if (bc == OR_BC_NOSLIP) {   /* boundary.jl:245-262 */
                const double Tw = s->d.Twall[d][side];
                FOR_PLANE(s, d, side, I) {
                    double rho = 0.0;
                    for (int k = 0; k < s->Ns; k++) rho += Qc(s, k)[I];
                    double m1 = Qc(s, s->im1)[I], m2 = Qc(s, s->im2)[I], m3 = Qc(s, s->im3)[I];
                    double ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
                    Qc(s, s->ie)[I] -= ke;
                    Qc(s, s->im1)[I] = 0;
                    Qc(s, s->im2)[I] = 0;
                    Qc(s, s->im3)[I] = 0;
                    if (Tw > 0) {   /* boundary.jl:203-221 wall_internal_energy */
                        double re = 0.0;
                        if (s->d.eos_kind == 2) {   /* boundary.jl:213-214 */
                            re = Qc(s, 0)[I] * s->cvk[0] * Tw + s->d.p_inf;
                        } else if (s->d.eos_kind == 1) {
                            for (int k = 0; k < s->Ns; k++) re += Qc(s, k)[I] * n9_e(s, k, Tw);
                        } else {
                            for (int k = 0; k < s->Ns; k++) re += Qc(s, k)[I] * s->cvk[k];
                            re *= Tw;
                        }
                        Qc(s, s->ie)[I] = re;
                    }
                }
            } else if (bc == OR_BC_EXTRAP) {   /* boundary.jl:264-276 */
                const int64_t st = d == 0 ? 1 : d == 1 ? s->pad[0] : s->pad[0] * s->pad[1];
                FOR_PLANE(s, d, side, I) {
                    const int64_t J = side == 0 ? I + st : I - st;
                    for (int c = 0; c < s->C; c++) Qc(s, c)[I] = Qc(s, c)[J];
                }
            }
        }
}

int or_solver_switch_bc(or_solver *s, int dim, int side)
{
    if (dim < 0 || dim > 2 || side < 0 || side > 1 || !s->d.switchable[dim][side]) return -1;
    s->d.bc[dim][side] = s->d.bc_after[dim][side];   /* idempotent */
    return 0;
}

/* timestep.jl:125-175 max_rate */
void or_max_rate(or_solver *s, double *rate_out, double *rho_min_out)
{
    exchange_state(s);
    primitives(s);
    const or_desc *d = &s->d;
    double rate = 0.0, rho_min = INFINITY;
    FOR_INTERIOR(s, I) {
        double rQ = 0.0;
        for (int sp = 0; sp < s->Ns; sp++) rQ += Qc(s, sp)[I];
        rho_min = jmin(rho_min, rQ);
        double rho = s->rho[I];
        double ri = 1 / rho;
        double c = s->c[I], cp = s->cp[I];
        double uv[3] = {s->u[I], s->v[I], s->w[I]};
        double acc = 0.0, dsum = 0.0;
        for (int dd = 0; dd < 3; dd++) {
            if (!s->active[dd]) continue;
            /* inv_h[d] / h[d] */
            const double idx = s->geom ? s->inv_h[dd][I] / s->h[dd] : 1.0 / s->h[dd];
            acc += (fabs(uv[dd]) + c) * idx;
            dsum += idx * idx;
        }
        /* timestep.jl:186-197 curvature_rate: collapsed angular dimensions */
        if (s->d.metric == 1 && !s->active[1]) acc += fabs(uv[1]) * s->inv_r[I];
        if (s->d.metric == 2) {
            double a = 0.0;
            if (!s->active[1]) a += fabs(uv[1]) * s->inv_r[I];
            if (!s->active[2]) a += fabs(uv[2]) * (s->inv_r[I] + fabs(s->cotr[I]));
            acc += a;
        }
        double Dmax = d->mu0 / (d->Sc * rho);
        for (int sp = 0; sp < s->Ns; sp++) Dmax = jmax(Dmax, d->mu0 / (d->Sc * rho) + s->D_art[sp][I]);
        double nu = (d->mu0 + s->mu_art[I] + s->beta_art[I]) * ri +
                    (d->mu0 * cp / d->Pr + s->kappa_art[I]) * ri / cp + Dmax;
        acc += 2 * nu * dsum;
        rate = jmax(rate, acc);
    }
    *rate_out = rate;
    *rho_min_out = rho_min;
}

/* timestep.jl:72-99 step! */
void or_step(or_solver *s, double dt, int prepared)
{
    for (int st = 0; st < RKN; st++) {
        s->tstage = s->t + RKC[st] * dt;
        int first_prepared = prepared && st == 0;
        if (!first_prepared) or_apply_bcs(s);
        or_compute_rhs(s, first_prepared);
        double A = RKA[st], B = RKB[st];
        for (int c = 0; c < s->C; c++) {
            double *du = duc(s, c), *dq = dQc(s, c), *q = Qc(s, c);
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < s->n[2]; k++)
                for (int64_t j = 0; j < s->n[1]; j++)
                    for (int64_t i = 0; i < s->n[0]; i++) {
                        int64_t I = pidx(s, i, j, k);
                        double vv = A * du[I] + dt * dq[I];
                        du[I] = vv;
                        q[I] += B * vv;
                    }
        }
    }
    s->tstage = s->t + dt;
    or_apply_bcs(s);
}

/* timestep.jl:414-426 filter_state! */
void or_filter_state(or_solver *s)
{
    for (int d = 0; d < 3; d++) {
        if (!s->active[d]) continue;
        for (int c = 0; c < s->C; c++) exchange_dim(s, Qc(s, c), d);
        for (int c = 0; c < s->C; c++) {
            filt_along_p(s, s->tmp_a, Qc(s, c), d, cons_parity(s, d, c));
            copy_interior(s, Qc(s, c), s->tmp_a);
        }
    }
}

/* stepcontrol.jl:184-204 check_step */
static int check_step(const or_desc *d, double dt, double rho_min, double dt_seen)
{
    if (!isfinite(dt)) return 1;
    if (!(rho_min > 0)) return 2;
    if (!(dt > 5.391247e-44)) return 3;
    if (!(dt >= d->dt_min)) return 4;
    if (d->dt_min_ratio > 0 && dt_seen > 0)
        if (!(dt >= d->dt_min_ratio * dt_seen)) return 5;
    return 0;
}

/* timestep.jl:277-400 run!: fresh Workspace; StepControl recovery
 * (timestep.jl:286-351: savepoint of Q, t, step on a state that has just
 * passed its check, never at or before a failed step; a failure rolls back,
 * multiplies cfl by cfl_backoff, clears the rate history and dt_seen) and the
 * soft landing on the next scheduled instant (timestep.jl:352-383) */
int or_run_ctl(or_solver *s, double tfinal, int64_t nmax, const or_run_hooks *hooks, double *dts, int64_t ndts,
               int *retries_used)
{
    memset(s->dQ, 0, sizeof(double) * s->NP * s->C);
    memset(s->du, 0, sizeof(double) * s->NP * s->C);
    const int recover = s->d.retries > 0;
    const size_t qn = (size_t)s->NP * s->C;
    double *save_q = NULL, save_t = s->t;
    int64_t save_step = s->step;
    if (recover) {
        save_q = malloc(qn * sizeof(double));
        memcpy(save_q, s->Q, qn * sizeof(double));
    }
    const int landing = s->d.landing_steps > 0 ? s->d.landing_steps : 2;
    int attempts = 0, rc = 0;
    int64_t guard_step = -1;
    double dt_seen = 0.0;
    int64_t taken = 0;
    while (s->t < tfinal && s->step < nmax) {
        s->tstage = s->t;
        or_apply_bcs(s);
        double rate, rho_min;
        or_max_rate(s, &rate, &rho_min);
        /* timestep.jl:29-42 predicted_dt */
        double r = rate;
        if (s->d.predict > 0 && s->rate_prev > 0) r = jmax(r, r + s->d.predict * (r - s->rate_prev));
        double dt = s->d.cfl / r;
        if (s->d.max_growth > 0 && s->dt_prev > 0) dt = jmin(dt, s->d.max_growth * s->dt_prev);
        int fail = check_step(&s->d, dt, rho_min, dt_seen);
        if (fail) {
            if (!recover || attempts >= s->d.retries) {
                rc = fail;
                break;
            }
            attempts++;
            memcpy(s->Q, save_q, qn * sizeof(double));
            guard_step = s->step;
            s->t = save_t;
            s->step = save_step;
            s->d.cfl *= s->d.cfl_backoff;
            s->dt_prev = 0.0;
            s->rate_prev = 0.0;
            dt_seen = 0.0;
            if (hooks && hooks->rewind) hooks->rewind(hooks->user, save_t, save_step);
            continue;
        }
        if (recover && s->d.savepoint_interval > 0 && s->step > guard_step &&
            s->step % s->d.savepoint_interval == 0) {
            memcpy(save_q, s->Q, qn * sizeof(double));
            save_t = s->t;
            save_step = s->step;
        }
        dt_seen = jmax(dt_seen, dt);
        dt = jmin(dt, tfinal - s->t);
        if (hooks && hooks->next_time) {
            double next = hooks->next_time(hooks->user, s->t, s->step);
            double gap = next - s->t;
            if (next <= tfinal && gap > 0 && gap < dt * landing) dt = gap / ceil(gap / dt);
        }
        or_step(s, dt, 1);
        if (dts && taken < ndts) dts[taken] = dt;
        taken++;
        s->t += dt;
        s->step += 1;
        s->dt_prev = dt;
        s->rate_prev = rate;
        if (s->d.filter_interval > 0 && s->step % s->d.filter_interval == 0) or_filter_state(s);
        if (hooks && hooks->after_step && hooks->after_step(hooks->user, s->t, s->step, dt)) break;
    }
    free(save_q);
    if (retries_used) *retries_used = attempts;
    return rc;
}

int or_run(or_solver *s, double tfinal, int64_t nmax, double *dts, int64_t ndts)
{
    return or_run_ctl(s, tfinal, nmax, NULL, dts, ndts, NULL);
}

double or_solver_cfl(const or_solver *s) { return s->d.cfl; }
void or_solver_set_cfl(or_solver *s, double cfl) { s->d.cfl = cfl; }

/* diagnostics.jl:41-80 quad_weight / volume_integral (Cartesian, inv_J == 1) */
static inline double quad_weight(const or_solver *s, int d, int64_t i)
{
    if (!(s->active[d] && !s->periodic[d])) return 1.0;
    if (i == 0) return s->fold[d].lo ? 1.0 : 0.5;   /* a folded edge takes a full weight */
    if (i == s->n[d] - 1) return s->fold[d].hi ? 1.0 : 0.5;
    return 1.0;
}

double or_volume_integral(const or_solver *s, const double *f)
{
    double dV = 1.0;
    for (int d = 0; d < 3; d++) dV *= s->active[d] ? s->h[d] : 1.0;
    double acc = 0.0;
    for (int64_t k = 0; k < s->n[2]; k++) {
        double wk = quad_weight(s, 2, k);
        for (int64_t j = 0; j < s->n[1]; j++) {
            double wj = wk * quad_weight(s, 1, j);
            for (int64_t i = 0; i < s->n[0]; i++) {
                int64_t I = pidx(s, i, j, k);
                const double inv_J = s->geom ? s->inv_J[I] : 1.0;
                acc += wj * quad_weight(s, 0, i) * f[I] / inv_J;
            }
        }
    }
    return acc * dV;
}

/* Kinetic energy as examples/taylor_green.jl:51-60 (divides by the mixture
 * density; for one species that is Q[:,1]); enstrophy ∫½|ω|² with ω from the
 * compact grad_u of the current state (io.jl:332-340). */
void or_diagnostics(or_solver *s, double *ke, double *ens)
{
    double *f = s->tmp_b;
    memset(f, 0, sizeof(double) * s->NP);
    FOR_INTERIOR(s, I) {
        double rho = 0.0;
        for (int sp = 0; sp < s->Ns; sp++) rho += Qc(s, sp)[I];
        double m1 = Qc(s, s->im1)[I], m2 = Qc(s, s->im2)[I], m3 = Qc(s, s->im3)[I];
        f[I] = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
    }
    *ke = or_volume_integral(s, f);
    primitives_and_gradients(s, 0);
    FOR_INTERIOR(s, I) {
        double w1 = s->gu[1 * 3 + 2][I] - s->gu[2 * 3 + 1][I];
        double w2 = s->gu[2 * 3 + 0][I] - s->gu[0 * 3 + 2][I];
        double w3 = s->gu[0 * 3 + 1][I] - s->gu[1 * 3 + 0][I];
        f[I] = 0.5 * (w1 * w1 + w2 * w2 + w3 * w3);
    }
    *ens = or_volume_integral(s, f);
}
