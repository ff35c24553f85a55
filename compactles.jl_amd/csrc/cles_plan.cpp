// Host-side construction of the line-operator tables (no kernels here).
//
// Steps, for one (scheme, n, periodic, h):
//   1. global band matrix with closure substitution and the 1/h prescale of
//      the RHS stencils -- src/operators.jl:44-101, src/operators_banded.jl:29-125;
//   2. balanced split of the line into P = 2^p chunks;
//   3. per chunk: Thomas factors (src/tridiag.jl:31-43) or the unpivoted banded
//      LU (src/banded.jl:38-61), and the spike blocks V = T^-1 AL, W = T^-1 CR
//      (src/tridiag.jl:85-99, src/banded.jl:108-128) with chunks as "ranks";
//   4. block-PCR coefficients for the reduced system in (head, tail) unknowns
//      (src/tridiag.jl:100-124 generalised from a dense LU to a cyclic
//      reduction whose blocks are fixed by the scheme).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "cles_internal.h"

namespace cles {

namespace {

using Mat = std::vector<long double>;   // small dense B x B, row-major

Mat mat_zero(int B) { return Mat(B * B, 0.0L); }
Mat mat_eye(int B)
{
    Mat m = mat_zero(B);
    for (int i = 0; i < B; i++) m[i * B + i] = 1.0L;
    return m;
}
Mat mat_mul(const Mat &a, const Mat &b, int B)
{
    Mat c = mat_zero(B);
    for (int i = 0; i < B; i++)
        for (int k = 0; k < B; k++)
            for (int j = 0; j < B; j++) c[i * B + j] += a[i * B + k] * b[k * B + j];
    return c;
}
Mat mat_add(const Mat &a, const Mat &b)
{
    Mat c = a;
    for (size_t i = 0; i < c.size(); i++) c[i] += b[i];
    return c;
}
Mat mat_neg(const Mat &a)
{
    Mat c = a;
    for (auto &x : c) x = -x;
    return c;
}
// Gauss-Jordan with partial pivoting (long double).
Mat mat_inv(const Mat &a, int B)
{
    Mat m = a, inv = mat_eye(B);
    for (int c = 0; c < B; c++) {
        int p = c;
        for (int r = c + 1; r < B; r++)
            if (std::fabs(m[r * B + c]) > std::fabs(m[p * B + c])) p = r;
        if (m[p * B + c] == 0.0L) throw Error{CLES_ERR_INVALID, "singular reduced block in line plan"};
        if (p != c)
            for (int j = 0; j < B; j++) {
                std::swap(m[p * B + j], m[c * B + j]);
                std::swap(inv[p * B + j], inv[c * B + j]);
            }
        long double d = m[c * B + c];
        for (int j = 0; j < B; j++) { m[c * B + j] /= d; inv[c * B + j] /= d; }
        for (int r = 0; r < B; r++) {
            if (r == c) continue;
            long double f = m[r * B + c];
            if (f == 0.0L) continue;
            for (int j = 0; j < B; j++) { m[r * B + j] -= f * m[c * B + j]; inv[r * B + j] -= f * inv[c * B + j]; }
        }
    }
    return inv;
}

// Banded matrix of one line: A(i, i+s) = band[(s+q)*n + i], s in -q..q.
struct Band {
    int n, q;
    std::vector<double> a;
    double &at(int i, int s) { return a[(size_t)(s + q) * n + i]; }
    double get(int i, int s) const { return a[(size_t)(s + q) * n + i]; }
};

// Chunk-local factorisation of rows [s0, s0+len).
struct ChunkFactor {
    int s0, len, q;
    // q = 1: l, dinv, c ; q = 2: L1, L2, U0inv, U1, U2 (all length len)
    std::vector<double> f[5];
};

ChunkFactor factor_chunk(const Band &A, int s0, int len)
{
    ChunkFactor F;
    F.s0 = s0;
    F.len = len;
    F.q = A.q;
    int q = A.q;
    if (q == 1) {
        // src/tridiag.jl:31-43 on the chunk block
        for (auto &v : F.f) v.assign(len, 0.0);
        double d = A.get(s0, 0);
        F.f[1][0] = 1.0 / d;
        for (int i = 1; i < len; i++) {
            F.f[0][i] = A.get(s0 + i, -1) * F.f[1][i - 1];
            d = A.get(s0 + i, 0) - F.f[0][i] * A.get(s0 + i - 1, 1);
            F.f[1][i] = 1.0 / d;
        }
        for (int i = 0; i < len; i++) F.f[2][i] = A.get(s0 + i, 1);
        return F;
    }
    // src/banded.jl:38-61 on the chunk block (q = 2)
    std::vector<double> Wk((size_t)(2 * q + 1) * len);
    for (int r = 0; r < 2 * q + 1; r++)
        for (int i = 0; i < len; i++) Wk[(size_t)r * len + i] = A.get(s0 + i, r - q);
    auto WK = [&](int r, int i) -> double & { return Wk[(size_t)r * len + i]; };
    std::vector<double> L((size_t)q * len, 0.0);
    for (int k = 0; k < len - 1; k++) {
        double piv = WK(q, k);
        int mmax = std::min(q, len - 1 - k);
        for (int m = 1; m <= mmax; m++) {
            int i = k + m;
            double lm = WK(q - m, i) / piv;
            L[(size_t)(m - 1) * len + k] = lm;
            for (int t = 1; t <= mmax; t++) WK(q + t - m, i) -= lm * WK(q + t, k);
        }
    }
    for (auto &v : F.f) v.assign(len, 0.0);
    for (int i = 0; i < len; i++) {
        F.f[0][i] = L[i];
        F.f[1][i] = L[(size_t)len + i];
        F.f[2][i] = 1.0 / WK(q, i);
        F.f[3][i] = (i + 1 < len) ? WK(q + 1, i) : 0.0;
        F.f[4][i] = (i + 2 < len) ? WK(q + 2, i) : 0.0;
    }
    return F;
}

// Solve T x = b with the chunk factor (same operation order as the kernel).
void chunk_solve(const ChunkFactor &F, std::vector<double> &x)
{
    int len = F.len;
    if (F.q == 1) {
        for (int i = 1; i < len; i++) x[i] -= F.f[0][i] * x[i - 1];
        x[len - 1] *= F.f[1][len - 1];
        for (int i = len - 2; i >= 0; i--) x[i] = (x[i] - F.f[2][i] * x[i + 1]) * F.f[1][i];
        return;
    }
    for (int k = 0; k < len - 1; k++) {
        x[k + 1] -= F.f[0][k] * x[k];
        if (k + 2 < len) x[k + 2] -= F.f[1][k] * x[k];
    }
    for (int i = len - 1; i >= 0; i--) {
        double acc = x[i];
        if (i + 1 < len) acc -= F.f[3][i] * x[i + 1];
        if (i + 2 < len) acc -= F.f[4][i] * x[i + 2];
        x[i] = acc * F.f[2][i];
    }
}

}  // namespace

void build_line_tables(LineTables &out, const cles_scheme_desc &s, int n, int periodic, double h, int max_chunk,
                       int dim, bool chunk_major, int lo_fold, int hi_fold)
{
    const int q = s.q, M = s.M, nc = s.nclosure;
    if (q < 1 || q > 2) throw Error{CLES_ERR_UNSUPPORTED, "scheme LHS half-bandwidth q must be 1 or 2"};
    if (M < 1 || M > 4) throw Error{CLES_ERR_UNSUPPORTED, "scheme RHS half-width M must be 1..4"};
    if (nc < 0 || nc > 4) throw Error{CLES_ERR_UNSUPPORTED, "at most 4 closure rows per edge are supported"};
    for (int j = 0; j < nc; j++)
        if (s.closure_rhs_len[j] < 1 || s.closure_rhs_len[j] > 8)
            throw Error{CLES_ERR_UNSUPPORTED, "closure RHS stencils of 1..8 entries are supported"};
    // src/operators.jl:50-53, src/operators_banded.jl:369-374 (n_halo = 4)
    int nmin = std::max(2 * nc + 1, 2 * M + 1);
    if (q > 1) nmin = std::max(nmin, 2 * q + 1);
    if (n < nmin) {
        std::ostringstream os;
        os << "local extent " << n << " along dim " << dim
           << " too small for scheme (need >= " << nmin << "); use fewer ranks in this dimension";
        throw Error{CLES_ERR_TOO_SMALL, os.str()};
    }
    const bool closed = !periodic;

    // 1. band matrix (closure substitution as operators.jl:59-85 / operators_banded.jl:381-427)
    Band A{n, q, std::vector<double>((size_t)(2 * q + 1) * n, 0.0)};
    for (int i = 0; i < n; i++) {
        A.at(i, 0) = 1.0;
        for (int sd = 1; sd <= q; sd++) { A.at(i, sd) = s.lhs[sd - 1]; A.at(i, -sd) = s.lhs[sd - 1]; }
    }
    if ((lo_fold || hi_fold) && (periodic || nc < M || nc + M > 8))
        throw Error{CLES_ERR_UNSUPPORTED, "a parity fold needs a closed line and nclosure >= M with nclosure + M <= 8"};
    if (closed) {
        if (lo_fold) {
            // operators_banded.jl:55-64: ghost column r - m < 0 of row r maps onto
            // interior column m - r - 1 with sign sigma_g (q = 1: b[1] += sigma_g alpha)
            for (int r = 0; r < q; r++)
                for (int mm = 1; mm <= q; mm++)
                    if (r - mm < 0) A.at(r, (mm - r - 1) - r) += (double)lo_fold * s.lhs[mm - 1];
        } else {
            for (int j = 0; j < nc; j++)
                for (int sd = -q; sd <= q; sd++) A.at(j, sd) = s.closure_lhs[j][q + sd];
        }
        if (hi_fold) {
            // operators_banded.jl:65-73: ghost column r + m >= n maps onto 2n - 1 - (r + m)
            for (int r = n - q; r < n; r++)
                for (int mm = 1; mm <= q; mm++)
                    if (r + mm >= n) A.at(r, (2 * n - 1 - (r + mm)) - r) += (double)hi_fold * s.lhs[mm - 1];
        } else {
            for (int j = 0; j < nc; j++) {
                int r = n - 1 - j;
                for (int sd = -q; sd <= q; sd++) A.at(r, sd) = s.closure_lhs[j][q - sd];
            }
        }
        // columns outside the line do not exist at a closed edge
        for (int i = 0; i < q; i++)
            for (int sd = -q; sd < 0; sd++)
                if (i + sd < 0) A.at(i, sd) = 0.0;
        for (int i = n - q; i < n; i++)
            for (int sd = 1; sd <= q; sd++)
                if (i + sd >= n) A.at(i, sd) = 0.0;
    }

    // 2. chunks
    int P = 1;
    while ((n + P - 1) / P > max_chunk) P *= 2;
    while (P > 1 && n / P < std::max(2 * q + 1, std::max(nc, 4))) P /= 2;
    const int m = n / P, extra = n % P;
    int nsteps = 0;
    while ((1 << nsteps) < P) nsteps++;
    auto cstart = [&](int k) { return k * m + std::min(k, extra); };
    auto clen = [&](int k) { return m + (k < extra ? 1 : 0); };

    const int B = 2 * q;
    const int nrow = q == 1 ? (int)R1_N : (int)R2_N;
    std::vector<double> row((size_t)nrow * n, 0.0);
    std::vector<Mat> Lb(P), Ub(P);

    // 3. per-chunk factors and spikes
    for (int k = 0; k < P; k++) {
        int s0 = cstart(k), len = clen(k);
        ChunkFactor F = factor_chunk(A, s0, len);
        // couplings to the previous chunk's tail (AL) and the next chunk's head (CR)
        std::vector<double> AL((size_t)q * q, 0.0), CR((size_t)q * q, 0.0);
        int prev_tail0 = s0 - q, next_head0 = s0 + len;
        for (int r = 0; r < q; r++)
            for (int sd = -q; sd < 0; sd++) {
                int col = s0 + r + sd;
                if (col >= s0) continue;
                if (closed && col < 0) continue;
                double a = A.get(s0 + r, sd);
                int t = col - prev_tail0;             // 0..q-1 (mod n when periodic)
                if (t < 0 || t >= q) throw Error{CLES_ERR_INVALID, "chunk coupling outside the previous tail"};
                AL[(size_t)r * q + t] += a;
            }
        for (int r = 0; r < q; r++) {
            int gr = s0 + len - q + r;
            for (int sd = 1; sd <= q; sd++) {
                int col = gr + sd;
                if (col < s0 + len) continue;
                if (closed && col >= n) continue;
                double a = A.get(gr, sd);
                int t = col - next_head0;
                if (t < 0 || t >= q) throw Error{CLES_ERR_INVALID, "chunk coupling outside the next head"};
                CR[(size_t)r * q + t] += a;
            }
        }
        std::vector<std::vector<double>> V(q, std::vector<double>(len, 0.0)), W = V;
        for (int t = 0; t < q; t++) {
            bool anyv = false, anyw = false;
            for (int r = 0; r < q; r++) {
                V[t][r] = AL[(size_t)r * q + t];
                W[t][len - q + r] = CR[(size_t)r * q + t];
                anyv |= V[t][r] != 0.0;
                anyw |= W[t][len - q + r] != 0.0;
            }
            if (anyv) chunk_solve(F, V[t]);
            if (anyw) chunk_solve(F, W[t]);
        }
        for (int i = 0; i < len; i++) {
            int g = s0 + i;
            if (q == 1) {
                row[(size_t)R1_L * n + g] = F.f[0][i];
                row[(size_t)R1_DINV * n + g] = F.f[1][i];
                row[(size_t)R1_C * n + g] = F.f[2][i];
                row[(size_t)R1_V * n + g] = V[0][i];
                row[(size_t)R1_W * n + g] = W[0][i];
            } else {
                for (int f = 0; f < 5; f++) row[(size_t)f * n + g] = F.f[f][i];
                row[(size_t)R2_V0 * n + g] = V[0][i];
                row[(size_t)R2_V1 * n + g] = V[1][i];
                row[(size_t)R2_W0 * n + g] = W[0][i];
                row[(size_t)R2_W1 * n + g] = W[1][i];
            }
        }
        // block rows: z_k + Lb_k z_{k-1} + Ub_k z_{k+1} = E_k, z = (head[q], tail[q])
        Mat L = mat_zero(B), U = mat_zero(B);
        for (int r = 0; r < q; r++)
            for (int t = 0; t < q; t++) {
                L[(size_t)r * B + q + t] = V[t][r];
                L[(size_t)(q + r) * B + q + t] = V[t][len - q + r];
                U[(size_t)r * B + t] = W[t][r];
                U[(size_t)(q + r) * B + t] = W[t][len - q + r];
            }
        Lb[k] = L;
        Ub[k] = U;
    }

    // 4. block PCR
    std::vector<Mat> D(P, mat_eye(B)), L = Lb, U = Ub;
    std::vector<double> pcr;
    pcr.reserve((size_t)(nsteps * P * 2 + P) * B * B);
    std::vector<double> steps((size_t)nsteps * P * 2 * B * B, 0.0);
    for (int st = 0, d = 1; st < nsteps; st++, d *= 2) {
        std::vector<Mat> D2(P), L2(P), U2(P);
        for (int k = 0; k < P; k++) {
            int km = k - d, kp = k + d;
            bool hm = true, hp2 = true;
            if (periodic) { km = ((km % P) + P) % P; kp = kp % P; }
            else { hm = km >= 0; hp2 = kp < P; }
            Mat alpha = mat_zero(B), beta = mat_zero(B);
            if (hm) alpha = mat_neg(mat_mul(L[k], mat_inv(D[km], B), B));
            if (hp2) beta = mat_neg(mat_mul(U[k], mat_inv(D[kp], B), B));
            Mat Dn = D[k];
            if (hm) Dn = mat_add(Dn, mat_mul(alpha, U[km], B));
            if (hp2) Dn = mat_add(Dn, mat_mul(beta, L[kp], B));
            D2[k] = Dn;
            L2[k] = hm ? mat_mul(alpha, L[km], B) : mat_zero(B);
            U2[k] = hp2 ? mat_mul(beta, U[kp], B) : mat_zero(B);
            double *dst = &steps[((size_t)st * P + k) * 2 * B * B];
            for (int i = 0; i < B * B; i++) { dst[i] = (double)alpha[i]; dst[B * B + i] = (double)beta[i]; }
        }
        D = D2; L = L2; U = U2;
    }
    pcr = steps;
    for (int k = 0; k < P; k++) {
        Mat Mk = periodic ? mat_add(mat_add(D[k], L[k]), U[k]) : D[k];
        Mat inv = mat_inv(Mk, B);
        for (int i = 0; i < B * B; i++) pcr.push_back((double)inv[i]);
    }

    // coefficients (src/operators.jl:87-92): derivative RHS prescaled by 1/h,
    // high-edge stencils negated for derivatives
    std::vector<double> coef(COEF_N, 0.0);
    double hinv = 1.0 / h;
    double sgn = s.symmetric ? 1.0 : -1.0;
    double scale = s.symmetric ? 1.0 : hinv;
    for (int mm = 0; mm < M; mm++) coef[COEF_CI + mm] = s.coeffs[mm] * scale;
    coef[COEF_A0] = s.a0;
    for (int j = 0; j < nc; j++)
        for (int kk = 0; kk < s.closure_rhs_len[j]; kk++) {
            coef[COEF_CLO + j * 8 + kk] = s.closure_rhs[j][kk] * scale;
            coef[COEF_CHI + j * 8 + kk] = s.closure_rhs[j][kk] * (scale * sgn);
        }
    if (hi_fold) {
        // rows n-1-jr (jr < nc) from fe[kap] = f[n-1-kap]: the interior stencil
        // with each ghost f(n + g) read as sigma_f f(n - 1 - g)
        const double sf = s.symmetric ? (double)hi_fold : -(double)hi_fold;
        for (int jr = 0; jr < nc; jr++) {
            double *w = &coef[COEF_CHI + jr * 8];
            for (int kk = 0; kk < 8; kk++) w[kk] = 0.0;
            // offset o from row n-1-jr reaches f[n-1-jr+o], i.e. kap = jr - o; kap < 0
            // is the ghost n + (-kap-1), which mirrors kap' = -kap - 1
            auto add = [&](int kap, double v) { w[kap < 0 ? -kap - 1 : kap] += kap < 0 ? sf * v : v; };
            if (s.symmetric) add(jr, s.a0);
            for (int mm = 1; mm <= M; mm++) {
                add(jr - mm, s.coeffs[mm - 1] * scale);
                add(jr + mm, sgn * s.coeffs[mm - 1] * scale);
            }
        }
    }
    if (lo_fold) {
        // the interior stencil of rows 0..nc-1 with each ghost f(j < 0) read as
        // sigma_f f(-j-1) (folds.jl:79-105 fold_fill! on a half-offset grid)
        const double sf = s.symmetric ? (double)lo_fold : -(double)lo_fold;
        for (int j = 0; j < nc; j++) {
            double *w = &coef[COEF_CLO + j * 8];
            for (int kk = 0; kk < 8; kk++) w[kk] = 0.0;
            auto add = [&](int col, double v) { w[col < 0 ? -col - 1 : col] += col < 0 ? sf * v : v; };
            if (s.symmetric) add(j, s.a0);
            for (int mm = 1; mm <= M; mm++) {
                add(j + mm, s.coeffs[mm - 1] * scale);
                add(j - mm, sgn * s.coeffs[mm - 1] * scale);
            }
        }
    }

    // chunk-major copy of the row constants for the ragged kernel
    std::vector<double> rowt;
    if (chunk_major) {
        if (max_chunk > RAG_MAX_CHUNK) throw Error{CLES_ERR_INVALID, "chunk-major tables hold chunks of at most 32 points"};
        rowt.assign((size_t)nrow * RAG_MAX_CHUNK * P, 0.0);
        for (int f = 0; f < nrow; f++)
            for (int k = 0; k < P; k++)
                for (int t = 0; t < clen(k); t++)
                    rowt[((size_t)f * RAG_MAX_CHUNK + t) * P + k] = row[(size_t)f * n + cstart(k) + t];
    }

    // host image: coef64 | coef32 | row64 | row32 | pcr | rowt64 | rowt32 (256-byte aligned parts)
    size_t b_coef64 = COEF_N * sizeof(double), b_coef32 = COEF_N * sizeof(float);
    size_t b_row64 = row.size() * sizeof(double), b_row32 = row.size() * sizeof(float);
    size_t b_pcr = pcr.size() * sizeof(double);
    size_t b_rowt64 = rowt.size() * sizeof(double), b_rowt32 = rowt.size() * sizeof(float);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    out.o_coef32 = al(b_coef64);
    out.o_row64 = out.o_coef32 + al(b_coef32);
    out.o_row32 = out.o_row64 + al(b_row64);
    out.o_pcr = out.o_row32 + al(b_row32);
    out.o_rowt64 = out.o_pcr + al(std::max<size_t>(b_pcr, 8));
    out.o_rowt32 = out.o_rowt64 + al(b_rowt64);
    const size_t total = out.o_rowt32 + al(b_rowt32);
    out.image.assign(total, 0);
    if (!rowt.empty()) {
        std::vector<float> rowt32(rowt.begin(), rowt.end());
        std::memcpy(out.image.data() + out.o_rowt64, rowt.data(), b_rowt64);
        std::memcpy(out.image.data() + out.o_rowt32, rowt32.data(), b_rowt32);
    }
    std::memcpy(out.image.data(), coef.data(), b_coef64);
    std::vector<float> coef32(coef.begin(), coef.end()), row32(row.begin(), row.end());
    std::memcpy(out.image.data() + out.o_coef32, coef32.data(), b_coef32);
    std::memcpy(out.image.data() + out.o_row64, row.data(), b_row64);
    std::memcpy(out.image.data() + out.o_row32, row32.data(), b_row32);
    if (b_pcr) std::memcpy(out.image.data() + out.o_pcr, pcr.data(), b_pcr);
    LineDev &L0 = out.dev;
    L0 = LineDev{};
    L0.n = n; L0.q = q; L0.M = M; L0.sym = s.symmetric ? 1 : 0; L0.nc = closed ? nc : 0;
    L0.periodic = periodic ? 1 : 0; L0.closed = closed ? 1 : 0;
    L0.P = P; L0.nsteps = nsteps; L0.m = m; L0.extra = extra;
}

void bind_line_tables(LineDev &dev, const LineTables &t, unsigned char *base)
{
    dev = t.dev;
    dev.coef64 = (const double *)base;
    dev.coef32 = (const float *)(base + t.o_coef32);
    dev.row64 = (const double *)(base + t.o_row64);
    dev.row32 = (const float *)(base + t.o_row32);
    dev.pcr = (const double *)(base + t.o_pcr);
    dev.rowt64 = t.o_rowt64 != t.o_rowt32 ? (const double *)(base + t.o_rowt64) : nullptr;
    dev.rowt32 = t.o_rowt64 != t.o_rowt32 ? (const float *)(base + t.o_rowt32) : nullptr;
}

void build_line_plan(HostPlan &hp, const cles_scheme_desc &s, int n, int periodic, double h, int max_chunk,
                     int lo_fold, int hi_fold)
{
    LineTables t;
    build_line_tables(t, s, n, periodic, h, max_chunk, hp.dim, false, lo_fold, hi_fold);
    void *d = nullptr;
    CLES_HIP(hipMalloc(&d, t.image.size()));
    hipError_t e = hipMemcpy(d, t.image.data(), t.image.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        throw Error{CLES_ERR_HIP, std::string("hipMemcpy(plan tables): ") + hipGetErrorString(e)};
    }
    hp.dbuf = d;
    hp.dbytes = t.image.size();
    bind_line_tables(hp.dev, t, (unsigned char *)d);
    hp.h = h;
}

void free_line_plan(HostPlan &hp)
{
    if (hp.dbuf) (void)hipFree(hp.dbuf);
    hp.dbuf = nullptr;
}

}  // namespace cles
