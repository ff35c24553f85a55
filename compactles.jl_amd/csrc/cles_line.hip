// Line-operator dispatch: (q, M, symmetric) -> explicit instantiation
// (compiled in the li_*.hip translation units).
#include "cles_line.hpp"

namespace cles {

#define CLES_EXTERN(T, Q, M, S)                                                                      \
    extern template void launch_epi<T, Q, M, S>(const HostPlan &, int, const View3 &, const T *,   \
                                                const EpiArgs &, hipStream_t, const LineFields &);
CLES_EXTERN(double, 1, 2, 0)
CLES_EXTERN(double, 1, 4, 1)
CLES_EXTERN(double, 2, 3, 0)
CLES_EXTERN(double, 1, 1, 0)
CLES_EXTERN(double, 1, 4, 0)
CLES_EXTERN(double, 2, 4, 0)
CLES_EXTERN(float, 1, 2, 0)
CLES_EXTERN(float, 1, 4, 1)
CLES_EXTERN(float, 2, 3, 0)
CLES_EXTERN(float, 1, 1, 0)
CLES_EXTERN(float, 1, 4, 0)
CLES_EXTERN(float, 2, 4, 0)
#undef CLES_EXTERN

template <typename T>
static void launch_typed(const HostPlan &hp, int dim, const View3 &v, const T *f, const EpiArgs &e, hipStream_t st,
                         const LineFields &mf)
{
    const LineDev &p = hp.dev;
    const int q = p.q, M = p.M, sym = p.sym;
    if (q == 1 && M == 2 && !sym) return launch_epi<T, 1, 2, 0>(hp, dim, v, f, e, st, mf);   // Lele C6
    if (q == 1 && M == 3 && !sym) return launch_epi<T, 1, 4, 0>(hp, dim, v, f, e, st, mf);
    if (q == 1 && M == 1 && !sym) return launch_epi<T, 1, 1, 0>(hp, dim, v, f, e, st, mf);   // Pade C4
    if (q == 1 && !sym) return launch_epi<T, 1, 4, 0>(hp, dim, v, f, e, st, mf);
    if (q == 1 && sym) return launch_epi<T, 1, 4, 1>(hp, dim, v, f, e, st, mf);              // G-V filters
    if (q == 2 && M == 3 && !sym) return launch_epi<T, 2, 3, 0>(hp, dim, v, f, e, st, mf);   // Lele C10
    if (q == 2 && !sym) return launch_epi<T, 2, 4, 0>(hp, dim, v, f, e, st, mf);
    throw Error{CLES_ERR_UNSUPPORTED, "pentadiagonal symmetric (filter) schemes are not supported"};
}

void launch_line(const HostPlan &hp, int dim, const View3 &v, const void *f, const EpiArgs &e, hipStream_t st,
                 const LineFields *mf)
{
    static const LineFields one{};
    const LineFields &m = mf ? *mf : one;
    if (m.nf < 1 || m.nf > LINE_MF) throw Error{CLES_ERR_INVALID, "fields per line launch must be 1..12"};
    if (m.nf > 1) {
        const int k = e.kind;
        if (k != EPI_STORE && k != EPI_STORE_NEG && k != EPI_SUB && k != EPI_STORE_GATE && k != EPI_LSRK &&
            k != EPI_LSRK_DEV && k != EPI_ART_MULTI)
            throw Error{CLES_ERR_INVALID, "several fields per launch need a store, subtract, LSRK or multi epilogue"};
        for (int i = 1; i < m.nf; i++) {
            if (!m.in[i]) throw Error{CLES_ERR_INVALID, "NULL field input"};
            if (k != EPI_LSRK && k != EPI_LSRK_DEV && k != EPI_ART_MULTI && !m.out[i])
                throw Error{CLES_ERR_INVALID, "NULL field output"};
        }
    }
    if (e.kind == EPI_ART_MULTI && !e.multi) throw Error{CLES_ERR_INVALID, "the multi epilogue needs its table"};
    if (hp.dtype == CLES_F64)
        launch_typed<double>(hp, dim, v, (const double *)f, e, st, m);
    else
        launch_typed<float>(hp, dim, v, (const float *)f, e, st, m);
}

}  // namespace cles
