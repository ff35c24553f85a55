// Ragged-batch line kernels: schemes C6, C4, C10 (derivatives) and the
// eighth-order filter family, FP64 and FP32.
#include "cles_line_rag.hpp"

namespace cles {

namespace {
template <typename T>
void launch_rag_typed(const RagLines &r, const T *f, const EpiArgs &e, hipStream_t st)
{
    const int q = r.q, M = r.M, sym = r.sym;
    if (q == 1 && M == 2 && !sym) return launch_rag_epi<T, 1, 2, 0>(r, f, e, st);   // Lele C6
    if (q == 1 && M == 1 && !sym) return launch_rag_epi<T, 1, 1, 0>(r, f, e, st);   // Pade C4
    if (q == 2 && M == 3 && !sym) return launch_rag_epi<T, 2, 3, 0>(r, f, e, st);   // Lele C10
    if (q == 1 && sym) return launch_rag_epi<T, 1, 4, 1>(r, f, e, st);              // G-V filters
    throw Error{CLES_ERR_UNSUPPORTED, "scheme shape not available for batched instances"};
}
}  // namespace

void launch_line_rag(const RagLines &r, int dtype, const void *f, const EpiArgs &e, hipStream_t st)
{
    if (dtype == CLES_F64)
        launch_rag_typed<double>(r, (const double *)f, e, st);
    else
        launch_rag_typed<float>(r, (const float *)f, e, st);
}

}  // namespace cles
