// Explicit instantiation: float line operators, q=1 M=4 sym=1.
#include "cles_line.hpp"

namespace cles {
template void launch_epi<float, 1, 4, 1>(const HostPlan &, int, const View3 &, const float *, const EpiArgs &,
                                            hipStream_t, const LineFields &);
}  // namespace cles
