// Explicit instantiation: float line operators, q=2 M=4 sym=0.
#include "cles_line.hpp"

namespace cles {
template void launch_epi<float, 2, 4, 0>(const HostPlan &, int, const View3 &, const float *, const EpiArgs &,
                                            hipStream_t, const LineFields &);
}  // namespace cles
