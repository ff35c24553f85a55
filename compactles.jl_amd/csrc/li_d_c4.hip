// Explicit instantiation: double line operators, q=1 M=1 sym=0.
#include "cles_line.hpp"

namespace cles {
template void launch_epi<double, 1, 1, 0>(const HostPlan &, int, const View3 &, const double *, const EpiArgs &,
                                            hipStream_t, const LineFields &);
}  // namespace cles
