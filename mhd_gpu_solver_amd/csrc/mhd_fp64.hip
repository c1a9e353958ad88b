// fp64 instantiation of the step kernels. Compiled with -ffp-contract=off: every product and sum
// rounds separately, in the reference CPU path's operation order.
#include "mhd_launch_impl.hpp"

namespace mhdhip {
MHDHIP_INSTANTIATE(double)
// the state comparison for every pairing of storage types (fp64 differences, no contraction)
template hipError_t launch_rel_l1<double, double>(const double*, const double*, const Geom&,
                                                  DD*, double*, hipStream_t);
template hipError_t launch_rel_l1<float, double>(const float*, const double*, const Geom&, DD*,
                                                 double*, hipStream_t);
template hipError_t launch_rel_l1<double, float>(const double*, const float*, const Geom&, DD*,
                                                 double*, hipStream_t);
template hipError_t launch_rel_l1<float, float>(const float*, const float*, const Geom&, DD*,
                                                double*, hipStream_t);
}  // namespace mhdhip
