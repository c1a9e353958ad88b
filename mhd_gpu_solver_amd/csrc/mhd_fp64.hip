// fp64 instantiation of the step kernels. Compiled with -ffp-contract=off: every product and sum
// rounds separately, in the reference CPU path's operation order.
#include "mhd_launch_impl.hpp"

namespace mhdhip {
MHDHIP_INSTANTIATE(double)
}  // namespace mhdhip
