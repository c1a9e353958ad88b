// fp32-bulk instantiation of the step kernels (pressure recovery and the fast-speed discriminant
// stay in fp64, see MixedOps<float>). Contraction into FMA is allowed here.
#include "mhd_launch_impl.hpp"

namespace mhdhip {
MHDHIP_INSTANTIATE(float)
}  // namespace mhdhip
