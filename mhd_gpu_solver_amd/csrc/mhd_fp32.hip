// fp32-bulk instantiation of the single-run step kernels (pressure recovery and the
// fast-speed discriminant stay in fp64, see MixedOps<float>). Contraction into FMA within each
// expression, plus the fused multiply-adds the sources spell out: the results do not depend on
// the direction a row segment is marched in.
#include "mhd_launch_impl.hpp"

namespace mhdhip {
MHDHIP_INSTANTIATE_SINGLE(float)
}  // namespace mhdhip
