// Fast fp64 build of the stepping kernels (precision MHD_HIP_FP64_FAST): fp64 storage and
// arithmetic in the same operation order, with contraction into FMA allowed and the division
// and square root of the flux and speed formulas taken from the hardware reciprocal and
// reciprocal square root refined by Newton steps (MHD_FAST_FP64; about 1 ulp instead of
// correctly rounded). Results stay within 1e-12 relative L1 of the exact fp64 path; the exact
// path (mhd_fp64.hip) remains the one bit-identical to the CPU.
#define MHD_FAST_FP64 1
#include "mhd_launch_impl.hpp"

namespace mhdhip {
MHDHIP_INSTANTIATE_FAST(double)
}  // namespace mhdhip
