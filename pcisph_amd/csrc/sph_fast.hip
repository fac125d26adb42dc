// sph_fast.hip - fast FP32 instantiation of the step kernels (built with
// -ffp-contract=fast): fused multiply-adds, hardware reciprocal and square root.
#include "sph_launch_impl.h"

namespace pcisph {
template struct SphLaunch<float, true>;
}  // namespace pcisph
