// FP32-bulk instantiation of the fused step kernel (built with
// -ffp-contract=off: the fused multiply-adds are the explicit ones of the
// kernel source, so results do not depend on the tile shape).
#include "erod_fused_kernel.hpp"

namespace erod {
  int fused_launch_fp32(erod_batch* b, const fused::Consts<float>& C, int steps) {
    return fused::launch_all<float>(b, C, steps);
  }
}  // namespace erod
