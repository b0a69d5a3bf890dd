// FP32-bulk instantiation of the fused step kernel (FMA contraction allowed).
#include "erod_fused_kernel.hpp"

namespace erod {
  int fused_launch_fp32(erod_batch* b, const fused::Consts<float>& C, int steps) {
    return fused::launch_all<float>(b, C, steps);
  }
}  // namespace erod
