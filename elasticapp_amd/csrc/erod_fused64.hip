// FP64 instantiation of the fused step kernel (built with -ffp-contract=off:
// every product and sum rounded on its own, in the reference's order).
#include "erod_fused_kernel.hpp"

namespace erod {
  int fused_launch_fp64(erod_batch* b, const fused::Consts<double>& C, int steps) {
    return fused::launch_all<double>(b, C, steps);
  }
}  // namespace erod
