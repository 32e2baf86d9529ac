// Mixed-precision instantiations: FP32 fluxes with Kahan-compensated sums.
#include "euler3d_stage.hpp"

namespace atmo {
bool e3_launch_stage_fp32_kahan(const E3Launch& L, hipStream_t s) {
  return dispatch_stage<float, true>(L, s);
}
}  // namespace atmo
