// Mixed-precision instantiations: FP32 fluxes with plain FP32 sums.
#include "euler3d_stage.hpp"

namespace atmo {
bool e3_launch_stage_fp32(const E3Launch& L, hipStream_t s) {
  return dispatch_stage<float, false>(L, s);
}
}  // namespace atmo
