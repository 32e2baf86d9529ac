// FP64 arithmetic with the contravariant vectors stored in FP32 (the direct
// stage form reads the FP32 copy and widens each value where it is staged).
#include "euler3d_stage.hpp"

namespace atmo {
bool e3_launch_stage_fp64_geo32(const E3Launch& L, hipStream_t s) {
  return dispatch_stage<double, false, float>(L, s);
}
}  // namespace atmo
