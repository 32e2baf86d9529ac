// FP64 instantiations of the fused stage kernel (compared with the oracle).
#include "euler3d_stage.hpp"

namespace atmo {
bool e3_launch_stage_fp64(const E3Launch& L, hipStream_t s) {
  return dispatch_stage<double, false>(L, s);
}
}  // namespace atmo
