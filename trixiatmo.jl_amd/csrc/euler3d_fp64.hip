// FP64 instantiations of the fused stage kernel (compared with the oracle).
#include "euler3d_stage.hpp"

namespace atmo {
bool e3_launch_stage_fp64(const E3Launch& L, hipStream_t s) {
  return dispatch_stage<double, false>(L, s);
}

bool e3_has_two_element_groups(int n, int precision) {
  // the FP32-geometry form computes in FP64 (precision 3)
  const bool f64 = precision == 0 || precision == 3;
  switch (n) {
    case 4: return f64 ? has_two_element_groups<4, double>() : has_two_element_groups<4, float>();
    case 5: return f64 ? has_two_element_groups<5, double>() : has_two_element_groups<5, float>();
    default: return false;
  }
}
}  // namespace atmo
