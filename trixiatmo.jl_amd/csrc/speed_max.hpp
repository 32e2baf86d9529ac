// The maximum of the StepsizeCallback wave-speed reductions.
#ifndef ATMO_SPEED_MAX_HPP
#define ATMO_SPEED_MAX_HPP

#include <hip/hip_runtime.h>

namespace atmo {

// max that propagates NaN, as Julia's max does in the reference's max_dt
// (fmax would drop it): a wave speed that is not a number makes the whole
// reduction NaN, and the step-size control then reports a non-finite state
// instead of stepping on.  Exact and order-independent like fmax.
__device__ __forceinline__ double speed_max(double a, double b) {
  return (a != a || b != b) ? a + b : (a > b ? a : b);
}

}  // namespace atmo

#endif
