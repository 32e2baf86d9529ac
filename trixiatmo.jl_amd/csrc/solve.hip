// The time loop of the examples' solve(ode, CarpenterKennedy2N54(williamson_condition
// = false); callback = StepsizeCallback(cfl)) (e.g.
// examples/shallow_water/covariant/elixir_barotropic_instability.jl:79-80)
// kept on the device.  Every step is the maximum-wave-speed reduction of the
// resident state, a one-thread-per-member control kernel that turns it into
// the step size (cut to end at t_end, as OrdinaryDiffEq's tstop handling
// does) and the five fused stages, which read the step size from device
// memory.  Two steps are captured as one graph and replayed; the host reads
// the clock back only every `check_every` steps.  A member that has reached
// t_end takes steps of size 0, which leave its state unchanged, so the
// members of an ensemble each follow their own step sizes and finish
// together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "handle.hpp"
#include "speed_max.hpp"

// clock layout (M members): t[M] (start of the step in flight), dt[M], steps
// taken [M], status (1: non-finite wave speed), t_end, cfl
constexpr int kClockExtra = 3;

namespace atmo {

namespace {

// the step just taken is added to t, then the next step size is
// dt = min(cfl * (2 / (n * max speed)), t_end - t) as atmo_max_dt and the host
// loop compute it (same operations in the same order), or 0 once t >= t_end
__device__ __forceinline__ void step_control_kernel_body(double max_speed, int m, int M, int n,
                                                         double* clock, double* dt) {
  const double t_end = clock[3 * M + 1], cfl = clock[3 * M + 2];
  const double t = clock[m] + dt[m];
  double d = 0.0;
  if (t < t_end) {
    double mx = max_speed;
    if (!isfinite(mx)) {
      clock[3 * M] = 1.0;  // every member that finds one stores the same value
    } else {
      if (mx < 5e-324) mx = 5e-324;
      d = fmin(cfl * (2.0 / (n * mx)), t_end - t);
      clock[2 * M + m] += 1.0;
    }
  }
  clock[m] = t;
  dt[m] = d;
}

__global__ void step_control_kernel(const double* __restrict__ max_speed, int M, int n,
                                    double* clock, double* dt) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < M) step_control_kernel_body(max_speed[m], m, M, n, clock, dt);
}

// the same for one member whose maximum speed is the maximum over per-element
// partials (written by the last stage of the step just taken)
__global__ __launch_bounds__(1024) void step_control_reduce_kernel(const double* __restrict__ partial,
                                                                   int n_items, int n,
                                                                   double* clock, double* dt) {
  __shared__ double s[1024 / 64];
  const int t = threadIdx.x;
  double m = 0.0;
  // eight independent loads in flight per lane, then their maximum
  for (int i0 = t; i0 < n_items; i0 += 8 * 1024) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = i0 + k * 1024 < n_items ? partial[i0 + k * 1024] : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) m = speed_max(m, v[k]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = speed_max(m, __shfl_xor(m, o, 64));
  if ((t & 63) == 0) s[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < 1024 / 64; ++w) m = speed_max(m, s[w]);
    s[0] = m;
  }
  __syncthreads();
  if (t == 0) step_control_kernel_body(s[0], 0, 1, n, clock, dt);
}

// one workgroup per ensemble member: its maximum speed is the maximum over the
// per-element partials of its elements [member_first[m], member_first[m + 1])
__global__ __launch_bounds__(256) void step_control_members_reduce_kernel(
    const double* __restrict__ partial, const int* __restrict__ member_first, int M, int n,
    double* clock, double* dt) {
  __shared__ double s[256 / 64];
  const int m = blockIdx.x, t = threadIdx.x;
  double mx = 0.0;
  const int e1 = member_first[m + 1];
  for (int e0 = member_first[m] + t; e0 < e1; e0 += 8 * 256) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = e0 + k * 256 < e1 ? partial[e0 + k * 256] : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mx = speed_max(mx, v[k]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = speed_max(mx, __shfl_xor(mx, o, 64));
  if ((t & 63) == 0) s[t >> 6] = mx;
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < 256 / 64; ++w) mx = speed_max(mx, s[w]);
    step_control_kernel_body(mx, m, M, n, clock, dt);
  }
}

}  // namespace

void launch_step_control_members_reduce(const double* partial, const int* member_first,
                                        int n_members, int n, double* clock, double* dt,
                                        hipStream_t s) {
  hipLaunchKernelGGL(step_control_members_reduce_kernel, dim3(n_members), dim3(256), 0, s, partial,
                     member_first, n_members, n, clock, dt);
}

void launch_step_control(const double* max_speed, int n_members, int n, double* clock, double* dt,
                         hipStream_t s) {
  hipLaunchKernelGGL(step_control_kernel, dim3((n_members + 63) / 64), dim3(64), 0, s, max_speed,
                     n_members, n, clock, dt);
}

namespace {
__global__ void set_scalar_kernel(double* ptr, double value) {
  if (threadIdx.x == 0) *ptr = value;
}
}  // namespace

void launch_set_scalar(double* ptr, double value, hipStream_t s) {
  hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(64), 0, s, ptr, value);
}

void launch_step_control_reduce(const double* partial, int n_items, int n, double* clock,
                                double* dt, hipStream_t s) {
  hipLaunchKernelGGL(step_control_reduce_kernel, dim3(1), dim3(1024), 0, s, partial, n_items, n,
                     clock, dt);
}

}  // namespace atmo

namespace {

using atmo::fail;

int enqueue_resident_steps(atmo_handle* h, int n_steps) {
  for (int s = 0; s < n_steps; ++s) {
    const int rc = h->eq == ATMO_EQ_COVARIANT_SWE_2D ? atmo::swe_enqueue_resident_step(h)
                   : h->eq == ATMO_EQ_RAINY_EULER_2D ? atmo::rainy_enqueue_resident_step(h)
                                                     : atmo::e3_enqueue_resident_step(h);
    if (rc != ATMO_OK) return rc;
  }
  return ATMO_OK;
}

// two resident steps as one graph (the ping-pong state buffers return to
// their places after an even number of stages)
int capture_two_steps(atmo_handle* h, hipGraphExec_t* exec) {
  hipGraph_t g = nullptr;
  HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_resident_steps(h, 2);
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (rc != ATMO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(h, ec);
  const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_TRY(h, ei);
  return ATMO_OK;
}

}  // namespace

extern "C" {

int atmo_solve(atmo_handle* h, double t_end, double cfl, int32_t check_every, int64_t* steps,
               double* t_reached) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  if (!(cfl > 0) || !std::isfinite(cfl) || !std::isfinite(t_end) || check_every < 1)
    return fail(&h->err, ATMO_ERR_INVALID, "bad solve arguments");
  const bool swe = h->eq == ATMO_EQ_COVARIANT_SWE_2D;
  const int M = swe ? h->n_members : 1;
  const size_t clock_len = 3 * (size_t)M + kClockExtra;
  if (!h->d_clock) HIP_TRY(h, hipMalloc(&h->d_clock, sizeof(double) * clock_len));
  // the step-size array the stages read: the shallow-water stage reads its
  // per-member array, the others the clock's
  double* d_dt = swe ? h->d_dt_member : h->d_clock + M;
  std::vector<double> clk(clock_len, 0.0);
  for (int m = 0; m < M; ++m) clk[m] = h->time;
  clk[3 * M + 1] = t_end;
  clk[3 * M + 2] = cfl;
  HIP_TRY(h, hipMemcpyAsync(h->d_clock, clk.data(), sizeof(double) * clock_len,
                            hipMemcpyHostToDevice, h->stream));
  if (swe) {
    HIP_TRY(h, hipMemcpyAsync(h->d_dt_member, clk.data() + M, sizeof(double) * M,
                              hipMemcpyHostToDevice, h->stream));
    if (!h->swe_tuned) {
      const int rc = atmo::swe_autotune(h, nullptr);
      if (rc != ATMO_OK) return rc;
    }
  }
  if (h->eq == ATMO_EQ_EULER_ENERGY_GRAVITY_3D && !h->e3_tuned) {
    const int rc = atmo::e3_autotune(h, nullptr);
    if (rc != ATMO_OK) return rc;
  }
  h->du_valid = false;
  bool done = !(h->time < t_end);
  if (!done) {
    const int rc = h->eq == ATMO_EQ_EULER_ENERGY_GRAVITY_3D ? atmo::e3_resident_prologue(h)
                   : swe                                  ? atmo::swe_resident_prologue(h)
                                                          : atmo::rainy_resident_prologue(h);
    if (rc != ATMO_OK) return rc;
  }
  hipGraphExec_t graph = nullptr;
  if (!h->timing && !done) {
    const int rc = capture_two_steps(h, &graph);
    if (rc != ATMO_OK) return rc;
  }
  std::vector<double> back(clock_len), dt_back(M);
  int rc = ATMO_OK;
  // a first short chunk gives the step size the later chunks are sized by
  int64_t chunk = std::min<int64_t>(check_every, 2);
  while (!done) {
    if (graph) {
      for (int64_t s = 0; s < chunk; s += 2) {
        const hipError_t e = hipGraphLaunch(graph, h->stream);
        if (e != hipSuccess) {
          rc = fail(&h->err, ATMO_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
          break;
        }
      }
    } else {
      rc = enqueue_resident_steps(h, (int)chunk);
    }
    if (rc != ATMO_OK) break;
    hipError_t e = hipMemcpyAsync(back.data(), h->d_clock, sizeof(double) * clock_len,
                                  hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && swe)
      e = hipMemcpyAsync(dt_back.data(), d_dt, sizeof(double) * M, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      rc = fail(&h->err, ATMO_ERR_DEVICE, std::string("clock read-back: ") + hipGetErrorString(e));
      break;
    }
    if (back[3 * M] != 0.0) {
      rc = fail(&h->err, ATMO_ERR_NONFINITE, "non-finite wave speed in state");
      break;
    }
    // members still running after the step in flight, and the steps they
    // still need at their last step size: the next chunk is sized to that
    done = true;
    double need = 0.0;
    for (int m = 0; m < M; ++m) {
      const double dtm = swe ? dt_back[m] : back[M + m];
      const double tm = back[m] + dtm;
      if (tm < t_end) {
        done = false;
        need = std::max(need, dtm > 0 ? std::ceil((t_end - tm) / dtm) : (double)check_every);
      }
    }
    chunk = (int64_t)std::min<double>(std::max(need, 1.0), (double)check_every);
    if (graph) chunk += chunk & 1;
  }
  if (graph) (void)hipGraphExecDestroy(graph);
  if (rc != ATMO_OK && rc != ATMO_ERR_NONFINITE) return rc;
  // final clock: each member's time is t + dt of its last step
  HIP_TRY(h, hipMemcpyAsync(back.data(), h->d_clock, sizeof(double) * clock_len,
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(dt_back.data(), d_dt, sizeof(double) * M, hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  double t_min = 0.0;
  for (int m = 0; m < M; ++m) {
    const double tm = back[m] + dt_back[m];
    t_min = m == 0 ? tm : std::min(t_min, tm);
    if (steps) steps[m] = (int64_t)back[2 * M + m];
  }
  if (M == 0) t_min = h->time;
  h->time = t_min;
  if (t_reached) *t_reached = t_min;
  return rc;
}

}  // extern "C"
