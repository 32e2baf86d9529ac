// atmo_max_dt3d_noalloc: StepsizeCallback's max_dt of a 3D state held by the
// caller in device memory, for callers that cannot allocate device memory
// (code under HIP graph capture, tight loops, harnesses that forbid it).  It
// runs the kernels of atmo_max_dt (per-element maximum scaled speeds, then
// their maximum) with the per-element partials and the maximum in `scratch`,
// a file-scope device array of fixed size, instead of buffers of a handle.
//
// Each GPU holds its own copy of `scratch` at its own address, found once per
// device with hipGetSymbolAddress while that device is current; kernels and
// copies take that address, never the array's name (which in host code is a
// host-side stand-in, not a device address).  Every call on one GPU shares
// that GPU's array: calls that overlap on one GPU (two streams, two threads)
// race, and the caller serialises them.  Every entry of the array that a call
// reads it has written first (one partial per element, then the maximum), so
// nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/atmo_hip.h"
#include "euler3d_kernels.hpp"
#include "handle.hpp"

namespace atmo {

namespace {

// per-element partials of up to ATMO_NOALLOC_MAX_ELEMENTS elements, then the
// maximum (2 MiB + 256 B per GPU)
constexpr int kScratchLen = ATMO_NOALLOC_MAX_ELEMENTS + 32;
static __device__ __align__(256) double scratch[kScratchLen];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's `scratch` (looked up once per device)
int scratch_address(double** ptr) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(nullptr, ATMO_ERR_DEVICE, "no current HIP device");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, ATMO_ERR_DEVICE,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return ATMO_OK;
}

// dt = cfl * (2 / (n * max speed)) from the reduced maximum (NaN when that
// maximum is not finite, where atmo_max_dt returns ATMO_ERR_NONFINITE)
__global__ void max_dt_finish_kernel(const double* __restrict__ mx_in, double cfl, int n,
                                     double* __restrict__ dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mx = mx_in[0];
  if (!isfinite(mx)) {
    dt[0] = __builtin_nan("");
    return;
  }
  if (mx < 5e-324) mx = 5e-324;
  dt[0] = cfl * (2.0 / (n * mx));
}

}  // namespace

}  // namespace atmo

using namespace atmo;

extern "C" int atmo_max_dt3d_noalloc(const atmo_config* cfg, const double* u,
                                     const double* contravariant_vectors,
                                     const double* inverse_jacobian, double cfl, double* dt,
                                     void* stream) {
  if (!cfg || !u || !contravariant_vectors || !inverse_jacobian || !dt)
    return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (cfg->equations != ATMO_EQ_EULER_ENERGY_GRAVITY_3D)
    return fail(nullptr, ATMO_ERR_INVALID, "atmo_max_dt3d_noalloc takes the 3D Euler equations");
  if (cfg->polydeg < 1 || cfg->polydeg > 7)
    return fail(nullptr, ATMO_ERR_INVALID, "polydeg must be in 1..7");
  if (cfg->n_elements < 1 || cfg->n_elements > ATMO_NOALLOC_MAX_ELEMENTS)
    return fail(nullptr, ATMO_ERR_INVALID,
                "n_elements must be in 1.." + std::to_string(ATMO_NOALLOC_MAX_ELEMENTS) +
                    " (the size of the library's fixed scratch array)");
  if (!(cfg->c_v > 0.0) || !(cfg->c_p > cfg->c_v))
    return fail(nullptr, ATMO_ERR_INVALID, "need c_p > c_v > 0");
  if (!(cfl > 0.0) || !std::isfinite(cfl)) return fail(nullptr, ATMO_ERR_INVALID, "bad cfl");
  double* ptr = nullptr;
  const int rc = scratch_address(&ptr);
  if (rc != ATMO_OK) return rc;
  E3Launch L{};
  const double gamma = cfg->c_p / cfg->c_v;
  L.q.gamma = gamma;
  L.q.gm1 = gamma - 1.0;
  L.q.inv_gm1 = 1.0 / (gamma - 1.0);
  L.n = cfg->polydeg + 1;
  L.n_elements = cfg->n_elements;
  L.u = u;
  L.cv = contravariant_vectors;
  L.invJ = inverse_jacobian;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = ptr;
  double* result = ptr + ATMO_NOALLOC_MAX_ELEMENTS;
  if (!e3_launch_reduction(L, 0, nullptr, partial, result, s))
    return fail(nullptr, ATMO_ERR_INVALID, "no kernel for this polydeg");
  hipLaunchKernelGGL(max_dt_finish_kernel, dim3(1), dim3(64), 0, s, result, cfl, L.n, dt);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
  return ATMO_OK;
}
