// Test-only library: the FP64 arithmetic helpers of the stage kernels
// (trixiatmo.jl_amd/csrc/euler3d_physics.hpp) applied element-wise on the
// device, so that tests can compare them with correctly rounded host values,
// and the rainy two-point flux (rainy_flux.hpp) evaluated with its states in
// both orders.  Built by __graft_entry__.build() next to this file; no
// product code loads it.
#include <hip/hip_runtime.h>

#include "../../trixiatmo.jl_amd/csrc/euler3d_physics.hpp"
#include "../../trixiatmo.jl_amd/csrc/rainy_flux.hpp"

namespace {

__global__ void probe_kernel(int op, const double* __restrict__ a, const double* __restrict__ b,
                             double* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  double r = 0.0;
  switch (op) {
    case 0: r = atmo::log_(x); break;
    case 1: r = atmo::sqrt_(x); break;
    case 2: r = atmo::div_nr(x, y); break;
    case 3: r = atmo::ln_mean_logs(x, y, atmo::log_(x), atmo::log_(y)); break;
    case 4: r = atmo::inv_ln_mean_logs(x, y, atmo::log_(x) - atmo::log_(y)); break;
    default: r = 0.0;
  }
  out[i] = r;
}

// node fields [n][11] in ECNode order: rd rv rc rr v1 v2 it vr lrd lrv lit
__device__ atmo::rainy::ECNode load_node(const double* p) {
  atmo::rainy::ECNode a;
  a.rd = p[0]; a.rv = p[1]; a.rc = p[2]; a.rr = p[3]; a.v1 = p[4]; a.v2 = p[5];
  a.it = p[6]; a.vr = p[7]; a.lrd = p[8]; a.lrv = p[9]; a.lit = p[10];
  return a;
}

// thread k < n evaluates flux(l_k, r_k), thread n + k flux(r_k, l_k): one
// call site, as in the stage kernel, where the two nodes of a pair run the
// same instructions with their states in opposite order
__global__ void flux_kernel(const double* __restrict__ l, const double* __restrict__ r, int n,
                            double n1, double n2, double* __restrict__ f_lr,
                            double* __restrict__ f_rl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * n) return;
  const bool swap = k >= n;
  const int i = swap ? k - n : k;
  const double* pa = (swap ? r : l) + 11 * (size_t)i;
  const double* pb = (swap ? l : r) + 11 * (size_t)i;
  const atmo::rainy::ECNode a = load_node(pa), b = load_node(pb);
  double f[6];
  atmo::rainy::flux_ec_rain(a, b, n1, n2, f);
  double* out = (swap ? f_rl : f_lr) + 6 * (size_t)i;
  for (int v = 0; v < 6; ++v) out[v] = f[v];
}

}  // namespace

// flux_ec_rain(l_i, r_i) and flux_ec_rain(r_i, l_i) for n node pairs
// (fields [n][11] each, host arrays; results [n][6]); 0 or a HIP error code
extern "C" int atmo_probe_flux_ec_rain(const double* l, const double* r, int n, double n1,
                                       double n2, double* f_lr, double* f_rl) {
  if (n <= 0) return 0;
  const size_t bin = sizeof(double) * 11 * (size_t)n, bout = sizeof(double) * 6 * (size_t)n;
  double *dl = nullptr, *dr = nullptr, *d1 = nullptr, *d2 = nullptr;
  hipError_t e = hipMalloc(&dl, bin);
  if (e == hipSuccess) e = hipMalloc(&dr, bin);
  if (e == hipSuccess) e = hipMalloc(&d1, bout);
  if (e == hipSuccess) e = hipMalloc(&d2, bout);
  if (e == hipSuccess) e = hipMemcpy(dl, l, bin, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dr, r, bin, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(flux_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, 0, dl, dr, n, n1, n2,
                       d1, d2);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(f_lr, d1, bout, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(f_rl, d2, bout, hipMemcpyDeviceToHost);
  for (double* p : {dl, dr, d1, d2})
    if (p) (void)hipFree(p);
  return (int)e;
}

// op: 0 log_(a), 1 sqrt_(a), 2 div_nr(a, b), 3 ln_mean_logs(a, b), 4
// inv_ln_mean_logs(a, b); host arrays of n values; returns 0 or a HIP error code
extern "C" int atmo_probe_math(int op, const double* a, const double* b, double* out, int n) {
  if (n <= 0) return 0;
  const size_t bytes = sizeof(double) * (size_t)n;
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&da, bytes);
  if (e == hipSuccess) e = hipMalloc(&db, bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, bytes);
  if (e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, op, da, db, dout, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dout) (void)hipFree(dout);
  return (int)e;
}
