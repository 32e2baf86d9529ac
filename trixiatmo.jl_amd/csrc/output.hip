// Output variables of the resident state: the `solution_variables` of
// SaveSolutionCallback, computed on the device and returned variable-major
// ([V][nodes], every variable one contiguous vector in Trixi's node order), the
// layout of the solution file's `variables_v` datasets
// (src/callbacks_step/save_solution_2d_manifold_in_3d_cartesian.jl:90-97).
//
//  * cons2cons: the conservative state itself.
//  * cons2prim, 3D Euler (src/equations/compressible_euler_energy_with_gravity_3d.jl:568-578):
//    (rho, v1, v2, v3, p, phi).
//  * cons2prim, rainy Euler (src/equations/compressible_rainy_euler_2d.jl:114-129):
//    the densities, the velocities, the energy density and u[7:9].
//  * cons2prim_and_vorticity, covariant shallow water
//    (src/callbacks_step/save_solution_covariant.jl:65-76, 92-117): total height
//    H = h + h_s, the global velocity A vcon, h_s and the relative vorticity
//    (d1 v_2 - d2 v_1) / J of the covariant velocity v_i = G_ij vcon^j by the
//    collocation derivative.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "atmo_internal.hpp"
#include "handle.hpp"

namespace atmo {

namespace {

struct DerivMat {
  double d[64];  // D[i * n + j]
};

__global__ void cons2cons_kernel(const double* __restrict__ u, int V, long long n_nodes,
                                 double* __restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_nodes) return;
  for (int v = 0; v < V; ++v) out[v * n_nodes + q] = u[q * V + v];
}

__global__ void euler3d_cons2prim_kernel(const double* __restrict__ u, long long n_nodes,
                                         double gm1, double* __restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_nodes) return;
  const double* a = u + q * 6;
  const double rho = a[0], v1 = a[1] / rho, v2 = a[2] / rho, v3 = a[3] / rho;
  const double p = gm1 * (a[4] - 0.5 * (a[1] * v1 + a[2] * v2 + a[3] * v3));
  const double w[6] = {rho, v1, v2, v3, p, a[5]};
  for (int v = 0; v < 6; ++v) out[v * n_nodes + q] = w[v];
}

__global__ void rainy_cons2prim_kernel(const double* __restrict__ u, long long n_nodes,
                                       double* __restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_nodes) return;
  const double* a = u + q * 9;
  const double rho_inv = 1.0 / (a[0] + a[1] + a[2]);
  const double w[9] = {a[0], a[1], a[2], a[3] * rho_inv, a[4] * rho_inv, a[5], a[6], a[7], a[8]};
  for (int v = 0; v < 9; ++v) out[v * n_nodes + q] = w[v];
}

// covariant velocity component c (0: v_1, 1: v_2) at node g
__device__ __forceinline__ double covariant_velocity(const double* u, const double* aux, long long g,
                                                     int c) {
  const double* a = aux + g * 26;
  const double inv_h = 1.0 / u[g * 3];
  const double vc1 = u[g * 3 + 1] * inv_h, vc2 = u[g * 3 + 2] * inv_h;
  return c == 0 ? a[13] * vc1 + a[14] * vc2 : a[14] * vc1 + a[15] * vc2;
}

template <int N>
__global__ void swe_prim_vorticity_kernel(const double* __restrict__ u,
                                          const double* __restrict__ aux, DerivMat D,
                                          long long n_nodes, double* __restrict__ out) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_nodes) return;
  const int q = (int)(g % (N * N));
  const long long base = g - q;
  const int i = q % N, j = q / N;
  const double* a = aux + g * 26;
  const double h = u[g * 3], vc1 = u[g * 3 + 1] / h, vc2 = u[g * 3 + 2] / h;
  const double hs = a[19];
  double d2 = 0.0, d1 = 0.0;
  for (int m = 0; m < N; ++m) {
    d2 += D.d[i * N + m] * covariant_velocity(u, aux, base + m + N * j, 1);
    d1 += D.d[j * N + m] * covariant_velocity(u, aux, base + i + N * m, 0);
  }
  const double w[6] = {h + hs,
                       a[0] * vc1 + a[3] * vc2,
                       a[1] * vc1 + a[4] * vc2,
                       a[2] * vc1 + a[5] * vc2,
                       hs,
                       (d2 - d1) / a[12]};
  for (int v = 0; v < 6; ++v) out[v * n_nodes + g] = w[v];
}

}  // namespace

}  // namespace atmo

using namespace atmo;

extern "C" {

int atmo_solution_variables(atmo_handle* h, int32_t kind, double* out, int32_t* n_vars) {
  if (!h || !out || kind < 0 || kind > 1)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad solution-variable arguments");
  const int V = h->eq == ATMO_EQ_COVARIANT_SWE_2D ? 3 : h->eq == ATMO_EQ_RAINY_EULER_2D ? 9 : 6;
  const long long n_nodes = h->len / V;
  const int V_out = kind == 0 ? V : (h->eq == ATMO_EQ_COVARIANT_SWE_2D ? 6 : V);
  double* d_out = nullptr;
  HIP_TRY(h, hipMalloc(&d_out, sizeof(double) * V_out * (size_t)n_nodes));
  const dim3 grid((unsigned)((n_nodes + 255) / 256)), block(256);
  if (kind == 0) {
    hipLaunchKernelGGL(cons2cons_kernel, grid, block, 0, h->stream, h->d_u, V, n_nodes, d_out);
  } else if (h->eq == ATMO_EQ_RAINY_EULER_2D) {
    hipLaunchKernelGGL(rainy_cons2prim_kernel, grid, block, 0, h->stream, h->d_u, n_nodes, d_out);
  } else if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) {
    DerivMat D{};
    Basis b(h->n);
    for (int q = 0; q < h->n * h->n; ++q) D.d[q] = b.D[q];
    switch (h->n) {
#define ATMO_OUT_CASE(NN)                                                                     \
  case NN:                                                                                    \
    hipLaunchKernelGGL(swe_prim_vorticity_kernel<NN>, grid, block, 0, h->stream, h->d_u,      \
                       h->d_aux, D, n_nodes, d_out);                                          \
    break;
      ATMO_OUT_CASE(2)
      ATMO_OUT_CASE(3)
      ATMO_OUT_CASE(4)
      ATMO_OUT_CASE(5)
      ATMO_OUT_CASE(6)
      ATMO_OUT_CASE(7)
      ATMO_OUT_CASE(8)
#undef ATMO_OUT_CASE
      default:
        (void)hipFree(d_out);
        return fail(&h->err, ATMO_ERR_INVALID, "polydeg out of range");
    }
  } else {
    const double gm1 = h->cfg.c_p / h->cfg.c_v - 1.0;
    hipLaunchKernelGGL(euler3d_cons2prim_kernel, grid, block, 0, h->stream, h->d_u, n_nodes, gm1,
                       d_out);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_out, sizeof(double) * V_out * (size_t)n_nodes, hipMemcpyDeviceToHost,
                       h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d_out);
  if (e != hipSuccess) return fail(&h->err, ATMO_ERR_DEVICE, hipGetErrorString(e));
  if (n_vars) *n_vars = V_out;
  return ATMO_OK;
}

}  // extern "C"
