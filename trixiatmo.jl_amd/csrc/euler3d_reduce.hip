// Reductions over the resident state: the per-element maximum scaled wave
// speed of StepsizeCallback's max_dt and the integrals of the analysis
// (mass, entropy, dS/dt = int w(u).du, every conserved variable).
#include <hip/hip_runtime.h>

#include "euler3d_kernels.hpp"
#include "euler3d_physics.hpp"
#include "speed_max.hpp"

namespace atmo {

template <int N>
struct Shape {
  static constexpr int N3 = N * N * N;
  static constexpr int BLOCK = ((N3 + 63) / 64) * 64;
};

struct WeightTable {
  double w[8];
};

// ---------------------------------------------------------------- reductions
// per element: max over nodes of |Ja^d . lambda| / J summed over d
template <int N>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_max_speed(E3Params q, const double* __restrict__ u,
                                                               const double* __restrict__ cv,
                                                               const double* __restrict__ invJ,
                                                               double* __restrict__ partial) {
  constexpr int N3 = Shape<N>::N3;
  constexpr int B = Shape<N>::BLOCK;
  __shared__ double s[3][B];
  const int e = blockIdx.x, t = threadIdx.x;
  double l1 = 0, l2 = 0, l3 = 0;
  if (t < N3) {
    const size_t g = (size_t)e * N3 + t;
    double uu[6], l[3];
    for (int v = 0; v < 6; ++v) uu[v] = u[g * 6 + v];
    node_scaled_speeds(q, uu, cv + g * 9, fabs(invJ[g]), l);
    l1 = l[0];
    l2 = l[1];
    l3 = l[2];
  }
  s[0][t] = l1; s[1][t] = l2; s[2][t] = l3;
  __syncthreads();
  for (int w = B / 2; w > 0; w >>= 1) {
    if (t < w) {
      s[0][t] = speed_max(s[0][t], s[0][t + w]);
      s[1][t] = speed_max(s[1][t], s[1][t + w]);
      s[2][t] = speed_max(s[2][t], s[2][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) partial[e] = s[0][0] + s[1][0] + s[2][0];
}

// per element: integrals of mass, entropy, w(u).du and every conserved variable
template <int N>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_integrals(E3Params q, WeightTable wt,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ du,
                                                               const double* __restrict__ invJ,
                                                               double* __restrict__ partial) {
  constexpr int N3 = Shape<N>::N3;
  constexpr int B = Shape<N>::BLOCK;
  constexpr int K = 9;
  __shared__ double s[K][B];
  const int e = blockIdx.x, t = threadIdx.x;
  double val[K] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (t < N3) {
    const size_t g = (size_t)e * N3 + t;
    const int i = t % N, j = (t / N) % N, k = t / (N * N);
    const double wJ = wt.w[i] * wt.w[j] * wt.w[k] / invJ[g];
    double uu[6];
    for (int v = 0; v < 6; ++v) uu[v] = u[g * 6 + v];
    const double rho = uu[0];
    const double v1 = uu[1] / rho, v2 = uu[2] / rho, v3 = uu[3] / rho;
    const double vsq = v1 * v1 + v2 * v2 + v3 * v3;
    const double p = q.gm1 * (uu[4] - 0.5 * rho * vsq);
    const double sth = log(p) - q.gamma * log(rho);
    val[0] = wJ * rho;
    val[1] = wJ * (-sth * rho * q.inv_gm1);
    if (du) {
      const double rp = rho / p;
      const double w0 = (q.gamma - sth) * q.inv_gm1 - 0.5 * rp * vsq;
      const double* d = du + g * 6;
      val[2] = wJ * (w0 * d[0] + rp * v1 * d[1] + rp * v2 * d[2] + rp * v3 * d[3] - rp * d[4]);
    }
    for (int v = 0; v < 6; ++v) val[3 + v] = wJ * uu[v];
  }
  for (int c = 0; c < K; ++c) s[c][t] = val[c];
  __syncthreads();
  for (int w = B / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int c = 0; c < K; ++c) s[c][t] += s[c][t + w];
    __syncthreads();
  }
  if (t < K) partial[(size_t)e * K + t] = s[t][0];
}

// single-block reduction of partial[n_items * width] (op 0: sum, 1: max)
__global__ __launch_bounds__(256) void reduce_columns(const double* __restrict__ partial,
                                                      int n_items, int width, int op,
                                                      double* __restrict__ result) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  for (int c = 0; c < width; ++c) {
    double acc = op == 0 ? 0.0 : -1e300;
    for (int i = t; i < n_items; i += 256) {
      const double x = partial[(size_t)i * width + c];
      acc = op == 0 ? acc + x : speed_max(acc, x);
    }
    s[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) s[t] = op == 0 ? s[t] + s[t + w] : speed_max(s[t], s[t + w]);
      __syncthreads();
    }
    if (t == 0) result[c] = s[0];
    __syncthreads();
  }
}

template <int N>
static void launch_red(const E3Launch& L, int which, const double* du, double* partial,
                       hipStream_t s) {
  // the quadrature weights enter the integrals only (the max-speed reduction
  // may be launched without them)
  WeightTable wt{};
  if (which != 0)
    for (int q = 0; q < N; ++q) wt.w[q] = L.weights[q];
  if (which == 0)
    hipLaunchKernelGGL((e3_max_speed<N>), dim3(L.n_elements), dim3(Shape<N>::BLOCK), 0, s, L.q,
                       L.u, L.cv, L.invJ, partial);
  else
    hipLaunchKernelGGL((e3_integrals<N>), dim3(L.n_elements), dim3(Shape<N>::BLOCK), 0, s, L.q,
                       wt, L.u, du, L.invJ, partial);
}

bool e3_launch_reduction(const E3Launch& L, int which, const double* du, double* partial,
                         double* result, hipStream_t s) {
  switch (L.n) {
    case 2: launch_red<2>(L, which, du, partial, s); break;
    case 3: launch_red<3>(L, which, du, partial, s); break;
    case 4: launch_red<4>(L, which, du, partial, s); break;
    case 5: launch_red<5>(L, which, du, partial, s); break;
    case 6: launch_red<6>(L, which, du, partial, s); break;
    case 7: launch_red<7>(L, which, du, partial, s); break;
    case 8: launch_red<8>(L, which, du, partial, s); break;
    default: return false;
  }
  const int width = which == 0 ? 1 : 9;
  hipLaunchKernelGGL(reduce_columns, dim3(1), dim3(256), 0, s, partial, L.n_elements, width,
                     which == 0 ? 1 : 0, result);
  return true;
}

}  // namespace atmo

namespace atmo {

__global__ void background_state_kernel(E3Params q, Background bg, const double* __restrict__ u,
                                        double* __restrict__ ubg, long n_nodes) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const double phi = u[i * 6 + 5];
  const double pb = bg.p0 * exp(-phi * bg.inv_RT0);
  ubg[i * 6 + 0] = pb * bg.inv_RT0;
  ubg[i * 6 + 1] = 0.0;
  ubg[i * 6 + 2] = 0.0;
  ubg[i * 6 + 3] = 0.0;
  ubg[i * 6 + 4] = pb * q.inv_gm1;
  ubg[i * 6 + 5] = phi;
}

__global__ void extract_momentum_kernel(const double* __restrict__ du, float* __restrict__ B,
                                        long n_nodes) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  B[i * 3 + 0] = (float)du[i * 6 + 1];
  B[i * 3 + 1] = (float)du[i * 6 + 2];
  B[i * 3 + 2] = (float)du[i * 6 + 3];
}

void e3_launch_background_state(const E3Launch& L, const double* u, double* ubg, hipStream_t s) {
  const long n = (long)L.n_elements * L.n * L.n * L.n;
  hipLaunchKernelGGL(background_state_kernel, dim3((n + 255) / 256), dim3(256), 0, s, L.q, L.bg, u,
                     ubg, n);
}

void e3_launch_extract_momentum(const double* du, float* B, long n_nodes, hipStream_t s) {
  hipLaunchKernelGGL(extract_momentum_kernel, dim3((n_nodes + 255) / 256), dim3(256), 0, s, du, B,
                     n_nodes);
}

}  // namespace atmo
