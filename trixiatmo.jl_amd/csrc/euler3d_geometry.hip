// Metric terms of curved hexahedra on the device: from the node coordinates
// of each element, the Jacobian matrix by the collocation derivative, the
// contravariant vectors Ja^d in Kopriva's invariant curl form (Kopriva 2006,
// J. Sci. Comput. 26, "Metric identities and the discontinuous spectral
// element method on curvilinear meshes") and 1/J from the determinant; the
// MetricTermsInvariantCurl choice named at
// src/semidiscretization/metric_terms.jl:25.  One workgroup per element, the
// coordinates and the Jacobian matrix staged in LDS.  Every sum runs in the
// order of the host mesh builder, without fused multiply-adds, so the result
// equals the host's bit for bit.
#include <hip/hip_runtime.h>

#include "euler3d_kernels.hpp"

namespace atmo {

namespace {

struct DerivativeMatrix {
  double d[64];  // D[i * n + j]
};

template <int N>
struct GeoShape {
  static constexpr int N3 = N * N * N;
  static constexpr int BLOCK = ((N3 + 63) / 64) * 64;
};

// node index with the index along direction d replaced by q
template <int N>
__device__ __forceinline__ int along(int d, int i, int j, int k, int q) {
  return d == 0 ? q + N * (j + N * k) : d == 1 ? i + N * (q + N * k) : i + N * (j + N * q);
}

template <int N>
__global__ __launch_bounds__(GeoShape<N>::BLOCK) void e3_metric_terms_kernel(
    DerivativeMatrix D, const double* __restrict__ X, double* __restrict__ cv,
    double* __restrict__ invJ) {
#pragma clang fp contract(off)
  constexpr int N3 = GeoShape<N>::N3, B = GeoShape<N>::BLOCK;
  __shared__ double sx[3][N3];
  __shared__ double sj[9][N3];  // dX_c / dxi_d at [c + 3 d]
  const int e = blockIdx.x, t = threadIdx.x;
  for (int idx = t; idx < 3 * N3; idx += B) sx[idx % 3][idx / 3] = X[(size_t)e * 3 * N3 + idx];
  __syncthreads();
  const bool active = t < N3;
  const int i = t % N, j = (t / N) % N, k = t / (N * N);
  if (active) {
    for (int d = 0; d < 3; ++d) {
      const int own = d == 0 ? i : d == 1 ? j : k;
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int q = 0; q < N; ++q) s += D.d[own * N + q] * sx[c][along<N>(d, i, j, k, q)];
        sj[c + 3 * d][t] = s;
      }
    }
  }
  __syncthreads();
  if (!active) return;
  // Ja^d_c = 1/2 [ (X_m dX_l/dxi_{d2} - X_l dX_m/dxi_{d2})_{,xi_{d1}}
  //              - (X_m dX_l/dxi_{d1} - X_l dX_m/dxi_{d1})_{,xi_{d2}} ]
  // with (c, m, l) and (d, d1, d2) cyclic
  const size_t g = (size_t)e * N3 + t;
  for (int d = 0; d < 3; ++d) {
    const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
    const int own1 = d1 == 0 ? i : d1 == 1 ? j : k;
    const int own2 = d2 == 0 ? i : d2 == 1 ? j : k;
    for (int c = 0; c < 3; ++c) {
      const int mm = (c + 1) % 3, ll = (c + 2) % 3;
      double term1 = 0.0, term2 = 0.0;
      for (int q = 0; q < N; ++q) {
        const int nd = along<N>(d1, i, j, k, q);
        const double gq = sx[mm][nd] * sj[ll + 3 * d2][nd] - sx[ll][nd] * sj[mm + 3 * d2][nd];
        term1 += D.d[own1 * N + q] * gq;
      }
      for (int q = 0; q < N; ++q) {
        const int nd = along<N>(d2, i, j, k, q);
        const double gq = sx[mm][nd] * sj[ll + 3 * d1][nd] - sx[ll][nd] * sj[mm + 3 * d1][nd];
        term2 += D.d[own2 * N + q] * gq;
      }
      cv[9 * g + c + 3 * d] = 0.5 * (term1 - term2);
    }
  }
  double J[9];
  for (int r = 0; r < 9; ++r) J[r] = sj[r][t];
  // determinant by the rule of Sarrus, in the host builder's order
  const double det = J[0] * J[4] * J[8] + J[3] * J[7] * J[2] + J[6] * J[1] * J[5] -
                     J[6] * J[4] * J[2] - J[3] * J[1] * J[8] - J[0] * J[7] * J[5];
  invJ[g] = 1.0 / det;
}

__global__ void narrow_to_float_kernel(const double* __restrict__ a, float* __restrict__ b,
                                       size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) b[q] = (float)a[q];
}

}  // namespace

bool e3_launch_metric_terms(int n, int n_elements, const double* D, const double* xyz,
                            double* cv, double* invJ, hipStream_t s) {
  DerivativeMatrix dm{};
  for (int q = 0; q < n * n; ++q) dm.d[q] = D[q];
  switch (n) {
#define ATMO_GEO_CASE(NN)                                                                    \
  case NN:                                                                                   \
    hipLaunchKernelGGL(e3_metric_terms_kernel<NN>, dim3(n_elements),                         \
                       dim3(GeoShape<NN>::BLOCK), 0, s, dm, xyz, cv, invJ);                  \
    return true;
    ATMO_GEO_CASE(2)
    ATMO_GEO_CASE(3)
    ATMO_GEO_CASE(4)
    ATMO_GEO_CASE(5)
    ATMO_GEO_CASE(6)
    ATMO_GEO_CASE(7)
    ATMO_GEO_CASE(8)
#undef ATMO_GEO_CASE
    default:
      return false;
  }
}

void launch_narrow_to_float(const double* a, float* b, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(narrow_to_float_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     a, b, n);
}

}  // namespace atmo
