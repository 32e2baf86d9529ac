// Geometry of the covariant form on the device: the 26 auxiliary variables
// per node of init_auxiliary_node_variables!
// (src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl:232-305,
// GlobalCartesianCoordinates) from the node coordinates of
// P4estMeshCubedSphere2D(...; element_local_mapping = true): the covariant
// basis A of the bilinear map of the element's corners projected onto the
// sphere (:310-337), G = A^T A, G^-1, the contravariant basis B = G^-1 A^T,
// J = sqrt(det G), and the Christoffel symbols of the second kind (:430-489)
// from dG/dxi -- exact derivatives of the analytic metric (forward-mode dual
// numbers: exact to rounding, as the host builder's complex step) or the
// collocation derivative of the nodal metric.  Then the stage kernel's
// compact rows and the per-face-node maps B_own A_nbr.  Every value the host
// builder computes by the same operations is computed here in its order with
// floating-point contraction off, so it equals the host's bit for bit; the
// exact Christoffel symbols agree with the host's to rounding.
#include <hip/hip_runtime.h>

#include "swe_kernels.hpp"

namespace atmo {

namespace {

// a + b eps with eps^2 = 0: the derivative in one direction rides along
struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const double r = sqrt(a.v);
  return {r, a.d / (2.0 * r)};
}
__device__ __forceinline__ Dual cst(double c) { return {c, 0.0}; }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }

template <typename T>
__device__ __forceinline__ T tc(double c);
template <>
__device__ __forceinline__ double tc<double>(double c) { return c; }
template <>
__device__ __forceinline__ Dual tc<Dual>(double c) { return cst(c); }

// basis_covariant (:310-337) of the bilinear map of the corners v1..v4
// projected onto the sphere of the given radius, at (xi1, xi2)
template <typename T>
__device__ __forceinline__ void basis_covariant(const double* v1, const double* v2,
                                                const double* v3, const double* v4, T xi1, T xi2,
                                                double radius, T A[3][2]) {
#pragma clang fp contract(off)
  T xe[3], d1[3], d2[3];
  const T one = tc<T>(1.0), q = tc<T>(0.25);
  for (int k = 0; k < 3; ++k) {
    const T a1 = tc<T>(v1[k]), a2 = tc<T>(v2[k]), a3 = tc<T>(v3[k]), a4 = tc<T>(v4[k]);
    xe[k] = q * ((one - xi1) * (one - xi2) * a1 + (one + xi1) * (one - xi2) * a2 +
                 (one + xi1) * (one + xi2) * a3 + (one - xi1) * (one + xi2) * a4);
    d1[k] = q * (-(one - xi2) * a1 + (one - xi2) * a2 + (one + xi2) * a3 - (one + xi2) * a4);
    d2[k] = q * (-(one - xi1) * a1 - (one + xi1) * a2 + (one + xi1) * a3 + (one - xi1) * a4);
  }
  const T nx = dsqrt(xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2]);
  const T p1 = (xe[0] * d1[0] + xe[1] * d1[1] + xe[2] * d1[2]) / (nx * nx);
  const T p2 = (xe[0] * d2[0] + xe[1] * d2[1] + xe[2] * d2[2]) / (nx * nx);
  const T rn = tc<T>(radius) / nx;
  for (int k = 0; k < 3; ++k) {
    A[k][0] = rn * (d1[k] - p1 * xe[k]);
    A[k][1] = rn * (d2[k] - p2 * xe[k]);
  }
}

template <typename T>
__device__ __forceinline__ void metric_covariant(const double* v1, const double* v2,
                                                 const double* v3, const double* v4, T xi1, T xi2,
                                                 double radius, T G[3]) {
#pragma clang fp contract(off)
  T A[3][2];
  basis_covariant(v1, v2, v3, v4, xi1, xi2, radius, A);
  G[0] = A[0][0] * A[0][0] + A[1][0] * A[1][0] + A[2][0] * A[2][0];
  G[1] = A[0][0] * A[0][1] + A[1][0] * A[1][1] + A[2][0] * A[2][1];
  G[2] = A[0][1] * A[0][1] + A[1][1] * A[1][1] + A[2][1] * A[2][1];
}

// calc_christoffel_symbols (:430-453): first kind from dG/dxi, raised with G^-1
__device__ __forceinline__ void christoffel(const double dG1[3], const double dG2[3],
                                            const double* Gc, double* out) {
#pragma clang fp contract(off)
  const double G11_1 = dG1[0], G12_1 = dG1[1], G22_1 = dG1[2];
  const double G11_2 = dG2[0], G12_2 = dG2[1], G22_2 = dG2[2];
  const double g1[2][2] = {{0.5 * G11_1, 0.5 * G11_2}, {0.5 * G11_2, G12_2 - 0.5 * G22_1}};
  const double g2[2][2] = {{G12_1 - 0.5 * G11_2, 0.5 * G22_1}, {0.5 * G22_1, 0.5 * G22_2}};
  const double Gc11 = Gc[0], Gc12 = Gc[1], Gc22 = Gc[2];
  out[0] = Gc11 * g1[0][0] + Gc12 * g2[0][0];
  out[1] = Gc11 * g1[0][1] + Gc12 * g2[0][1];
  out[2] = Gc11 * g1[1][1] + Gc12 * g2[1][1];
  out[3] = Gc12 * g1[0][0] + Gc22 * g2[0][0];
  out[4] = Gc12 * g1[0][1] + Gc22 * g2[0][1];
  out[5] = Gc12 * g1[1][1] + Gc22 * g2[1][1];
}

struct NodeTable {
  double x[8];   // LGL nodes
  double D[64];  // D[i * n + j]
};

// one 64-lane workgroup per element, one lane per node (n^2 <= 64)
template <int N>
__global__ __launch_bounds__(64) void swe_aux_kernel(NodeTable nt, const double* __restrict__ X,
                                                     const double* __restrict__ topography,
                                                     int collocation,
                                                     const int* __restrict__ member,
                                                     const int* __restrict__ member_first,
                                                     double* __restrict__ aux) {
#pragma clang fp contract(off)
  constexpr int N2 = N * N;
  __shared__ double sG[3][N2];
  const int e = blockIdx.x, t = threadIdx.x;
  const bool active = t < N2;
  const int i = t % N, j = t / N;
  // corners as the reference reads them from tree_node_coordinates, and the
  // radius of the first node of the element's mesh (ensemble member)
  const double* v1 = X + 3 * ((size_t)N2 * e);
  const double* v2 = X + 3 * ((N - 1) + (size_t)N2 * e);
  const double* v3 = X + 3 * ((N - 1) + N * (N - 1) + (size_t)N2 * e);
  const double* v4 = X + 3 * (N * (N - 1) + (size_t)N2 * e);
  const double* X0 = member ? X + 3 * (size_t)N2 * member_first[member[e]] : X;
  const double R = sqrt(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
  double* a = aux + 26 * ((size_t)N2 * e + t);
  if (active) {
    double A[3][2];
    basis_covariant<double>(v1, v2, v3, v4, nt.x[i], nt.x[j], R, A);
    for (int k = 0; k < 3; ++k) {
      a[k] = A[k][0];
      a[3 + k] = A[k][1];
    }
    const double G11 = A[0][0] * A[0][0] + A[1][0] * A[1][0] + A[2][0] * A[2][0];
    const double G12 = A[0][0] * A[0][1] + A[1][0] * A[1][1] + A[2][0] * A[2][1];
    const double G22 = A[0][1] * A[0][1] + A[1][1] * A[1][1] + A[2][1] * A[2][1];
    const double det = G11 * G22 - G12 * G12;
    const double Gc11 = G22 / det, Gc12 = -G12 / det, Gc22 = G11 / det;
    // contravariant basis B = G^-1 A^T (2 x 3, column-major in aux[6..11])
    for (int k = 0; k < 3; ++k) {
      a[6 + 2 * k] = Gc11 * A[k][0] + Gc12 * A[k][1];
      a[7 + 2 * k] = Gc12 * A[k][0] + Gc22 * A[k][1];
    }
    a[12] = sqrt(det);
    a[13] = G11;
    a[14] = G12;
    a[15] = G22;
    a[16] = Gc11;
    a[17] = Gc12;
    a[18] = Gc22;
    a[19] = topography ? topography[(size_t)N2 * e + t] : 0.0;
    sG[0][t] = G11;
    sG[1][t] = G12;
    sG[2][t] = G22;
  }
  __syncthreads();
  if (!active) return;
  double dG1[3], dG2[3];
  if (!collocation) {
    // exact derivatives of the analytic metric
    Dual G[3];
    metric_covariant<Dual>(v1, v2, v3, v4, Dual{nt.x[i], 1.0}, cst(nt.x[j]), R, G);
    for (int k = 0; k < 3; ++k) dG1[k] = G[k].d;
    metric_covariant<Dual>(v1, v2, v3, v4, cst(nt.x[i]), Dual{nt.x[j], 1.0}, R, G);
    for (int k = 0; k < 3; ++k) dG2[k] = G[k].d;
  } else {
    // collocation derivative of the nodal covariant metric (:380-416)
    for (int k = 0; k < 3; ++k) {
      dG1[k] = 0.0;
      dG2[k] = 0.0;
    }
    for (int q = 0; q < N; ++q)
      for (int k = 0; k < 3; ++k) {
        dG1[k] += nt.D[i * N + q] * sG[k][q + N * j];
        dG2[k] += nt.D[j * N + q] * sG[k][i + N * q];
      }
  }
  christoffel(dG1, dG2, a + 16, a + 20);
}

// the stage kernel's compact rows from the auxiliary variables (the host
// builder's packing: G, h_s, the six Christoffel symbols, z / r, J)
__global__ void swe_compact_rows_kernel(const double* __restrict__ aux,
                                        const double* __restrict__ X, long long nn,
                                        double* __restrict__ geo) {
#pragma clang fp contract(off)
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nn) return;
  const double* a = aux + 26 * g;
  for (int c = 0; c < 3; ++c) geo[c * nn + g] = a[13 + c];
  geo[3 * nn + g] = a[19];
  for (int c = 0; c < 6; ++c) geo[(4 + c) * nn + g] = a[20 + c];
  const double* x = X + 3 * g;
  geo[10 * nn + g] = x[2] / sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  geo[11 * nn + g] = a[12];
}

// per face node the 2 x 2 map B_own A_nbr of the neighbour's contravariant
// momentum into the own basis (contravariant2global, then
// global2contravariant: covariant_shallow_water.jl:140-153), in the host
// builder's order
template <int N>
__global__ void swe_face_maps_kernel(const double* __restrict__ aux, const int* __restrict__ nbr,
                                     int n_elements, double* __restrict__ fxf) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 4LL * N * n_elements) return;
  const int k = (int)(q % N);
  const int f = (int)((q / N) % 4);
  const long long e = q / (4 * N);
  const int i = f == 0 ? 0 : (f == 1 ? N - 1 : k), j = f == 2 ? 0 : (f == 3 ? N - 1 : k);
  const double* bo = aux + ((size_t)e * N * N + i + N * j) * 26 + 6;  // B (2 x 3)
  const double* an = aux + (size_t)nbr[q] * 26;                        // A (3 x 2)
  double* M = fxf + q * 4;
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c)
      M[2 * r + c] = bo[r] * an[3 * c] + bo[r + 2] * an[1 + 3 * c] + bo[r + 4] * an[2 + 3 * c];
}

}  // namespace

bool swe_launch_aux(int n, int n_elements, const double* x_lgl, const double* D, const double* X,
                    const double* topography, int collocation, const int* member,
                    const int* member_first, double* aux, hipStream_t s) {
  NodeTable nt{};
  for (int q = 0; q < n; ++q) nt.x[q] = x_lgl[q];
  for (int q = 0; q < n * n; ++q) nt.D[q] = D[q];
  switch (n) {
#define ATMO_AUX_CASE(NN)                                                                   \
  case NN:                                                                                  \
    hipLaunchKernelGGL(swe_aux_kernel<NN>, dim3(n_elements), dim3(64), 0, s, nt, X,         \
                       topography, collocation, member, member_first, aux);                 \
    return true;
    ATMO_AUX_CASE(2)
    ATMO_AUX_CASE(3)
    ATMO_AUX_CASE(4)
    ATMO_AUX_CASE(5)
    ATMO_AUX_CASE(6)
    ATMO_AUX_CASE(7)
    ATMO_AUX_CASE(8)
#undef ATMO_AUX_CASE
    default:
      return false;
  }
}

void swe_launch_compact_rows(const double* aux, const double* X, long long nn, double* geo,
                             hipStream_t s) {
  hipLaunchKernelGGL(swe_compact_rows_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                     aux, X, nn, geo);
}

bool swe_launch_face_maps(int n, int n_elements, const double* aux, const int* nbr, double* fxf,
                          hipStream_t s) {
  const long long total = 4LL * n * n_elements;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  switch (n) {
#define ATMO_FXF_CASE(NN)                                                                   \
  case NN:                                                                                  \
    hipLaunchKernelGGL(swe_face_maps_kernel<NN>, grid, block, 0, s, aux, nbr, n_elements, fxf); \
    return true;
    ATMO_FXF_CASE(2)
    ATMO_FXF_CASE(3)
    ATMO_FXF_CASE(4)
    ATMO_FXF_CASE(5)
    ATMO_FXF_CASE(6)
    ATMO_FXF_CASE(7)
    ATMO_FXF_CASE(8)
#undef ATMO_FXF_CASE
    default:
      return false;
  }
}

}  // namespace atmo
