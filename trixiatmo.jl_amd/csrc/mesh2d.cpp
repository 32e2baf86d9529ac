// Host-side construction of the 2D cubed-sphere mesh for the covariant
// shallow-water form and its 26 auxiliary variables per node.
//
//  * P4estMeshCubedSphere2D(trees_per_face_dimension, radius; polydeg,
//    element_local_mapping = true): src/meshes/p4est_cubed_sphere.jl:47-79,
//    corners by the equiangular gnomonic projection, nodes by the bilinear map
//    of the corners projected to the sphere (:345-401, Guba et al. 2014,
//    Appendix A).  Trees are numbered panel-major with x fastest, as the
//    LinearIndices((n_cells_x, n_cells_y, 6)) of the reference.
//  * Auxiliary variables: src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl
//    :232-305 (GlobalCartesianCoordinates basis, G, G^-1, contravariant basis,
//    J = sqrt(det G), bottom topography) and :380-489 (Christoffel symbols of
//    the second kind from dG/dxi: exact derivatives of the analytic metric,
//    here by complex-step differentiation, which is exact to rounding, or the
//    collocation derivative of the nodal metric).
//  * Connectivity: element edges matched by node coordinates; primary side =
//    (element, face) pair sorting first, indexed :i_forward.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <map>
#include <string>
#include <vector>

#include "atmo_internal.hpp"

namespace atmo {

// cubed_sphere_mapping of the six cube faces (xi, eta in [-1, 1]) to the
// sphere of the given radius, equiangular gnomonic projection
static void cubed_sphere_point(double xi, double eta, double radius, int direction, double out[3]) {
  const double x = std::tan(xi * M_PI / 4.0), y = std::tan(eta * M_PI / 4.0);
  double c[3];
  switch (direction) {
    case 1: c[0] = -1; c[1] = -x; c[2] = y; break;
    case 2: c[0] = 1; c[1] = x; c[2] = y; break;
    case 3: c[0] = x; c[1] = -1; c[2] = y; break;
    case 4: c[0] = -x; c[1] = 1; c[2] = y; break;
    case 5: c[0] = -x; c[1] = y; c[2] = -1; break;
    default: c[0] = x; c[1] = y; c[2] = 1; break;
  }
  const double r = std::sqrt(1.0 + x * x + y * y);
  for (int k = 0; k < 3; ++k) out[k] = radius / r * c[k];
}

// basis_covariant for GlobalCartesianCoordinates (:310-337): A[:, i] = dx/dxi_i
// of the bilinear map of the corners projected onto the sphere
template <typename T>
static void basis_covariant(const double* v1, const double* v2, const double* v3,
                            const double* v4, T xi1, T xi2, double radius, T A[3][2]) {
  T xe[3], d1[3], d2[3];
  for (int k = 0; k < 3; ++k) {
    xe[k] = 0.25 * ((1.0 - xi1) * (1.0 - xi2) * v1[k] + (1.0 + xi1) * (1.0 - xi2) * v2[k] +
                    (1.0 + xi1) * (1.0 + xi2) * v3[k] + (1.0 - xi1) * (1.0 + xi2) * v4[k]);
    d1[k] = 0.25 * (-(1.0 - xi2) * v1[k] + (1.0 - xi2) * v2[k] + (1.0 + xi2) * v3[k] -
                    (1.0 + xi2) * v4[k]);
    d2[k] = 0.25 * (-(1.0 - xi1) * v1[k] - (1.0 + xi1) * v2[k] + (1.0 + xi1) * v3[k] +
                    (1.0 - xi1) * v4[k]);
  }
  const T nx = std::sqrt(xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2]);
  const T p1 = (xe[0] * d1[0] + xe[1] * d1[1] + xe[2] * d1[2]) / (nx * nx);
  const T p2 = (xe[0] * d2[0] + xe[1] * d2[1] + xe[2] * d2[2]) / (nx * nx);
  for (int k = 0; k < 3; ++k) {
    A[k][0] = radius / nx * (d1[k] - p1 * xe[k]);
    A[k][1] = radius / nx * (d2[k] - p2 * xe[k]);
  }
}

template <typename T>
static void metric_covariant(const double* v1, const double* v2, const double* v3,
                             const double* v4, T xi1, T xi2, double radius, T G[3]) {
  T A[3][2];
  basis_covariant(v1, v2, v3, v4, xi1, xi2, radius, A);
  G[0] = A[0][0] * A[0][0] + A[1][0] * A[1][0] + A[2][0] * A[2][0];
  G[1] = A[0][0] * A[0][1] + A[1][0] * A[1][1] + A[2][0] * A[2][1];
  G[2] = A[0][1] * A[0][1] + A[1][1] * A[1][1] + A[2][1] * A[2][1];
}

// calc_christoffel_symbols (:430-453): first kind from dG/dxi, raised with G^-1
static void christoffel(const double dG1[3], const double dG2[3], const double Gc[3], double out[6]) {
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

bool build_cubed_sphere_2d(Mesh2D& m, const Basis& b, int nf, double radius, int christoffel_mode,
                           std::string* err) {
  const int n = m.n;
  const size_t n2 = (size_t)n * n;
  m.n_elements = 6 * nf * nf;
  const int E = m.n_elements;
  m.node_coordinates.assign(3 * n2 * E, 0.0);
  m.aux.assign(26 * n2 * E, 0.0);
  const double dx = 2.0 / nf;
  for (int dir = 1; dir <= 6; ++dir)
    for (int cy = 0; cy < nf; ++cy)
      for (int cx = 0; cx < nf; ++cx) {
        const int e = cx + nf * (cy + nf * (dir - 1));
        const double xo = -1.0 + cx * dx + dx / 2, yo = -1.0 + cy * dx + dx / 2;
        double v[4][3];
        cubed_sphere_point(xo - dx / 2, yo - dx / 2, radius, dir, v[0]);
        cubed_sphere_point(xo + dx / 2, yo - dx / 2, radius, dir, v[1]);
        cubed_sphere_point(xo + dx / 2, yo + dx / 2, radius, dir, v[2]);
        cubed_sphere_point(xo - dx / 2, yo + dx / 2, radius, dir, v[3]);
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const double a = b.x[i], c = b.x[j];
            double xb[3], nb = 0.0;
            for (int k = 0; k < 3; ++k) {
              xb[k] = 0.25 * ((1 - a) * (1 - c) * v[0][k] + (1 + a) * (1 - c) * v[1][k] +
                              (1 + a) * (1 + c) * v[2][k] + (1 - a) * (1 + c) * v[3][k]);
              nb += xb[k] * xb[k];
            }
            nb = std::sqrt(nb);
            for (int k = 0; k < 3; ++k)
              m.node_coordinates[3 * (i + n * j + n2 * e) + k] = radius * xb[k] / nb;
          }
      }
  // auxiliary variables from the corner nodes (as the reference reads them
  // from tree_node_coordinates) and the radius of the first node
  const double* X = m.node_coordinates.data();
  const double R = std::sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  for (int e = 0; e < E; ++e) {
    const double* v1 = X + 3 * (0 + n2 * e);
    const double* v2 = X + 3 * ((n - 1) + n2 * e);
    const double* v3 = X + 3 * ((n - 1) + n * (n - 1) + n2 * e);
    const double* v4 = X + 3 * (n * (n - 1) + n2 * e);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double* a = &m.aux[26 * (i + n * j + n2 * e)];
        double A[3][2];
        basis_covariant<double>(v1, v2, v3, v4, b.x[i], b.x[j], R, A);
        for (int k = 0; k < 3; ++k) { a[k] = A[k][0]; a[3 + k] = A[k][1]; }
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
        a[12] = std::sqrt(det);
        a[13] = G11; a[14] = G12; a[15] = G22;
        a[16] = Gc11; a[17] = Gc12; a[18] = Gc22;
        a[19] = 0.0;  // bottom topography: set by the caller
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double* a = &m.aux[26 * (i + n * j + n2 * e)];
        double dG1[3], dG2[3];
        if (christoffel_mode == 0) {
          // exact derivatives of the analytic metric by complex-step differentiation
          using C = std::complex<double>;
          const double h = 1e-30;
          C G[3];
          metric_covariant<C>(v1, v2, v3, v4, C(b.x[i], h), C(b.x[j], 0.0), R, G);
          for (int k = 0; k < 3; ++k) dG1[k] = G[k].imag() / h;
          metric_covariant<C>(v1, v2, v3, v4, C(b.x[i], 0.0), C(b.x[j], h), R, G);
          for (int k = 0; k < 3; ++k) dG2[k] = G[k].imag() / h;
        } else {
          // collocation derivative of the nodal covariant metric (:380-416)
          for (int k = 0; k < 3; ++k) { dG1[k] = 0.0; dG2[k] = 0.0; }
          for (int q = 0; q < n; ++q) {
            const double* ai = &m.aux[26 * (q + n * j + n2 * e)];
            const double* aj = &m.aux[26 * (i + n * q + n2 * e)];
            for (int k = 0; k < 3; ++k) {
              dG1[k] += b.D[i * n + q] * ai[13 + k];
              dG2[k] += b.D[j * n + q] * aj[13 + k];
            }
          }
        }
        christoffel(dG1, dG2, a + 16, a + 20);
      }
  }
  // connectivity: match edges by their node coordinates
  const double tol = 1e-9 * R;
  auto face_node = [&](int e, int f, int k) -> const double* {
    const int i = f == 0 ? 0 : (f == 1 ? n - 1 : k);
    const int j = f == 2 ? 0 : (f == 3 ? n - 1 : k);
    return X + 3 * (i + n * j + n2 * e);
  };
  std::map<std::array<long long, 3>, std::vector<std::pair<int, int>>> faces;
  const double quant = 1e-6 * R;
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 4; ++f) {
      double c[3] = {0, 0, 0};
      const double* p0 = face_node(e, f, 0);
      const double* p1 = face_node(e, f, n - 1);
      for (int k = 0; k < 3; ++k) c[k] = 0.5 * (p0[k] + p1[k]);
      std::array<long long, 3> key{std::llround(c[0] / quant), std::llround(c[1] / quant),
                                   std::llround(c[2] / quant)};
      faces[key].push_back({e, f});
    }
  std::vector<std::array<int, 4>> ifaces;
  for (auto& kv : faces) {
    auto v = kv.second;
    if (v.size() != 2) {
      if (err) *err = "cubed-sphere edge without exactly two elements";
      return false;
    }
    std::sort(v.begin(), v.end());
    ifaces.push_back({v[0].first, v[0].second, v[1].first, v[1].second});
  }
  std::sort(ifaces.begin(), ifaces.end());
  m.interface_neighbor_ids.clear();
  m.interface_node_indices.clear();
  auto codes = [](int f, int surf, int32_t out[2]) {
    if (f == 0) { out[0] = kBegin; out[1] = surf; }
    else if (f == 1) { out[0] = kEnd; out[1] = surf; }
    else if (f == 2) { out[0] = surf; out[1] = kBegin; }
    else { out[0] = surf; out[1] = kEnd; }
  };
  for (auto& t : ifaces) {
    const int ep = t[0], fp = t[1], es = t[2], fs = t[3];
    bool fwd = true;
    for (int k = 0; k < n; ++k) {
      const double* a = face_node(ep, fp, k);
      const double* c = face_node(es, fs, k);
      const double d = std::sqrt((a[0] - c[0]) * (a[0] - c[0]) + (a[1] - c[1]) * (a[1] - c[1]) +
                                 (a[2] - c[2]) * (a[2] - c[2]));
      if (d > tol) { fwd = false; break; }
    }
    if (!fwd) {
      for (int k = 0; k < n; ++k) {
        const double* a = face_node(ep, fp, k);
        const double* c = face_node(es, fs, n - 1 - k);
        const double d = std::sqrt((a[0] - c[0]) * (a[0] - c[0]) + (a[1] - c[1]) * (a[1] - c[1]) +
                                   (a[2] - c[2]) * (a[2] - c[2]));
        if (d > tol) {
          if (err) *err = "edge nodes do not match in either orientation";
          return false;
        }
      }
    }
    int32_t cp[2], cs[2];
    codes(fp, kIFwd, cp);
    codes(fs, fwd ? kIFwd : kIBwd, cs);
    m.interface_neighbor_ids.push_back(ep);
    m.interface_neighbor_ids.push_back(es);
    m.interface_node_indices.push_back(cp[0]);
    m.interface_node_indices.push_back(cp[1]);
    m.interface_node_indices.push_back(cs[0]);
    m.interface_node_indices.push_back(cs[1]);
  }
  m.n_interfaces = (int)ifaces.size();
  return true;
}

}  // namespace atmo
