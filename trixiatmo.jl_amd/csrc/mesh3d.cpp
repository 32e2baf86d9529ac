// Host-side generation of curved hexahedral meshes, their metric terms and
// their Trixi-style face connectivity.  This replaces, for the two fixed
// topologies of this library, p4est's connectivity construction plus Trixi's
// P4estMesh{3} element/interface/boundary initialisation.
//
//  * Cartesian brick (Trixi `P4estMesh(trees_per_dimension; coordinates_min,
//    coordinates_max, periodicity)`): affine tree mapping.
//  * Cubed-sphere shell: src/meshes/p4est_cubed_sphere_topography.jl:114-163
//    (constructor), :596-620 (cubed_sphere_mapping_topography with the
//    perfect-sphere topography and the Gal-Chen vertical coordinate, which is
//    then the identity), :622-670 (tree node coordinates at LGL nodes).
//    Trees are numbered column first (layer fastest), as at :271-283; within
//    a panel this library uses row-major cell order instead of the Gilbert
//    curve, which permutes elements and changes no value.
//  * Metric terms: Jacobian matrix by the collocation derivative, contravariant
//    vectors Ja^i in Kopriva's invariant curl form (Kopriva 2006, J. Sci.
//    Comput. 26, "Metric identities and the discontinuous spectral element
//    method on curvilinear meshes"), 1/J from the determinant; this is the
//    MetricTermsInvariantCurl choice named at
//    src/semidiscretization/metric_terms.jl:25.
//  * Connectivity: faces are matched by node coordinates; the primary side of
//    an interface is the (element, face) pair that sorts first and is indexed
//    (:i_forward, :j_forward); the secondary side's node-index codes are the
//    ones under which every surface node coincides.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "atmo_internal.hpp"

namespace atmo {

static inline size_t nidx(int n, int i, int j, int k, int e) {
  return (size_t)i + n * ((size_t)j + n * ((size_t)k + n * (size_t)e));
}

void compute_metric_terms(Mesh3D& m, const Basis& b) {
  const size_t n3 = (size_t)m.n * m.n * m.n;
  m.jacobian_matrix.assign(9 * n3 * m.n_elements, 0.0);
  m.contravariant.assign(9 * n3 * m.n_elements, 0.0);
  m.inverse_jacobian.assign(n3 * m.n_elements, 0.0);
  compute_metric_terms(m.n, m.n_elements, m.node_coordinates.data(), b.D.data(),
                       m.jacobian_matrix.data(), m.contravariant.data(), m.inverse_jacobian.data());
}

void compute_metric_terms(int n, int E, const double* X, const double* D, double* jacobian_matrix,
                          double* contravariant, double* inverse_jacobian) {
  auto at = [&](int c, int i, int j, int k, int e) { return X[3 * nidx(n, i, j, k, e) + c]; };
  auto step = [](int d, int& i, int& j, int& k) -> int& { return d == 0 ? i : (d == 1 ? j : k); };
  for (int e = 0; e < E; ++e) {
    // Jacobian matrix jac[c + 3 d] = dX_c / dxi_d by the collocation derivative
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double* J = &jacobian_matrix[9 * nidx(n, i, j, k, e)];
          for (int d = 0; d < 3; ++d) {
            int ii = i, jj = j, kk = k;
            int& ix = step(d, ii, jj, kk);
            const int own = ix;
            for (int c = 0; c < 3; ++c) {
              double s = 0.0;
              for (int q = 0; q < n; ++q) {
                ix = q;
                s += D[own * n + q] * at(c, ii, jj, kk, e);
              }
              ix = own;
              J[c + 3 * d] = s;
            }
          }
        }
    // invariant curl form:
    // Ja^d_c = 1/2 [ (X_m dX_l/dxi_{d+2} - X_l dX_m/dxi_{d+2})_{,xi_{d+1}}
    //              - (X_m dX_l/dxi_{d+1} - X_l dX_m/dxi_{d+1})_{,xi_{d+2}} ]
    // with (c, m, l) cyclic and d, d+1, d+2 cyclic.
    for (int d = 0; d < 3; ++d) {
      const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      for (int c = 0; c < 3; ++c) {
        const int mm = (c + 1) % 3, ll = (c + 2) % 3;
        for (int k = 0; k < n; ++k)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
              double term1 = 0.0, term2 = 0.0;
              {  // derivative along d1 of (X_m X_l,d2 - X_l X_m,d2)
                int ii = i, jj = j, kk = k;
                int& ix = step(d1, ii, jj, kk);
                const int own = ix;
                for (int q = 0; q < n; ++q) {
                  ix = q;
                  const size_t nd = nidx(n, ii, jj, kk, e);
                  const double* Jq = &jacobian_matrix[9 * nd];
                  const double g = X[3 * nd + mm] * Jq[ll + 3 * d2] - X[3 * nd + ll] * Jq[mm + 3 * d2];
                  term1 += D[own * n + q] * g;
                }
              }
              {  // derivative along d2 of (X_m X_l,d1 - X_l X_m,d1)
                int ii = i, jj = j, kk = k;
                int& ix = step(d2, ii, jj, kk);
                const int own = ix;
                for (int q = 0; q < n; ++q) {
                  ix = q;
                  const size_t nd = nidx(n, ii, jj, kk, e);
                  const double* Jq = &jacobian_matrix[9 * nd];
                  const double g = X[3 * nd + mm] * Jq[ll + 3 * d1] - X[3 * nd + ll] * Jq[mm + 3 * d1];
                  term2 += D[own * n + q] * g;
                }
              }
              contravariant[9 * nidx(n, i, j, k, e) + c + 3 * d] = 0.5 * (term1 - term2);
            }
      }
    }
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double* J = &jacobian_matrix[9 * nidx(n, i, j, k, e)];
          // J[c + 3 d]; determinant by the rule of Sarrus
          const double det = J[0] * J[4] * J[8] + J[3] * J[7] * J[2] + J[6] * J[1] * J[5] -
                             J[6] * J[4] * J[2] - J[3] * J[1] * J[8] - J[0] * J[7] * J[5];
          inverse_jacobian[nidx(n, i, j, k, e)] = 1.0 / det;
        }
  }
}

// element node index of surface node (a, b) under a node-index code
static inline int code_index(int code, int a, int b, int n) {
  switch (code) {
    case kBegin: return 0;
    case kEnd: return n - 1;
    case kIFwd: return a;
    case kIBwd: return n - 1 - a;
    case kJFwd: return b;
    default: return n - 1 - b;
  }
}

static void face_codes(int face, int c1, int c2, int32_t out[3]) {
  const int d = face / 2;
  int r = 0;
  for (int q = 0; q < 3; ++q) {
    if (q == d) out[q] = (face % 2 == 0) ? kBegin : kEnd;
    else out[q] = (r++ == 0) ? c1 : c2;
  }
}

bool build_connectivity(Mesh3D& m, std::string* err) {
  const int n = m.n, E = m.n_elements;
  const double* X = m.node_coordinates.data();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t q = 0; q < m.node_coordinates.size() / 3; ++q)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], X[3 * q + c]);
      hi[c] = std::max(hi[c], X[3 * q + c]);
    }
  const double diam = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const double tol = 1e-8 * diam;
  const double quant = 1e-5 * diam;
  auto wrap = [&](int c, double v) {
    if (m.period[c] > 0) v -= m.period[c] * std::floor(v / m.period[c] + 1e-9);
    return v;
  };
  auto dist = [&](const double* a, const double* b) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      double dd = a[c] - b[c];
      if (m.period[c] > 0) dd -= m.period[c] * std::round(dd / m.period[c]);
      s += dd * dd;
    }
    return std::sqrt(s);
  };
  auto face_node = [&](int e, int face, int a, int b) -> const double* {
    int32_t codes[3];
    face_codes(face, kIFwd, kJFwd, codes);
    return X + 3 * nidx(n, code_index(codes[0], a, b, n), code_index(codes[1], a, b, n),
                        code_index(codes[2], a, b, n), e);
  };
  std::map<std::array<long long, 3>, std::vector<std::pair<int, int>>> faces;
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 6; ++f) {
      double ctr[3] = {0, 0, 0};
      const int cs[4][2] = {{0, 0}, {n - 1, 0}, {0, n - 1}, {n - 1, n - 1}};
      for (auto& ab : cs) {
        const double* p = face_node(e, f, ab[0], ab[1]);
        for (int c = 0; c < 3; ++c) ctr[c] += 0.25 * p[c];
      }
      std::array<long long, 3> key;
      for (int c = 0; c < 3; ++c) key[c] = std::llround(wrap(c, ctr[c]) / quant);
      faces[key].push_back({e, f});
    }
  m.interface_neighbor_ids.clear();
  m.interface_node_indices.clear();
  m.boundary_element.clear();
  m.boundary_node_indices.clear();
  m.boundary_kind.clear();
  std::vector<std::tuple<int, int, int, int>> ifaces;  // (ep, fp, es, fs)
  std::vector<std::pair<int, int>> bfaces;
  for (auto& kv : faces) {
    auto v = kv.second;
    std::sort(v.begin(), v.end());
    if (v.size() == 2) ifaces.emplace_back(v[0].first, v[0].second, v[1].first, v[1].second);
    else if (v.size() == 1) bfaces.push_back(v[0]);
    else { if (err) *err = "face matched more than two elements"; return false; }
  }
  std::sort(ifaces.begin(), ifaces.end());
  std::sort(bfaces.begin(), bfaces.end());
  const int opts[8][2] = {{kIFwd, kJFwd}, {kIBwd, kJFwd}, {kIFwd, kJBwd}, {kIBwd, kJBwd},
                          {kJFwd, kIFwd}, {kJBwd, kIFwd}, {kJFwd, kIBwd}, {kJBwd, kIBwd}};
  for (auto& t : ifaces) {
    const int ep = std::get<0>(t), fp = std::get<1>(t), es = std::get<2>(t), fs = std::get<3>(t);
    int32_t cp[3], cs[3];
    face_codes(fp, kIFwd, kJFwd, cp);
    bool found = false;
    for (auto& o : opts) {
      face_codes(fs, o[0], o[1], cs);
      bool ok = true;
      for (int b = 0; b < n && ok; ++b)
        for (int a = 0; a < n && ok; ++a) {
          const double* pp = X + 3 * nidx(n, code_index(cp[0], a, b, n), code_index(cp[1], a, b, n),
                                          code_index(cp[2], a, b, n), ep);
          const double* ps = X + 3 * nidx(n, code_index(cs[0], a, b, n), code_index(cs[1], a, b, n),
                                          code_index(cs[2], a, b, n), es);
          if (dist(pp, ps) > tol) ok = false;
        }
      if (ok) { found = true; break; }
    }
    if (!found) { if (err) *err = "no orientation matches the face nodes"; return false; }
    m.interface_neighbor_ids.push_back(ep);
    m.interface_neighbor_ids.push_back(es);
    for (int q = 0; q < 3; ++q) m.interface_node_indices.push_back(cp[q]);
    for (int q = 0; q < 3; ++q) m.interface_node_indices.push_back(cs[q]);
  }
  for (auto& bf : bfaces) {
    int32_t cb[3];
    face_codes(bf.second, kIFwd, kJFwd, cb);
    m.boundary_element.push_back(bf.first);
    for (int q = 0; q < 3; ++q) m.boundary_node_indices.push_back(cb[q]);
    // -z face of a shell column is :inside, +z is :outside
    m.boundary_kind.push_back(bf.second == 4 ? 0 : (bf.second == 5 ? 1 : 2));
  }
  m.n_interfaces = (int)ifaces.size();
  m.n_boundaries = (int)bfaces.size();
  return true;
}

// Cartesian brick: trees (cx, cy, cz), tree node = affine image of the LGL nodes
void build_cartesian(Mesh3D& m, const Basis& b, const int trees[3], const double xmin[3],
                     const double xmax[3], const int periodic[3]) {
  const int n = m.n;
  m.n_elements = trees[0] * trees[1] * trees[2];
  m.node_coordinates.assign(3 * (size_t)n * n * n * m.n_elements, 0.0);
  for (int c = 0; c < 3; ++c) m.period[c] = periodic[c] ? (xmax[c] - xmin[c]) : 0.0;
  int e = 0;
  for (int cz = 0; cz < trees[2]; ++cz)
    for (int cy = 0; cy < trees[1]; ++cy)
      for (int cx = 0; cx < trees[0]; ++cx, ++e) {
        const int cell[3] = {cx, cy, cz};
        for (int k = 0; k < n; ++k)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
              const int ijk[3] = {i, j, k};
              for (int c = 0; c < 3; ++c) {
                // reference coordinate on [-1, 1] of the whole brick, then affine map
                const double dxi = 2.0 / trees[c];
                const double xi = -1.0 + cell[c] * dxi + 0.5 * dxi + 0.5 * dxi * b.x[ijk[c]];
                m.node_coordinates[3 * nidx(n, i, j, k, e) + c] =
                    xmin[c] + 0.5 * (xi + 1.0) * (xmax[c] - xmin[c]);
              }
            }
      }
}

// p4est_cubed_sphere_topography.jl:596-620 with zero topography
static void shell_mapping(double xi, double eta, double zeta, double inner_radius,
                          double thickness, int direction, double out[3]) {
  const double x = std::tan(xi * M_PI / 4.0), y = std::tan(eta * M_PI / 4.0);
  double cube[3];
  switch (direction) {
    case 1: cube[0] = -y; cube[1] = -x; cube[2] = -1; break;
    case 2: cube[0] = 1; cube[1] = x; cube[2] = y; break;
    case 3: cube[0] = -x; cube[1] = 1; cube[2] = y; break;
    case 4: cube[0] = -1; cube[1] = -x; cube[2] = y; break;
    case 5: cube[0] = x; cube[1] = -1; cube[2] = y; break;
    default: cube[0] = x; cube[1] = y; cube[2] = 1; break;
  }
  const double r = std::sqrt(1.0 + x * x + y * y);
  const double z = thickness * 0.5 * (zeta + 1.0);
  const double s = (inner_radius + z) / r;
  for (int c = 0; c < 3; ++c) out[c] = s * cube[c];
}

void build_cubed_sphere_shell(Mesh3D& m, const Basis& b, int trees_per_face, int layers,
                              double inner_radius, double thickness) {
  const int n = m.n, nf = trees_per_face;
  m.n_elements = 6 * nf * nf * layers;
  m.node_coordinates.assign(3 * (size_t)n * n * n * m.n_elements, 0.0);
  const double dx = 2.0 / nf, dz = 2.0 / layers;
  for (int dir = 1; dir <= 6; ++dir)
    for (int cy = 0; cy < nf; ++cy)
      for (int cx = 0; cx < nf; ++cx)
        for (int cz = 0; cz < layers; ++cz) {
          const int e = cz + layers * ((cx + nf * cy) + nf * nf * (dir - 1));
          const double xo = -1.0 + cx * dx + dx / 2, yo = -1.0 + cy * dx + dx / 2,
                       zo = -1.0 + cz * dz + dz / 2;
          for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
              for (int i = 0; i < n; ++i)
                shell_mapping(xo + dx / 2 * b.x[i], yo + dx / 2 * b.x[j], zo + dz / 2 * b.x[k],
                              inner_radius, thickness, dir,
                              &m.node_coordinates[3 * nidx(n, i, j, k, e)]);
        }
}

}  // namespace atmo
