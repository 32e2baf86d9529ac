// Internal (non-ABI) declarations shared by the host-side C++ of libatmo_hip.
#ifndef ATMO_INTERNAL_HPP
#define ATMO_INTERNAL_HPP

#include <cstdint>
#include <string>
#include <vector>

namespace atmo {

// Lobatto-Legendre basis: nodes, weights, derivative matrix D[i*n+j] and the
// flux-differencing matrix Dsplit = 2D + diag(1/w_1, 0, ..., -1/w_n).
struct Basis {
  int n;
  std::vector<double> x, w, D, Dsplit, bary;
  explicit Basis(int n);
};

// Trixi-style node-index codes of P4estInterfaceContainer / BoundaryContainer
enum NodeCode : int32_t {
  kBegin = 0, kEnd = 1, kIFwd = 2, kIBwd = 3, kJFwd = 4, kJBwd = 5
};

// A curved hexahedral mesh with one element per p4est tree (level 0), its
// geometry and its Trixi-style face connectivity.
struct Mesh3D {
  int polydeg = 0, n = 0, n_elements = 0;
  std::vector<double> node_coordinates;  // [3, n, n, n, E]
  std::vector<double> jacobian_matrix;   // [3, 3, n, n, n, E]  (dx_c / dxi_d)
  std::vector<double> contravariant;     // [3, 3, n, n, n, E]  (Ja^d)_c
  std::vector<double> inverse_jacobian;  // [n, n, n, E]
  int n_interfaces = 0, n_boundaries = 0;
  std::vector<int32_t> interface_neighbor_ids;  // [2, I]
  std::vector<int32_t> interface_node_indices;  // [3, 2, I]
  std::vector<int32_t> boundary_element;        // [B]
  std::vector<int32_t> boundary_node_indices;   // [3, B]
  std::vector<int32_t> boundary_kind;           // [B]  0 inside, 1 outside, ...
  double period[3] = {0, 0, 0};                 // > 0: periodic length
};

// A 2D cubed-sphere mesh (one element per tree) with the 26 auxiliary
// variables per node of the covariant form and 2D Trixi-style connectivity.
struct Mesh2D {
  int polydeg = 0, n = 0, n_elements = 0, n_interfaces = 0;
  std::vector<double> node_coordinates;         // [3, n, n, E]
  std::vector<double> aux;                      // [26, n, n, E]
  std::vector<int32_t> interface_neighbor_ids;  // [2, I]
  std::vector<int32_t> interface_node_indices;  // [2, 2, I]
};

// christoffel_mode 0: exact derivatives of the analytic metric; 1: collocation
bool build_cubed_sphere_2d(Mesh2D& m, const Basis& b, int trees_per_face, double radius,
                           int christoffel_mode, std::string* err);

// node coordinates -> Jacobian matrix, invariant-curl metric terms, 1/J
void compute_metric_terms(Mesh3D& m, const Basis& b);
// the same on raw arrays: X [3,n,n,n,E], D [n,n] -> jacobian_matrix
// [3,3,n,n,n,E] (scratch), contravariant [3,3,n,n,n,E], inverse_jacobian [n,n,n,E]
void compute_metric_terms(int n, int n_elements, const double* X, const double* D,
                          double* jacobian_matrix, double* contravariant,
                          double* inverse_jacobian);
// face matching by node coordinates -> interfaces + boundaries
bool build_connectivity(Mesh3D& m, std::string* err);

}  // namespace atmo

#endif
