// Host-code checks of libatmo_hip under AddressSanitizer and
// UndefinedBehaviorSanitizer (CPU only; no device is touched): the mesh
// builders (mesh2d.cpp, mesh3d.cpp, basis.cpp) at every polydeg, and the
// connectivity validation of atmo_create / atmo_swe_create / atmo_rainy_create
// on the invalid configurations and malformed connectivity the ctypes tests
// use, each of which must be rejected before any device call.  Exits nonzero
// when an expectation fails; the sanitizers abort on a memory or UB error.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/atmo_hip.h"

static int failures = 0;
#define EXPECT(cond, what)                                   \
  do {                                                       \
    if (!(cond)) {                                           \
      std::fprintf(stderr, "FAILED: %s\n", what);            \
      ++failures;                                            \
    }                                                        \
  } while (0)

static void meshes() {
  for (int p = 1; p <= 7; ++p) {
    const int n = p + 1;
    std::vector<double> x(n), w(n), D(n * n), Ds(n * n);
    EXPECT(atmo_lgl_basis(n, x.data(), w.data(), D.data(), Ds.data()) == ATMO_OK, "lgl basis");
    const int32_t trees[3] = {2, 3, 2}, per[3] = {1, 0, 1};
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 3, 2};
    atmo_mesh3d* m = nullptr;
    EXPECT(atmo_mesh3d_cartesian(p, trees, lo, hi, per, &m) == ATMO_OK, "cartesian mesh");
    int32_t E = 0, I = 0, B = 0;
    atmo_mesh3d_sizes(m, &E, &I, &B);
    const double *xyz, *cv, *ij;
    const int32_t *ids, *idx, *be, *bidx, *bk;
    EXPECT(atmo_mesh3d_arrays(m, &xyz, &cv, &ij, &ids, &idx, &be, &bidx, &bk) == ATMO_OK,
           "cartesian arrays");
    double s = 0;
    for (int q = 0; q < 9 * n * n * n * E; ++q) s += cv[q];
    EXPECT(s == s, "finite metric terms");
    atmo_mesh3d_destroy(m);
    EXPECT(atmo_mesh3d_cubed_sphere_shell(p, 2, 2, 6.371e6, 3.0e4, &m) == ATMO_OK, "shell mesh");
    atmo_mesh3d_destroy(m);
    for (int col = 0; col <= 1; ++col) {
      atmo_mesh2d* m2 = nullptr;
      EXPECT(atmo_mesh2d_cubed_sphere(p, 2, 6.371e6, col, &m2) == ATMO_OK, "cubed sphere 2D");
      atmo_mesh2d_destroy(m2);
    }
  }
  atmo_mesh3d* m = nullptr;
  EXPECT(atmo_mesh3d_cubed_sphere_shell(16, 2, 2, 1.0, 0.1, &m) != ATMO_OK && !m, "mesh polydeg 16 refused");
  EXPECT(atmo_lgl_basis(1, nullptr, nullptr, nullptr, nullptr) != ATMO_OK, "n = 1 refused");
}

static void euler3d() {
  const int32_t trees[3] = {2, 2, 2}, per[3] = {1, 1, 1};
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  atmo_mesh3d* m = nullptr;
  atmo_mesh3d_cartesian(2, trees, lo, hi, per, &m);
  int32_t E = 0, I = 0, B = 0;
  atmo_mesh3d_sizes(m, &E, &I, &B);
  const double *xyz, *cv, *ij;
  const int32_t *ids, *idx, *be, *bidx, *bk;
  atmo_mesh3d_arrays(m, &xyz, &cv, &ij, &ids, &idx, &be, &bidx, &bk);
  atmo_config good{};
  good.equations = ATMO_EQ_EULER_ENERGY_GRAVITY_3D;
  good.polydeg = 2;
  good.n_elements = E;
  good.c_p = 1004.0;
  good.c_v = 717.0;
  good.gravity = 9.81;
  // invalid configurations
  for (int k = 0; k < 7; ++k) {
    atmo_config c = good;
    if (k == 0) c.polydeg = 9;
    if (k == 1) c.volume_flux = 7;
    if (k == 2) c.surface_flux = 9;
    if (k == 3) c.precision = 5;
    if (k == 4) c.c_v = 2000.0;
    if (k == 5) c.n_elements = 0;
    if (k == 6) c.source_terms = 4;
    atmo_geometry3d g{};
    atmo_handle* h = nullptr;
    EXPECT(atmo_create(&c, &g, &h) == ATMO_ERR_INVALID && !h, "invalid 3D configuration refused");
  }
  // malformed connectivity: a dropped interface, a bad node-index code, a
  // neighbour outside the mesh, a face used twice
  std::vector<int32_t> i2(ids, ids + 2 * I), x2(idx, idx + 6 * I);
  for (int k = 0; k < 4; ++k) {
    std::vector<int32_t> a = i2, b = x2;
    int nif = I;
    if (k == 0) nif = I - 1;
    if (k == 1) b[0] = 9;
    if (k == 2) a[0] = 1000;
    if (k == 3) { a[2] = a[0]; std::memcpy(&b[6], &b[0], 6 * sizeof(int32_t)); }
    atmo_geometry3d g{};
    g.contravariant_vectors = cv;
    g.inverse_jacobian = ij;
    g.n_interfaces = nif;
    g.interface_neighbor_ids = a.data();
    g.interface_node_indices = b.data();
    atmo_handle* h = nullptr;
    EXPECT(atmo_create(&good, &g, &h) == ATMO_ERR_INVALID && !h, "malformed 3D connectivity refused");
  }
  atmo_mesh3d_destroy(m);
}

static void covariant() {
  atmo_mesh2d* m = nullptr;
  atmo_mesh2d_cubed_sphere(3, 2, 1.0, 0, &m);
  int32_t E = 0, I = 0;
  atmo_mesh2d_sizes(m, &E, &I);
  const double *xyz, *aux;
  const int32_t *ids, *idx;
  atmo_mesh2d_arrays(m, &xyz, &aux, &ids, &idx);
  atmo_swe_config good{};
  good.polydeg = 3;
  good.n_elements = E;
  good.source_terms = 1;
  good.gravity = 9.81;
  good.rotation_rate = 7.292e-5;
  atmo_geometry2d g{};
  g.aux_node_vars = aux;
  g.node_coordinates = xyz;
  g.n_interfaces = I;
  g.interface_neighbor_ids = ids;
  g.interface_node_indices = idx;
  for (int k = 0; k < 8; ++k) {
    atmo_swe_config c = good;
    if (k == 0) c.polydeg = 8;
    if (k == 1) c.surface_flux = 4;
    if (k == 2) c.surface_noncons = 2;
    if (k == 3) c.precision = 1;
    if (k == 4) c.gravity = 0.0;
    if (k == 5) c.n_elements = 0;
    if (k == 6) c.source_terms = 3;
    if (k == 7) { c.volume_flux = 2; c.surface_flux = 1; c.source_terms = 2; }
    atmo_handle* h = nullptr;
    EXPECT(atmo_swe_create(&c, &g, &h) == ATMO_ERR_INVALID && !h, "invalid covariant configuration refused");
  }
  std::vector<int32_t> i2(ids, ids + 2 * I), x2(idx, idx + 4 * I);
  for (int k = 0; k < 5; ++k) {
    std::vector<int32_t> a = i2, b = x2;
    atmo_geometry2d gg = g;
    gg.interface_neighbor_ids = a.data();
    gg.interface_node_indices = b.data();
    const int32_t first_a[2] = {0, E - 1}, first_b[3] = {0, 12, E};
    if (k == 0) gg.n_interfaces = I - 1;
    if (k == 1) a[0] = E;
    if (k == 2) b[0] = 7;
    if (k == 3) { gg.n_members = 1; gg.member_first_element = first_a; }
    if (k == 4) { gg.n_members = 2; gg.member_first_element = first_b; }
    atmo_handle* h = nullptr;
    EXPECT(atmo_swe_create(&good, &gg, &h) == ATMO_ERR_INVALID && !h, "malformed covariant connectivity refused");
  }
  atmo_mesh2d_destroy(m);
}

static void rainy() {
  atmo_rainy_config good{};
  good.polydeg = 3;
  good.cells_x = 4;
  good.cells_y = 3;
  good.x_max = 1.0;
  good.y_max = 1.0;
  good.volume_integral = ATMO_RAINY_VOL_FLUX_EC_RAIN;
  good.newton_tolerance = 1e-9;
  good.newton_max_iterations = 20;
  for (int k = 0; k < 5; ++k) {
    atmo_rainy_config c = good;
    if (k == 0) c.polydeg = 0;
    if (k == 1) c.cells_x = 0;
    if (k == 2) c.x_max = 0.0;
    if (k == 3) c.volume_integral = 5;
    if (k == 4) c.source_terms = 9;
    atmo_handle* h = nullptr;
    EXPECT(atmo_rainy_create(&c, &h) == ATMO_ERR_INVALID && !h, "invalid rainy configuration refused");
  }
}

int main() {
  meshes();
  euler3d();
  covariant();
  rainy();
  if (failures) std::fprintf(stderr, "%d expectation(s) failed\n", failures);
  else std::printf("host checks passed\n");
  return failures ? 1 : 0;
}
