// Launch interface of the covariant shallow-water kernels (internal).
#ifndef ATMO_SWE_KERNELS_HPP
#define ATMO_SWE_KERNELS_HPP

#include <hip/hip_runtime.h>

namespace atmo {

struct SweLaunch {
  int n, n_elements;
  int volume_flux;      // 0 flux_ec, 1 flux_central, 2 weak form of the standard form
  int surface_flux;     // 0 FluxPlusDissipation(flux_ec, LLF), 1 flux_lax_friedrichs, 2 flux_ec,
                        // 3 flux_lax_friedrichs of the standard form
  int surface_noncons;  // 0 surface_simplified, 1 flux_nonconservative_ec (split form only)
  int source_terms;     // 1 source_terms_geometric_coriolis, 2 the same of the standard form
  double gravity, rotation_rate;
  double dsplit[64];    // Dsplit[i * n + m]
  double dhat[64];      // Trixi's derivative_hat[i * n + m] (weak form)
  double inv_w;         // 1 / w_end
  const double* u;
  double* out;
  double* k;
  const double* aux;    // [26, n, n, E]
  const double* xyz;    // [3, n, n, E]
  // the stage kernel's geometry: geo[c * n_nodes + node] for c = G_11, G_12,
  // G_22, h_s, the six Christoffel symbols (aux rows 13-15, 19-25), z / r and
  // J (G^-1, and in FP64 J, are recomputed from G where they are needed); fxf[((e * 4 + f) * n + k) * 4 + 2 r + s] = (B_own A_nbr)_rs, the
  // 2 x 2 map of a neighbour's contravariant momentum into this element's
  // contravariant basis at face node k of face f
  const double* geo;
  const double* fxf;
  // the same two arrays rounded to FP32 (ATMO_PREC_FP32_GEOMETRY; NULL in
  // FP64): the stage and the max-speed kernels then read these, and every
  // operation stays FP64
  const float* geo32;
  const float* fxf32;
  // FP32 geometry with bottom topography: h_s [n_nodes] in FP64 (the FP32
  // row 3 is then not read), NULL otherwise
  const double* hs64;
  long long n_nodes;
  const int* nbr;       // [E, 4, n] neighbour global node of each face node
  int mode;             // 0 rhs, 1 RK stage
  double rk_a, rk_dt, rk_b;
  const double* dt_member;  // per-member time steps (NULL: rk_dt for all)
  double* speed_partial;    // RK stage, if set: per element the StepsizeCallback sum over
                            // directions of the maximum scaled speed of the new state
  const int* member;        // [E] member of each element
  int topography;           // 0: h_s is zero at every node and its row is not read
  int variant;              // workgroup shape: 0..2 = 1, 2, 4 waves; +3 persistent grid
};
constexpr int kSweVariants = 6;

// init_auxiliary_node_variables! on the device (swe_geometry.hip): aux
// [26,n,n,E] from the node coordinates X [3,n,n,E] (LGL nodes x_lgl[n],
// derivative matrix D[n*n]; topography [n,n,E] or NULL; collocation: the
// Christoffel symbols from the collocation derivative instead of the exact
// metric derivatives; the radius is that of the first node of the element's
// member, member[e] and member_first[m] device arrays, or of the first node
// of X when member is NULL); false for n > 8
bool swe_launch_aux(int n, int n_elements, const double* x_lgl, const double* D, const double* X,
                    const double* topography, int collocation, const int* member,
                    const int* member_first, double* aux, hipStream_t s);
// the stage kernel's compact rows geo [12][nn] from aux and X
void swe_launch_compact_rows(const double* aux, const double* X, long long nn, double* geo,
                             hipStream_t s);
// the face maps fxf [E][4][n][2][2] from aux and the neighbour links nbr [E][4][n]
bool swe_launch_face_maps(int n, int n_elements, const double* aux, const int* nbr, double* fxf,
                          hipStream_t s);
// b[q] = (float) a[q] (euler3d_geometry.hip)
void launch_narrow_to_float(const double* a, float* b, size_t n, hipStream_t s);
constexpr int kSweGeo = 12;

// J = sqrt(det G) and G^-1 = adj(G) / det G from the covariant metric tensor
// G = A^T A (the identities aux rows 12 and 16-18 satisfy: J is the area
// element |a_1 x a_2|); one sequence of operations, with an explicit fused
// multiply-add, on the host and in every kernel, so each use gets the same bits
__host__ __device__ inline void metric_from_gcov(double g11, double g12, double g22, double& J,
                                                 double& c11, double& c12, double& c22) {
  const double det = fma(g11, g22, -(g12 * g12));
  const double inv = 1.0 / det;
  J = sqrt(det);
  c11 = g22 * inv;
  c12 = -g12 * inv;
  c22 = g11 * inv;
}

bool swe_launch_stage(const SweLaunch& L, hipStream_t s);
// per-member max over elements of the summed scaled wave speeds -> out[n_members]
bool swe_launch_member_max_speed(const SweLaunch& L, double* partial, const int* member_first,
                                 int n_members, double* out, hipStream_t s);

// integrals (mass, entropy, dS/dt) -> out[3]; w = device copy of the LGL weights
bool swe_launch_integrals(const SweLaunch& L, const double* du, const double* w, double* partial,
                          double* out, hipStream_t s);

}  // namespace atmo

#endif
