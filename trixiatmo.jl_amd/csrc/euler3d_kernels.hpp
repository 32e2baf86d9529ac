// Launch interface of the 3D Euler kernels (internal to libatmo_hip).
#ifndef ATMO_EULER3D_KERNELS_HPP
#define ATMO_EULER3D_KERNELS_HPP

#include <hip/hip_runtime.h>

#include "euler3d_physics.hpp"

namespace atmo {

// hydrostatic background of the mixed path (see euler3d_stage.hpp)
struct Background {
  int on;
  double p0, inv_RT0;
  const float* B;  // [n nodes, 3] FP32 copy of the background's momentum tendency
};

// links of a stage to the resident time loop
struct ResidentLink {
  const double* dt;       // if set, the stage's time step is read here
  double* speed_partial;  // if set (RK stage): per element, StepsizeCallback's
                          // sum over directions of the maximum scaled speed of
                          // the new state, as the max-speed reduction computes it
  const double* cv64;     // FP64 contravariant vectors for that speed
};

struct E3Launch {
  E3Params q;
  int n, n_elements;
  int volume_flux, volume_noncons, surface_flux, surface_noncons;
  int mode;  // 0 rhs, 1 RK stage
  const double* u;
  double* out;
  double* k;              // RK register [5, nodes] (phi has no tendency)
  const double* cv;       // FP64 contravariant vectors (reductions)
  const void* cv_t;       // contravariant vectors in the stage kernel's compute type
  const double* invJ;
  const double* dsplit;   // host copy of Dsplit[n*n]
  const double* weights;  // host copy of the LGL weights[n]
  double inv_w;           // 1 / w_end
  const int* nbr;  // [E, 6, n*n] neighbour global node id, < 0 boundary
  const int* face_info;  // [E, 6] 0 primary/boundary, 8 | dp secondary of primary direction dp
  double rk_a, rk_dt, rk_b;
  ResidentLink res;      // resident time loop: device step size, speeds of the new state
  double coriolis;  // 2 Omega, 0 = no source
  Background bg;
  int epb;          // elements per workgroup of the direct stage form: 1, or 2 where built
};

// whether the stage kernel of this polydeg / precision has a two-element
// workgroup shape (the 3D autotuner's choice)
bool e3_has_two_element_groups(int n, int precision);

// background state from the geopotential of u: (rho_bg, 0, 0, 0, p_bg/(gamma-1), phi)
void e3_launch_background_state(const E3Launch& L, const double* u, double* ubg, hipStream_t s);
// FP32 copy of the momentum rows of du: B[node * 3 + c] = du[node * 6 + 1 + c]
void e3_launch_extract_momentum(const double* du, float* B, long n_nodes, hipStream_t s);

bool e3_launch_stage_fp64(const E3Launch& L, hipStream_t s);
// FP64 arithmetic, contravariant vectors read from the FP32 copy (cv_t)
bool e3_launch_stage_fp64_geo32(const E3Launch& L, hipStream_t s);
bool e3_launch_stage_fp32_kahan(const E3Launch& L, hipStream_t s);
bool e3_launch_stage_fp32(const E3Launch& L, hipStream_t s);
// which 0: per-element max scaled speed -> result[0] = max
// which 1: integrals (9 columns) -> result[0..8]
bool e3_launch_reduction(const E3Launch& L, int which, const double* du, double* partial,
                         double* result, hipStream_t s);
// contravariant vectors [3,3,n,n,n,E] and 1/J [n,n,n,E] from the node
// coordinates [3,n,n,n,E] (D: the n x n derivative matrix); false for n > 8
bool e3_launch_metric_terms(int n, int n_elements, const double* D, const double* xyz,
                            double* cv, double* invJ, hipStream_t s);
// b[q] = (float) a[q]
void launch_narrow_to_float(const double* a, float* b, size_t n, hipStream_t s);

}  // namespace atmo

#endif
