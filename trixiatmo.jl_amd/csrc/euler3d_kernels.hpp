// Launch interface of the 3D Euler kernels (internal to libatmo_hip).
#ifndef ATMO_EULER3D_KERNELS_HPP
#define ATMO_EULER3D_KERNELS_HPP

#include <hip/hip_runtime.h>

#include "euler3d_physics.hpp"

namespace atmo {

struct E3Launch {
  E3Params q;
  int n, n_elements;
  int volume_flux, volume_noncons, surface_flux, surface_noncons;
  int mode;  // 0 rhs, 1 RK stage
  const double* u;
  double* out;
  double* k;
  const double* cv;
  const double* invJ;
  const int* nbr;  // [E, 6, n*n] neighbour global node id, < 0 boundary
  const int* face_info;  // [E, 6] 0 primary/boundary, 8 | dp secondary of primary direction dp
  double rk_a, rk_dt, rk_b;
  double coriolis;  // 2 Omega, 0 = no source
};

// tables: dsplit_all[9][64] (entry [n][i*n+m]), weights_all[9][8]
hipError_t e3_upload_basis_tables(const double* dsplit_all, const double* weights_all);
bool e3_launch_stage_fp64(const E3Launch& L, hipStream_t s);
// which 0: per-element max scaled speed -> result[0] = max
// which 1: integrals (9 columns) -> result[0..8]
bool e3_launch_reduction(const E3Launch& L, int which, const double* du, double* partial,
                         double* result, hipStream_t s);

}  // namespace atmo

#endif
