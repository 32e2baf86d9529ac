// The entropy-conserving two-point flux of CompressibleRainyEulerEquations2D
// (flux_ec_rain, src/equations/compressible_rainy_euler_2d.jl:872-961) on
// per-node quantities formed once per node, and the equations' constants:
// shared by the stage kernels and by the tests' device probe.
#ifndef ATMO_RAINY_FLUX_HPP
#define ATMO_RAINY_FLUX_HPP

#include <hip/hip_runtime.h>

#include "euler3d_physics.hpp"

namespace atmo {
namespace rainy {

// CompressibleRainyEulerEquations2D() constants (:77-113)
constexpr double c_l = 4186, c_pd = 1004, c_vd = 717, c_pv = 1885, c_vv = 1424;
constexpr double R_d = c_pd - c_vd, R_v = c_pv - c_vv;
constexpr double ref_sp = 610.7, ref_T = 273.15, L_ref = 2500000;
constexpr double grav = 9.80616, N_0 = 8000000, v_0 = 130;
constexpr double kPi = 3.14159265358979323846;

// (a^2 + b^2) / 2 with both squares rounded before the sum: the value does
// not depend on the order of a and b (a contracted a a + b b rounds only one
// of the squares, and which one follows the order)
__device__ __forceinline__ double half_sum_squares(double a, double b) {
#pragma clang fp contract(off)
  return 0.5 * (a * a + b * b);
}

// the per-node quantities flux_ec_rain reads, formed once per node
struct ECNode {
  double rd, rv, rc, rr, v1, v2, it, vr;  // it = 1 / T, vr from rho_vapour + rho_cloud
  double lrd, lrv, lit;                   // log(rho_dry), log(rho_vapour), log(1 / T)
};

__device__ __forceinline__ void flux_ec_rain(const ECNode& l, const ECNode& r, double n1,
                                             double n2, double* f) {
  const double itl = l.it, itr = r.it;
  const double v1a = 0.5 * (l.v1 + r.v1), v2a = 0.5 * (l.v2 + r.v2);
  const double v1sq = half_sum_squares(l.v1, r.v1), v2sq = half_sum_squares(l.v2, r.v2);
  const double K = 0.5 * (v1sq + v2sq);
  const double vdn = n1 * v1a + n2 * v2a;
  const double vrn = 0.5 * n2 * (l.vr + r.vr);
  const double dry_a = 0.5 * (l.rd + r.rd);
  const double vap_a = 0.5 * (l.rv + r.rv);
  const double cld_a = 0.5 * (l.rc + r.rc);
  const double rain_a = 0.5 * (l.rr + r.rr);
  // logarithmic means from the node logarithms (the series branch where the
  // two values are close), no logarithm per pair
  double dry_log = 0.0, vap_log = 0.0;
  if (l.rd != 0.0 && r.rd != 0.0) dry_log = ln_mean_logs(l.rd, r.rd, l.lrd, r.lrd);
  if (l.rv != 0.0 && r.rv != 0.0) vap_log = ln_mean_logs(l.rv, r.rv, l.lrv, r.lrv);
  const double it_avg = 0.5 * (itl + itr);
  const double it_log = inv_ln_mean_logs(itl, itr, l.lit - r.lit);
  const double p_int = div_nr(1.0, it_avg) * (R_d * dry_a + R_v * vap_a);
  const double f_vap = vap_log * vdn;
  const double f_cld = cld_a * vdn;
  const double f_dry = dry_log * vdn;
  const double f_rain = rain_a * (vdn - vrn);
  const double f_moist = f_vap + f_cld;
  const double f_rho = f_dry + f_moist + f_rain;
  const double f_m1 = f_rho * v1a + p_int * n1;
  const double f_m2 = f_rho * v2a + p_int * n2;
  f[0] = f_dry;
  f[1] = f_moist;
  f[2] = f_rain;
  f[3] = f_m1;
  f[4] = f_m2;
  f[5] = (c_vd * it_log - K) * f_dry + (c_vv * it_log - K + L_ref) * f_vap +
         (c_l * it_log - K) * (f_cld + f_rain) + (v1a * f_m1 + v2a * f_m2);
}

}  // namespace rainy
}  // namespace atmo

#endif
