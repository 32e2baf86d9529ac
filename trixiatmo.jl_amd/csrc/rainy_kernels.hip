// gfx950 kernels of the compressible rainy Euler equations in 2D
// (CompressibleRainyEulerEquations2D, src/equations/compressible_rainy_euler_2d.jl)
// on a uniform Cartesian StructuredMesh, periodic in x, periodic or
// simple-slip-wall in y: the fused rhs!/RK stage with the saturation Newton
// solve of NonlinearSolveDG (src/callbacks_stage/nonlinear_solve_dg2d.jl)
// applied to every updated stage state, the standalone Newton solve with an
// optional MLP seed, the step-size reduction and the integrals.
//
// Layout: u [9, n, n, E] (rho_dry, rho_moist, rho_rain, rho_v1, rho_v2,
// energy, rho_vapour, rho_cloud, T), elements numbered x fastest.  A 64-lane
// workgroup holds EPB = 64 / n^2 elements (one lane per node); states and the
// per-node quantities a two-point flux needs are staged in LDS.
#include <hip/hip_runtime.h>

#include "euler3d_physics.hpp"
#include "rainy_flux.hpp"
#include "rainy_kernels.hpp"
#include "speed_max.hpp"

namespace atmo {

using namespace rainy;

namespace {

constexpr int NV = 9;


// cons2nonlinearsystemsol (:219-260)
__device__ __forceinline__ void nonlinear_sol(const double* u, double& rv, double& rc, double& T) {
  const double rho = u[0] + u[1] + u[2];
  const double inv = 1.0 / rho;
  if (u[1] == 0.0 && u[2] == 0.0) {
    const double v1 = u[3] * inv, v2 = u[4] * inv;
    const double ek = 0.5 * (v1 * v1 + v2 * v2) * rho;
    rv = 0.0;
    rc = 0.0;
    T = (u[5] - ek) / (c_vd * rho);
    return;
  }
  rv = u[6];
  rc = u[7];
  if (rv < 0.0 && rv >= -1e-15) rv = 0.0;
  if (rc < 0.0 && rc >= -1e-15) rc = 0.0;
  T = u[8];
}

// x^(1/8) by three square roots (pow(x, 0.125) of the reference to a few
// ulp, at a fraction of pow's cost)
__device__ __forceinline__ double root8(double x) { return sqrt_(sqrt_(sqrt_(x))); }

// terminal_velocity_rain (:336-351)
__device__ __forceinline__ double terminal_velocity(double rho_moist, double rho_rain) {
  if (rho_rain > 0.0)
    return v_0 * 1.9386213994279082 * root8(div_nr(rho_rain, kPi * (rho_moist + rho_rain) * N_0));
  return 0.0;
}

// saturation_vapour_pressure (:353-375), ref_sp (T / ref_T)^k1 exp(k2 (1 / ref_T
// - 1 / T)), formed as one exponential of the summed exponents
constexpr double kSvp1 = (c_pv - c_l) / R_v;
constexpr double kSvp2 = (L_ref - (c_pv - c_l) * ref_T) / R_v;
__device__ __forceinline__ double svp(double T) {
  return ref_sp * exp(kSvp1 * log(T / ref_T) + kSvp2 * (1 / ref_T - 1 / T));
}

// its derivative (:377-398), ref_sp / ref_T^k1 (k1 T^(k1-1) + k2 T^(k1-2))
// exp(...), which equals svp(T) (k1 + k2 / T) / T
__device__ __forceinline__ double svp_deriv_from(double sv, double T) {
  return sv * (kSvp1 + kSvp2 / T) / T;
}

// quantities of one node used by the two-point and surface fluxes
struct RNode {
  double rd, rm, rr, rho, inv, v1, v2, rv, rc, T, p, c, vr_m, vr_sol, E;
};

__device__ __forceinline__ RNode rnode(const double* u) {
  RNode a;
  a.rd = u[0];
  a.rm = u[1];
  a.rr = u[2];
  a.rho = u[0] + u[1] + u[2];
  a.inv = div_nr(1.0, a.rho);
  a.v1 = u[3] * a.inv;
  a.v2 = u[4] * a.inv;
  a.E = u[5];
  nonlinear_sol(u, a.rv, a.rc, a.T);
  a.p = (R_d * a.rd + R_v * a.rv) * a.T;
  // speed_of_sound (:299-334)
  const double q_v = div_nr(a.rv, a.rd);
  const double q_l = div_nr(a.rc + a.rr, a.rd);
  const double gm = 1 + div_nr(R_d + R_v * q_v, c_vd + c_vv * q_v + c_l * q_l);
  a.c = sqrt_(gm * a.p * a.inv);
  a.vr_m = terminal_velocity(a.rm, a.rr);
  a.vr_sol = terminal_velocity(a.rv + a.rc, a.rr);
  return a;
}

// flux in normal direction (:502-545)
__device__ __forceinline__ void flux_n(const RNode& a, double n1, double n2, double* f) {
  const double vn = a.v1 * n1 + a.v2 * n2;
  const double vrn = a.vr_m * n2;
  f[0] = a.rd * vn;
  f[1] = a.rm * vn;
  f[2] = a.rr * (vn - vrn);
  f[3] = a.rho * vn * a.v1 + a.p * n1 - a.rr * vrn * a.v1;
  f[4] = a.rho * vn * a.v2 + a.p * n2 - a.rr * vrn * a.v2;
  f[5] = (a.E + a.p) * vn - (c_l * a.T + 0.5 * (a.v1 * a.v1 + a.v2 * a.v2)) * a.rr * vrn;
  f[6] = f[7] = f[8] = 0.0;
}

// flux in orientation form (:448-500): orientation 1 has no rain terms
__device__ __forceinline__ void flux_o(const RNode& a, int orientation, double* f) {
  if (orientation == 1) {
    f[0] = a.rd * a.v1;
    f[1] = a.rm * a.v1;
    f[2] = a.rr * a.v1;
    f[3] = a.rho * a.v1 * a.v1 + a.p;
    f[4] = a.rho * a.v1 * a.v2;
    f[5] = (a.E + a.p) * a.v1;
  } else {
    const double vr = a.vr_m;
    f[0] = a.rd * a.v2;
    f[1] = a.rm * a.v2;
    f[2] = a.rr * (a.v2 - vr);
    f[3] = a.rho * a.v1 * a.v2 - a.rr * vr * a.v1;
    f[4] = a.rho * a.v2 * a.v2 + a.p - a.rr * vr * a.v2;
    f[5] = (a.E + a.p) * a.v2 - (c_l * a.T + 0.5 * (a.v1 * a.v1 + a.v2 * a.v2)) * a.rr * vr;
  }
  f[6] = f[7] = f[8] = 0.0;
}

// flux_ec_rain (:872-960)
__device__ __forceinline__ void flux_ec_rain(const RNode& l, const RNode& r, double n1, double n2,
                                             double* f) {
  const double itl = 1.0 / l.T, itr = 1.0 / r.T;
  const double v1a = 0.5 * (l.v1 + r.v1), v2a = 0.5 * (l.v2 + r.v2);
  const double v1sq = 0.5 * (l.v1 * l.v1 + r.v1 * r.v1), v2sq = 0.5 * (l.v2 * l.v2 + r.v2 * r.v2);
  const double K = 0.5 * (v1sq + v2sq);
  const double vdn = n1 * v1a + n2 * v2a;
  const double vrn = 0.5 * n2 * (l.vr_sol + r.vr_sol);
  const double dry_a = 0.5 * (l.rd + r.rd);
  const double vap_a = 0.5 * (l.rv + r.rv);
  const double cld_a = 0.5 * (l.rc + r.rc);
  const double rain_a = 0.5 * (l.rr + r.rr);
  double dry_log = 0.0, vap_log = 0.0;
  if (l.rd != 0.0 && r.rd != 0.0) dry_log = ln_mean(l.rd, r.rd);
  if (l.rv != 0.0 && r.rv != 0.0) vap_log = ln_mean(l.rv, r.rv);
  const double it_avg = 0.5 * (itl + itr);
  const double it_log = inv_ln_mean(itl, itr);
  const double p_int = (1.0 / it_avg) * (R_d * dry_a + R_v * vap_a);
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
  f[6] = f[7] = f[8] = 0.0;
}

__device__ __forceinline__ ECNode ecnode(const double* u) {
  ECNode a;
  const double rho = u[0] + u[1] + u[2];
  const double inv = div_nr(1.0, rho);
  double T;
  nonlinear_sol(u, a.rv, a.rc, T);
  a.rd = u[0];
  a.rr = u[2];
  a.v1 = u[3] * inv;
  a.v2 = u[4] * inv;
  a.it = div_nr(1.0, T);
  a.vr = terminal_velocity(a.rv + a.rc, a.rr);
  a.lrd = log_(a.rd);
  a.lrv = log_(a.rv);
  a.lit = -log_(T);
  return a;
}

// flux_lax_friedrichs with max_abs_speed_naive (:607-624) on all nine entries
__device__ __forceinline__ void flux_llf(const RNode& l, const RNode& r, const double* ul,
                                         const double* ur, double n1, double n2, double* f) {
  double fl[NV], fr[NV];
  flux_n(l, n1, n2, fl);
  flux_n(r, n1, n2, fr);
  double vl = fabs(l.v1 * n1 + l.v2 * n2);
  vl += fabs(l.vr_m * n2);
  double vr = fabs(r.v1 * n1 + r.v2 * n2);
  vr += fabs(r.vr_m * n2);
  const double lam = fmax(vl, vr) + fmax(l.c, r.c) * sqrt_(n1 * n1 + n2 * n2);
#pragma unroll
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * lam * (ur[v] - ul[v]);
}

__device__ __forceinline__ void surface_flux(const RainyLaunch& L, const double* ul,
                                             const double* ur, double n1, double n2, double* f) {
  const RNode l = rnode(ul), r = rnode(ur);
  if (L.surface_flux == 1) flux_ec_rain(l, r, n1, n2, f);
  else flux_llf(l, r, ul, ur, n1, n2, f);
}

// source_terms_rainy (:547-577)
__device__ __forceinline__ void source_rainy(const double* u, double* s) {
  const double rho = u[0] + u[1] + u[2];
  double rv, rc, T;
  nonlinear_sol(u, rv, rc, T);
  const double rho_vs = svp(T) / (R_v * T);
  const double rr = u[2];
  // rr^(3/16), rr^(1/2), rr^(7/8) from the chain of square roots
  const double r2 = sqrt_(rr), r4 = sqrt_(r2), r8 = sqrt_(r4), r16 = sqrt_(r8);
  double S_ev = (3.86e-3 - 9.41e-5 * (T - ref_T)) * (1 + 9.1 * (r16 * r8));
  S_ev *= (rho_vs - rv) * r2;
  const double S = 0.001 * rc + 1.72 * rc * (r2 * r4 * r8) - S_ev;
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  s[1] = -S;
  s[2] = S;
  s[4] = -rho * grav;
  s[5] = -u[4] * grav;
}

// examples/euler/precipitation/tests/convergence_test.jl:45-139
__device__ __forceinline__ void source_convergence(double x, double y, double t, double* s) {
  const double c = 2, A = 0.1, Lw = 2, fr = 1 / Lw, w = 2 * kPi * fr;
  const double c_help = v_0 * 1.9386213994279082 * pow(kPi * N_0, -0.125);
  const double si = sin(w * (x + y - t)), co = cos(w * (x + y - t));
  const double rho = c + A * si;
  const double rho_x = w * A * co;
  const double rho_t = -rho_x;
  const double T = rho + 250.0;
  const double sv = svp(T), dsv = svp_deriv_from(sv, T);
  const double rv = sv / (R_v * T);
  const double rc = rho / c_l * 3000;
  const double rr = rho / c_l * 1000;
  const double rm = rv + rc;
  const double rd = rho - rm - rr;
  const double vr = terminal_velocity(rm, rr);
  const double svp_t = rho_t * dsv, svp_x = rho_x * dsv;
  const double rv_t = (svp_t * T - rho_t * sv) / (R_v * T * T);
  const double rv_x = (svp_x * T - rho_x * sv) / (R_v * T * T);
  const double rc_t = rho_t / c_l * 3000, rc_x = rho_x / c_l * 3000;
  const double rr_t = rho_t / c_l * 1000, rr_x = rho_x / c_l * 1000;
  const double rm_t = rv_t + rc_t, rm_x = rv_x + rc_x;
  const double rd_t = rho_t - rm_t - rr_t, rd_x = rho_x - rm_x - rr_x;
  double E_t = (c_vd * rd_t + c_vv * rv_t + c_l * (rc_t + rr_t)) * T;
  E_t += (c_vd * rd + c_vv * rv + c_l * (rc + rr)) * rho_t;
  E_t += rv_t * L_ref + rho_t;
  double E_x = (c_vd * rd_x + c_vv * rv_x + c_l * (rc_x + rr_x)) * T;
  E_x += (c_vd * rd + c_vv * rv + c_l * (rc + rr)) * rho_x;
  E_x += rv_x * L_ref + rho_x;
  double p_x = (rd_x * R_d + rv_x * R_v) * T;
  p_x += (rd * R_d + rv * R_v) * rho_x;
  double vr_x = c_help * 0.125 * ((rr_x * rm - rr * rm_x) / ((rm + rr) * (rm + rr)));
  vr_x *= pow(rr / (rm + rr), -0.875);
  const double rrvr_x = rr_x * vr + rr * vr_x;
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  s[0] = rd_t + 2 * rd_x;
  s[1] = rm_t + 2 * rm_x;
  s[2] = rr_t + 2 * rr_x - rrvr_x;
  s[3] = rho_x + p_x - rrvr_x;
  s[4] = rho_x + p_x - rrvr_x;
  double sE = E_t + 2 * (E_x + p_x) - (c_l * rho_x * rr * vr);
  sE -= (c_l * T + 1) * rrvr_x;
  s[5] = sE;
}

// saturation_residual and its Jacobian (:738-809); J[row + 3 col]; both take
// sv = saturation_vapour_pressure(x[2])
__device__ __forceinline__ void residual(const double* u, const double* x, double sv, double* r) {
  const double rho = u[0] + u[1] + u[2];
  const double inv = 1.0 / rho;
  const double v1 = u[3] * inv, v2 = u[4] * inv;
  double r1 = (c_vd * u[0] + c_vv * x[0] + c_l * (x[1] + u[2])) * x[2];
  r1 += x[0] * L_ref;
  r1 -= (u[5] - rho * 0.5 * (v1 * v1 + v2 * v2));
  double r2 = fmin(sv / (R_v * x[2]), u[1]);
  r2 -= x[0];
  r2 *= 1.0e7;
  double r3 = u[1];
  r3 -= x[0] + x[1];
  r3 *= 1.0e7;
  r[0] = r1;
  r[1] = r2;
  r[2] = r3;
}

__device__ __forceinline__ void jacobian(const double* u, const double* x, double sv, double* J) {
  const double svt = svp_deriv_from(sv, x[2]);
  J[0] = c_vv * x[2] + L_ref;
  J[3] = c_l * x[2];
  J[6] = c_vd * u[0] + c_vv * x[0] + c_l * (x[1] + u[2]);
  J[1] = -10000000.0;
  J[4] = 0.0;
  J[7] = (sv / (R_v * x[2]) < u[1]) ? (svt * x[2] - sv) / (R_v * x[2] * x[2]) * 1.0e7 : 0.0;
  J[2] = -10000000.0;
  J[5] = -10000000.0;
  J[8] = 0.0;
}

// 3x3 Gaussian elimination with partial pivoting (the first largest pivot of
// each column, one row swap); rows are swapped with selects so the system
// stays in registers
__device__ __forceinline__ void swap_rows(double (&a)[3][4], int c, int p) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (r <= c) continue;
    const bool sw = p == r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double tc = a[c][j], tr = a[r][j];
      a[c][j] = sw ? tr : tc;
      a[r][j] = sw ? tc : tr;
    }
  }
}

__device__ __forceinline__ void solve3(const double* J, const double* r, double* dx) {
  double a[3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = J[i + 3 * j];
    a[i][3] = r[i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int p = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i <= c) continue;
      const double v = fabs(a[i][c]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    swap_rows(a, c, p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i <= c) continue;
      const double f = a[i][c] / a[c][c];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j >= c) a[i][j] -= f * a[c][j];
    }
  }
#pragma unroll
  for (int i = 2; i >= 0; --i) {
    double sacc = a[i][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (j > i) sacc -= a[i][j] * dx[j];
    dx[i] = sacc / a[i][i];
  }
}

// one node of nonlinear_solve_dg_2d!: solve from guess x (updated in place),
// the residual reading `un` as it was before the rain clamp; returns the
// number of updates or -1 when max_iter updates did not reach the tolerance
__device__ __forceinline__ int newton_node(const double* un, double* x, double tol, int max_iter) {
  for (int it = 1; it <= max_iter; ++it) {
    double r[3];
    const double sv = svp(x[2]);
    residual(un, x, sv, r);
    if (fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2]))) < tol) return it - 1;
    double J[9], dx[3];
    jacobian(un, x, sv, J);
    solve3(J, r, dx);
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] += -dx[k];
  }
  return -1;
}

template <int N>
struct RShape {
  static constexpr int N2 = N * N;
  static constexpr int EPB = N2 >= 64 ? 1 : 64 / N2;
  static constexpr int BLOCK = ((EPB * N2 + 63) / 64) * 64;
  static constexpr int NF = 4 * N;
};

// the Newton statistics of one wave (updates, failures, largest update count,
// fallbacks), added to the wave's own slot stats[4 * slot ..] by its lane 0:
// no two waves of a launch share a slot and launches on the handle's stream
// follow each other, so no atomics are needed (one 64-bit atomic per wave on
// three shared counters serialised in the L2 and cost ~0.3 ms per launch)
__device__ __forceinline__ void accumulate_stats(unsigned long long* stats, size_t slot, int it,
                                                 int fallback) {
  unsigned long long upd = it > 0 ? (unsigned long long)it : 0ull;
  unsigned long long fail = it < 0 ? 1ull : 0ull;
  unsigned long long fb = fallback ? 1ull : 0ull;
  unsigned long long mx = upd;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    upd += __shfl_xor(upd, o, 64);
    fail += __shfl_xor(fail, o, 64);
    fb += __shfl_xor(fb, o, 64);
    const unsigned long long other = __shfl_xor(mx, o, 64);
    mx = other > mx ? other : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    unsigned long long* s = stats + 4 * slot;
    s[0] += upd;
    s[1] += fail;
    s[2] = s[2] > mx ? s[2] : mx;
    s[3] += fb;
  }
}

}  // namespace

// network inputs of one state: (rho, internal energy density, total water)
__device__ __forceinline__ void mlp_inputs(const double* u, float* x) {
  const double rho = u[0] + u[1] + u[2];
  const double inv = 1.0 / rho;
  const double v1 = u[3] * inv, v2 = u[4] * inv;
  x[0] = (float)rho;
  x[1] = (float)(u[5] - 0.5 * rho * (v1 * v1 + v2 * v2));
  x[2] = (float)(u[1] + u[2]);
}

// network inputs of the lanes of a wave: the node's state where `valid`,
// else the normalisation means (a harmless input for lanes without a node)
__device__ __forceinline__ void mlp_inputs_or_mean(const double* u, bool valid, const float* norm,
                                                   float* x) {
  if (valid) {
    mlp_inputs(u, x);
  } else {
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = norm[i];
  }
}

// seeded guess of one node from the network's prediction y = (T, q_v): q_v
// splits rho_moist into vapour and cloud (all vapour when the prediction is
// within 1 % of it: the unsaturated branch), and T then follows from the
// energy row of the saturation residual, which is linear in T, so the guess
// satisfies the energy and total-water rows exactly and Newton only has the
// saturation row left to correct; returns false when the guess is unusable
__device__ __forceinline__ bool mlp_seed(const double* u, const float* y, double* x) {
  const double rho = u[0] + u[1] + u[2];
  const double inv = 1.0 / rho;
  const double v1 = u[3] * inv, v2 = u[4] * inv;
  const double e_int = u[5] - rho * 0.5 * (v1 * v1 + v2 * v2);
  double rv = fmin(fmax((double)y[1] * rho, 0.0), u[1]);
  if (rv >= 0.99 * u[1]) rv = u[1];
  const double rc = u[1] - rv;
  x[0] = rv;
  x[1] = rc;
  x[2] = (e_int - rv * L_ref) / (c_vd * u[0] + c_vv * rv + c_l * (rc + u[2]));
  return x[2] > 0.0;
}

// closed-form guess without a network: all moisture as vapour with T from
// the energy row (exact below saturation: no update is then needed); where
// that state is supersaturated, one saturation adjustment at that
// temperature (vapour at its saturation density, the rest cloud) with T from
// the energy row again.  false when T is not positive.
__device__ __forceinline__ bool closed_form_seed(const double* u, double* x) {
  const double rho = u[0] + u[1] + u[2];
  const double inv = 1.0 / rho;
  const double v1 = u[3] * inv, v2 = u[4] * inv;
  const double e_int = u[5] - rho * 0.5 * (v1 * v1 + v2 * v2);
  double rv = u[1], rc = 0.0;
  double T = (e_int - rv * L_ref) / (c_vd * u[0] + c_vv * rv + c_l * u[2]);
  if (T > 0.0) {
    const double rvs = svp(T) / (R_v * T);
    if (rv > rvs) {
      rv = rvs;
      rc = u[1] - rv;
      T = (e_int - rv * L_ref) / (c_vd * u[0] + c_vv * rv + c_l * (rc + u[2]));
    }
  }
  x[0] = rv;
  x[1] = rc;
  x[2] = T;
  return T > 0.0;
}

// the saturation solve of one node from the closed-form guess, redone from
// the default guess u[7:9] (x on entry) where it fails; returns the update
// count (-1: failed) and sets *fallback
__device__ __forceinline__ int closed_form_solve(const double* u, double* x, double tol,
                                                 int max_iter, int* fallback) {
  const double x0[3] = {x[0], x[1], x[2]};
  const bool ok = closed_form_seed(u, x);
  int it = ok ? newton_node(u, x, tol, max_iter) : -1;
  if (it < 0 || !(x[2] > 0.0)) {
    const int used = ok ? max_iter : 0;
    x[0] = x0[0];
    x[1] = x0[1];
    x[2] = x0[2];
    const int it2 = newton_node(u, x, tol, max_iter);
    it = it2 < 0 ? -1 : used + it2;
    *fallback = 1;
  }
  return it;
}

// t0 + c dt rounded as two separate operations (the host's order, no fused
// multiply-add), so a stage time read from the device clock equals the host's
__device__ __forceinline__ double stage_time(double t0, double c, double dt) {
#pragma clang fp contract(off)
  return t0 + c * dt;
}

// StepsizeCallback's scaled wave speeds at one node of the uniform mesh:
// |Ja^i . lambda| / J with lambda_1 = |v_1| + c, lambda_2 = |v_2| + c + |v_r|
// (the rain's terminal velocity); the max-speed reduction and the last stage
// of a resident step both call this one function, so they agree bit for bit
__device__ __forceinline__ void rainy_node_speeds(const RainyLaunch& L, const double* u,
                                                  double l[2]) {
  const RNode a = rnode(u);
  const double lam1 = fabs(a.v1) + a.c;
  const double lam2 = fabs(a.v2) + a.c + fabs(a.vr_m);
  l[0] = fabs(L.Ja11 * lam1 + 0.0 * lam2) * L.inv_J;
  l[1] = fabs(0.0 * lam1 + L.Ja22 * lam2) * L.inv_J;
}

// SPEED: the RK stage also writes L.speed_partial (the last stage of a
// resident step); a variant of its own, so the other stages keep their
// register budget
// polydeg 3 without the network: a four-wave register budget (128 VGPRs,
// 64 B of scratch per lane) measured 67.0 against 68.2 us per stage at 136
// VGPRs and three waves; the network-seeded variant, whose inference holds
// more values, keeps the compiler's choice
template <int N, bool SEEDED, bool SPEED>
__global__ __launch_bounds__(RShape<N>::BLOCK)
__attribute__((amdgpu_waves_per_eu(N == 4 && !SEEDED ? 4 : 1)))
void rainy_stage(RainyLaunch L) {
  using S = RShape<N>;
  constexpr int N2 = S::N2, EPB = S::EPB, B = S::BLOCK;
  // one buffer per element, used in turn for the volume terms (flux
  // differencing: the per-node quantities of flux_ec_rain [11][N2]; weak form:
  // the contravariant fluxes of the 6 flux rows, [2][6][N2] where they fit,
  // else one direction [6][N2] at a time) and for the face fluxes [9][4 n].
  // Sized by flux differencing and the faces: at polydeg 3 the workgroup then
  // takes 10,240 B of LDS and sixteen fit on a CU (160 KiB) where 10,752 B
  // let fifteen, which put the 8,192 workgroups of a 256 x 128 mesh in three
  // rounds instead of two
  constexpr int VOLW = 11 * N2 > NV * S::NF ? 11 * N2 : NV * S::NF;
  constexpr bool WEAK_SPLIT = 12 * N2 > VOLW;
  __shared__ double s_u[EPB][NV][N2];
  __shared__ double s_buf[EPB][VOLW];
  auto s_f = [&](int es, int o, int v, int q) -> double& { return s_buf[es][(o * 6 + v) * N2 + q]; };
  auto s_ec = [&](int es, int k, int q) -> double& { return s_buf[es][k * N2 + q]; };
  auto s_sf = [&](int es, int v, int r) -> double& { return s_buf[es][v * S::NF + r]; };

  const int t = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  double rk_dt = L.rk_dt, t_src = L.t;
  if (L.clock) {
    rk_dt = L.clock[1];
    t_src = stage_time(L.clock[0], L.c_stage, rk_dt);
    if (L.mode == 1 && rk_dt == 0.0) {  // finished: the state is carried over as it is
      for (int idx = t; idx < EPB * N2 * NV; idx += B) {
        const size_t g = (size_t)e0 * N2 * NV + idx;
        if (e0 + idx / (N2 * NV) < L.n_elements) L.out[g] = L.u[g];
      }
      return;
    }
  }
  for (int idx = t; idx < EPB * N2 * NV; idx += B) {
    const int es = idx / (N2 * NV), r = idx - es * N2 * NV;
    if (e0 + es < L.n_elements) s_u[es][r % NV][r / NV] = L.u[(size_t)(e0 + es) * N2 * NV + r];
  }
  // the network's parameters are read from memory where used: the reads are
  // uniform over the wave (cached), cheaper than staging them per workgroup
  const bool seeded = SEEDED && L.mode == 1 && L.newton_max_iter > 0;
  __syncthreads();

  const int es = t / N2, q = t - es * N2;
  const int e = e0 + es;
  const bool active = es < EPB && e < L.n_elements;
  const int ti = q % N, tj = q / N;
  double du[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) du[v] = 0.0;

  if (L.volume_integral == 0) {
    // weak form: node fluxes into LDS, then du_i = sum_m dhat[i][m] F_m per line
    double g2[6];  // the second direction's contravariant flux (split form)
    if (active) {
      double uo[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) uo[v] = s_u[es][v][q];
      const RNode a = rnode(uo);
      double f1[NV], f2[NV];
      flux_o(a, 1, f1);
      flux_o(a, 2, f2);
#pragma unroll
      for (int v = 0; v < 6; ++v) {
        s_f(es, 0, v, q) = L.Ja11 * f1[v] + 0.0 * f2[v];
        g2[v] = 0.0 * f1[v] + L.Ja22 * f2[v];
        if constexpr (!WEAK_SPLIT) s_f(es, 1, v, q) = g2[v];
      }
    }
    __syncthreads();
    if constexpr (!WEAK_SPLIT) {
      if (active) {
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const double d1 = L.dhat[ti * N + m], d2 = L.dhat[tj * N + m];
#pragma unroll
          for (int v = 0; v < 6; ++v)
            du[v] += d1 * s_f(es, 0, v, m + N * tj) + d2 * s_f(es, 1, v, ti + N * m);
        }
      }
    } else {
      // one direction at a time through the same rows of the buffer
      if (active) {
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const double d1 = L.dhat[ti * N + m];
#pragma unroll
          for (int v = 0; v < 6; ++v) du[v] += d1 * s_f(es, 0, v, m + N * tj);
        }
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int v = 0; v < 6; ++v) s_f(es, 0, v, q) = g2[v];
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const double d2 = L.dhat[tj * N + m];
#pragma unroll
          for (int v = 0; v < 6; ++v) du[v] += d2 * s_f(es, 0, v, ti + N * m);
        }
      }
    }
  } else {
    // flux differencing with flux_ec_rain, gathered per node.  flux_ec_rain
    // gives the same bits with its two states exchanged (every operation is
    // commutative, negates both operands of a quotient, or is kept from
    // contraction where the order of a contracted sum would show), so each
    // lane takes its partners in rotated order own + 1, ..., own + n - 1
    // (mod n) with its own node on the left: every lane has a partner in each
    // of the n - 1 steps and the wave evaluates one flux per step (a loop over
    // the line's n nodes that skips the own one and orders the pair by index
    // runs both orders of the call under divergent lanes in every step)
    if (active) {
      double uo[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) uo[v] = s_u[es][v][q];
      const ECNode a = ecnode(uo);
      s_ec(es, 0, q) = a.rd; s_ec(es, 1, q) = a.rv; s_ec(es, 2, q) = a.rc; s_ec(es, 3, q) = a.rr;
      s_ec(es, 4, q) = a.v1; s_ec(es, 5, q) = a.v2; s_ec(es, 6, q) = a.it; s_ec(es, 7, q) = a.vr;
      s_ec(es, 8, q) = a.lrd; s_ec(es, 9, q) = a.lrv; s_ec(es, 10, q) = a.lit;
    }
    __syncthreads();
    if (active) {
      ECNode a;
      a.rd = s_ec(es, 0, q); a.rv = s_ec(es, 1, q); a.rc = s_ec(es, 2, q); a.rr = s_ec(es, 3, q);
      a.v1 = s_ec(es, 4, q); a.v2 = s_ec(es, 5, q); a.it = s_ec(es, 6, q); a.vr = s_ec(es, 7, q);
      a.lrd = s_ec(es, 8, q); a.lrv = s_ec(es, 9, q); a.lit = s_ec(es, 10, q);
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int own = o == 0 ? ti : tj;
        const double n1 = o == 0 ? L.Ja11 : 0.0, n2 = o == 0 ? 0.0 : L.Ja22;
#pragma unroll 1
        for (int j = 1; j < N; ++j) {
          const int m = own + j >= N ? own + j - N : own + j;
          const int pq = o == 0 ? m + N * tj : ti + N * m;
          ECNode b;
          b.rd = s_ec(es, 0, pq); b.rv = s_ec(es, 1, pq); b.rc = s_ec(es, 2, pq);
          b.rr = s_ec(es, 3, pq); b.v1 = s_ec(es, 4, pq); b.v2 = s_ec(es, 5, pq);
          b.it = s_ec(es, 6, pq); b.vr = s_ec(es, 7, pq);
          b.lrd = s_ec(es, 8, pq); b.lrv = s_ec(es, 9, pq); b.lit = s_ec(es, 10, pq);
          double f[6];
          flux_ec_rain(a, b, n1, n2, f);
          const double ds = L.dsplit[own * N + m];
#pragma unroll
          for (int v = 0; v < 6; ++v) du[v] += ds * f[v];
        }
      }
    }
  }

  // surface fluxes, one lane per face node of the workgroup's elements; the
  // buffer of the volume terms is free once every lane has gathered
  __syncthreads();
  for (int fid = t; fid < EPB * S::NF; fid += B) {
    const int fe = fid / S::NF, r = fid - fe * S::NF;
    const int f = r / N, l = r - f * N;
    const int ee = e0 + fe;
    if (ee >= L.n_elements) continue;
    const int ex = ee % L.nx, ey = ee / L.nx;
    double ul[NV], ur[NV], fl[NV];
    double n1 = 0.0, n2 = 0.0;
    if (f < 2) {
      n1 = L.Ja11;
      const int en = f == 0 ? ((ex + L.nx - 1) % L.nx) + L.nx * ey : ((ex + 1) % L.nx) + L.nx * ey;
      const int qo = f == 0 ? N * l : (N - 1) + N * l;          // own node
      const int qn = f == 0 ? (N - 1) + N * l : N * l;          // neighbour node
      const double* un = L.u + ((size_t)en * N2 + qn) * NV;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double own = s_u[fe][v][qo];
        ul[v] = f == 0 ? un[v] : own;
        ur[v] = f == 0 ? own : un[v];
      }
    } else {
      n2 = L.Ja22;
      const bool bottom = f == 2;
      const int qo = bottom ? l : l + N * (N - 1);
      const bool wall = !L.periodic_y && (bottom ? ey == 0 : ey == L.ny - 1);
      double nb[NV];
      if (wall) {
        // boundary_condition_simple_slip_wall (:963-983): mirrored state
#pragma unroll
        for (int v = 0; v < NV; ++v) nb[v] = s_u[fe][v][qo];
        nb[4] = -nb[4];
      } else {
        const int en = bottom ? ex + L.nx * ((ey + L.ny - 1) % L.ny) : ex + L.nx * ((ey + 1) % L.ny);
        const int qn = bottom ? l + N * (N - 1) : l;
        const double* un = L.u + ((size_t)en * N2 + qn) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) nb[v] = un[v];
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double own = s_u[fe][v][qo];
        ul[v] = bottom ? nb[v] : own;
        ur[v] = bottom ? own : nb[v];
      }
    }
    surface_flux(L, ul, ur, n1, n2, fl);
    const double sgn = (f % 2 == 0) ? -L.inv_w0 : L.inv_wn;
#pragma unroll
    for (int v = 0; v < NV; ++v) s_sf(fe, v, r) = sgn * fl[v];
  }
  __syncthreads();
  int it_stat = 0, fallback_stat = 0;
  double un[NV];  // the RK-updated state (stage mode)
#pragma unroll
  for (int v = 0; v < NV; ++v) un[v] = 0.0;
  if (active) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const bool on = f == 0 ? ti == 0 : f == 1 ? ti == N - 1 : f == 2 ? tj == 0 : tj == N - 1;
      if (!on) continue;
      const int l = f < 2 ? tj : ti;
#pragma unroll
      for (int v = 0; v < NV; ++v) du[v] += s_sf(es, v, f * N + l);
    }
    double uo[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      uo[v] = s_u[es][v][q];
      du[v] *= -L.inv_J;
    }
    if (L.source) {
      double sv[NV];
      if (L.source == 1) {
        source_rainy(uo, sv);
      } else if (L.source == 2) {
        const int ex = e % L.nx, ey = e / L.nx;
        const double x = L.x0 + L.dx * (ex + 0.5 * (L.xi[ti] + 1.0));
        const double y = L.y0 + L.dy * (ey + 0.5 * (L.xi[tj] + 1.0));
        source_convergence(x, y, t_src, sv);
      } else {  // source_terms_no_phase_change (:579-593)
#pragma unroll
        for (int v = 0; v < NV; ++v) sv[v] = 0.0;
        sv[4] = -grav * (uo[0] + uo[1] + uo[2]);
        sv[5] = -grav * uo[4];
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) du[v] += sv[v];
    }
    if (L.mode == 0) {
      const size_t g = ((size_t)e * N2 + q) * NV;
#pragma unroll
      for (int v = 0; v < NV; ++v) L.out[g + v] = du[v];
    } else {
      const size_t g = ((size_t)e * N2 + q) * NV;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double kn = L.rk_a * L.k[g + v] + rk_dt * du[v];
        L.k[g + v] = kn;
        un[v] = uo[v] + L.rk_b * kn;
      }
    }
  }
  // the network's prediction for the updated states: every lane of the wave
  float ypred[kMlpOut] = {0.0f, 0.0f};
  if constexpr (SEEDED) {
    if (seeded) {
      float xi[kMlpIn];
      mlp_inputs_or_mean(un, active, L.mlp.norm, xi);
      mlp_eval_wave(L.mlp.params, L.mlp.norm, xi, ypred);
    }
  }
  if (active && L.mode != 0) {
    const size_t g = ((size_t)e * N2 + q) * NV;
    // the stage limiter: NonlinearSolveDG on the updated state
    {
      if (L.newton_max_iter > 0) {
        double x[3] = {un[6], un[7], un[8]};
        if constexpr (SEEDED) {
          // network seed; a failed seeded solve is redone from u[7:9]
          const bool ok = mlp_seed(un, ypred, x);
          it_stat = ok ? newton_node(un, x, L.newton_tol, L.newton_max_iter) : -1;
          if (it_stat < 0 || !(x[2] > 0.0)) {
            const int used = ok ? L.newton_max_iter : 0;
            x[0] = un[6];
            x[1] = un[7];
            x[2] = un[8];
            const int it2 = newton_node(un, x, L.newton_tol, L.newton_max_iter);
            it_stat = it2 < 0 ? -1 : used + it2;
            fallback_stat = 1;
          }
        } else {
          if (L.closed_form_seed)
            it_stat = closed_form_solve(un, x, L.newton_tol, L.newton_max_iter, &fallback_stat);
          else
            it_stat = newton_node(un, x, L.newton_tol, L.newton_max_iter);
        }
        if (un[2] < 0.0) un[2] = 0.0;
        un[6] = x[0];
        un[7] = x[1];
        un[8] = x[2];
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) L.out[g + v] = un[v];
      if constexpr (SPEED) {
        // the node's speeds go to its first two state fields in LDS, which
        // only this lane reads after the face phase and which it has read
        double l[2];
        rainy_node_speeds(L, un, l);
        s_u[es][0][q] = l[0];
        s_u[es][1][q] = l[1];
      }
    }
  }
  if constexpr (SPEED) {
    // per element: the maximum over its nodes of each speed (exact in any
    // order, NaN-propagating), then the sum over the directions
    __syncthreads();
    if (active && q == 0) {
      double m1 = s_u[es][0][0], m2 = s_u[es][1][0];
      for (int r = 1; r < N2; ++r) {
        m1 = speed_max(m1, s_u[es][0][r]);
        m2 = speed_max(m2, s_u[es][1][r]);
      }
      L.speed_partial[e] = m1 + m2;
    }
  }
  // every lane of the (single-wave) workgroup takes part in the reduction
  if (L.stats && L.mode == 1 && L.newton_max_iter > 0)
    accumulate_stats(L.stats, blockIdx.x, it_stat, fallback_stat);
}

// per-element maximum of the scaled wave speeds (stepsize_dg2d.jl:5-43 pattern)
template <int N>
__global__ __launch_bounds__(64) void rainy_max_speed(RainyLaunch L, double* partial) {
  constexpr int N2 = N * N;
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double s1[64], s2[64];
  double l1 = 0.0, l2 = 0.0;
  for (int q = t; q < N2; q += 64) {
    double l[2];
    rainy_node_speeds(L, L.u + ((size_t)e * N2 + q) * NV, l);
    l1 = speed_max(l1, l[0]);
    l2 = speed_max(l2, l[1]);
  }
  s1[t] = l1;
  s2[t] = l2;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (t < w) {
      s1[t] = speed_max(s1[t], s1[t + w]);
      s2[t] = speed_max(s2[t], s2[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) partial[e] = s1[0] + s2[0];
}

__global__ __launch_bounds__(256) void rainy_max_reduce(const double* partial, int n, double* out) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  double acc = 5e-324;
  for (int i = t; i < n; i += 256) acc = speed_max(acc, partial[i]);
  s[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) s[t] = speed_max(s[t], s[t + w]);
    __syncthreads();
  }
  if (t == 0) out[0] = s[0];
}

template <int N>
__global__ __launch_bounds__(64) void rainy_integrals(RainyLaunch L, const double* w,
                                                      double* partial) {
  constexpr int N2 = N * N;
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double s[NV][64];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  const double J = 1.0 / L.inv_J;
  for (int q = t; q < N2; q += 64) {
    const double wJ = w[q % N] * w[q / N] * J;
    const double* u = L.u + ((size_t)e * N2 + q) * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += wJ * u[v];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v][t] = acc[v];
  __syncthreads();
  for (int k = 32; k > 0; k >>= 1) {
    if (t < k)
      for (int v = 0; v < NV; ++v) s[v][t] += s[v][t + k];
    __syncthreads();
  }
  if (t < NV) partial[(size_t)e * NV + t] = s[t][0];
}

__global__ __launch_bounds__(256) void rainy_sum_columns(const double* partial, int n_items,
                                                         double* out) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  for (int c = 0; c < NV; ++c) {
    double acc = 0.0;
    for (int i = t; i < n_items; i += 256) acc += partial[(size_t)i * NV + c];
    s[t] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (t < k) s[t] += s[t + k];
      __syncthreads();
    }
    if (t == 0) out[c] = s[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void rainy_newton_kernel(RainyLaunch L, double* u,
                                                           MlpView mlp, int use_mlp,
                                                           int* iterations,
                                                           unsigned long long* stats) {
  const size_t nodes = (size_t)L.n_elements * L.n * L.n;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int it = 0, fallback = 0;
  const bool valid = q < nodes;
  double uc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) uc[v] = valid ? u[q * NV + v] : 0.0;
  // the network's prediction: every lane of the wave takes part
  float ypred[kMlpOut] = {0.0f, 0.0f};
  if (use_mlp) {
    float xi[kMlpIn];
    mlp_inputs_or_mean(uc, valid, mlp.norm, xi);
    mlp_eval_wave(mlp.params, mlp.norm, xi, ypred);
  }
  if (valid) {
    double* un = u + q * NV;
    double x[3] = {uc[6], uc[7], uc[8]};
    if (use_mlp) {
      const bool ok = mlp_seed(uc, ypred, x);
      it = ok ? newton_node(uc, x, L.newton_tol, L.newton_max_iter) : -1;
      if (it < 0 || !(x[2] > 0.0)) {
        // fall back to the default guess, as the unseeded solve
        fallback = 1;
        const int used = ok ? (it < 0 ? L.newton_max_iter : it) : 0;
        x[0] = uc[6];
        x[1] = uc[7];
        x[2] = uc[8];
        const int it2 = newton_node(uc, x, L.newton_tol, L.newton_max_iter);
        it = it2 < 0 ? -1 : used + it2;
      }
    } else if (L.closed_form_seed) {
      it = closed_form_solve(uc, x, L.newton_tol, L.newton_max_iter, &fallback);
    } else {
      it = newton_node(uc, x, L.newton_tol, L.newton_max_iter);
    }
    if (uc[2] < 0.0) un[2] = 0.0;
    un[6] = x[0];
    un[7] = x[1];
    un[8] = x[2];
    if (iterations) iterations[q] = it;
  }
  // every lane of the wave takes part (a lane past the last node adds nothing)
  if (stats) accumulate_stats(stats, q / 64, it, fallback);
}

// features for training: x = (rho, internal energy, total water), y = (T, q_v)
__global__ void rainy_features_kernel(const double* u, size_t nodes, float* x, float* y) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nodes) return;
  const double* un = u + q * NV;
  float xi[kMlpIn];
  mlp_inputs(un, xi);
  x[3 * q] = xi[0];
  x[3 * q + 1] = xi[1];
  x[3 * q + 2] = xi[2];
  const double rho = un[0] + un[1] + un[2];
  y[2 * q] = (float)un[8];
  y[2 * q + 1] = (float)(un[6] / rho);
}

#define ATMO_RAINY_SWITCH(MACRO) \
  switch (L.n) {                 \
    MACRO(2)                     \
    MACRO(3)                     \
    MACRO(4)                     \
    MACRO(5)                     \
    MACRO(6)                     \
    MACRO(7)                     \
    MACRO(8)                     \
    default: return false;       \
  }

bool rainy_launch_stage(const RainyLaunch& L, hipStream_t s) {
#define ATMO_CASE(NN)                                                                         \
  case NN: {                                                                                  \
    const int grid = (L.n_elements + RShape<NN>::EPB - 1) / RShape<NN>::EPB;                  \
    const bool seeded = L.mlp.params && L.mode == 1 && L.newton_max_iter > 0;                 \
    const bool speed = L.speed_partial && L.mode == 1;                                          \
    if (seeded && speed)                                                                        \
      hipLaunchKernelGGL((rainy_stage<NN, true, true>), dim3(grid), dim3(RShape<NN>::BLOCK), 0, s, L); \
    else if (seeded)                                                                            \
      hipLaunchKernelGGL((rainy_stage<NN, true, false>), dim3(grid), dim3(RShape<NN>::BLOCK), 0, s, L); \
    else if (speed)                                                                             \
      hipLaunchKernelGGL((rainy_stage<NN, false, true>), dim3(grid), dim3(RShape<NN>::BLOCK), 0, s, L); \
    else                                                                                        \
      hipLaunchKernelGGL((rainy_stage<NN, false, false>), dim3(grid), dim3(RShape<NN>::BLOCK), 0, s, L); \
    return true;                                                                              \
  }
  ATMO_RAINY_SWITCH(ATMO_CASE)
#undef ATMO_CASE
}

bool rainy_launch_max_speed(const RainyLaunch& L, double* partial, double* out, hipStream_t s) {
#define ATMO_CASE(NN)                                                                         \
  case NN:                                                                                    \
    hipLaunchKernelGGL(rainy_max_speed<NN>, dim3(L.n_elements), dim3(64), 0, s, L, partial); \
    break;
  ATMO_RAINY_SWITCH(ATMO_CASE)
#undef ATMO_CASE
  hipLaunchKernelGGL(rainy_max_reduce, dim3(1), dim3(256), 0, s, partial, L.n_elements, out);
  return true;
}

bool rainy_launch_integrals(const RainyLaunch& L, const double* w, double* partial, double* out,
                            hipStream_t s) {
#define ATMO_CASE(NN)                                                                            \
  case NN:                                                                                       \
    hipLaunchKernelGGL(rainy_integrals<NN>, dim3(L.n_elements), dim3(64), 0, s, L, w, partial); \
    break;
  ATMO_RAINY_SWITCH(ATMO_CASE)
#undef ATMO_CASE
  hipLaunchKernelGGL(rainy_sum_columns, dim3(1), dim3(256), 0, s, partial, L.n_elements, out);
  return true;
}

// sums (max for the largest update count) of the per-wave statistics slots
// into out[4], clearing the slots for the next accumulation
__global__ __launch_bounds__(256) void rainy_stats_reduce(unsigned long long* __restrict__ slots,
                                                          int n_slots,
                                                          unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s[4][256];
  const int t = threadIdx.x;
  unsigned long long a[4] = {0ull, 0ull, 0ull, 0ull};
  for (int i = t; i < n_slots; i += 256) {
    unsigned long long* p = slots + 4 * (size_t)i;
    a[0] += p[0];
    a[1] += p[1];
    a[2] = a[2] > p[2] ? a[2] : p[2];
    a[3] += p[3];
    p[0] = p[1] = p[2] = p[3] = 0ull;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) s[k][t] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s[0][t] += s[0][t + w];
      s[1][t] += s[1][t + w];
      s[2][t] = s[2][t] > s[2][t + w] ? s[2][t] : s[2][t + w];
      s[3][t] += s[3][t + w];
    }
    __syncthreads();
  }
  if (t < 4) out[t] = s[t][0];
}

void rainy_launch_stats_reduce(unsigned long long* slots, int n_slots, unsigned long long* out,
                               hipStream_t s) {
  hipLaunchKernelGGL(rainy_stats_reduce, dim3(1), dim3(256), 0, s, slots, n_slots, out);
}

void rainy_launch_newton(const RainyLaunch& L, double* u, const MlpView* mlp, int* iterations,
                         hipStream_t s) {
  const size_t nodes = (size_t)L.n_elements * L.n * L.n;
  const int grid = (int)((nodes + 255) / 256);
  MlpView m = mlp ? *mlp : MlpView{nullptr, nullptr};
  hipLaunchKernelGGL(rainy_newton_kernel, dim3(grid), dim3(256), 0, s, L, u, m, mlp ? 1 : 0,
                     iterations, L.stats);
}

void rainy_launch_features(const RainyLaunch& L, const double* u, float* x, float* y,
                           hipStream_t s) {
  const size_t nodes = (size_t)L.n_elements * L.n * L.n;
  hipLaunchKernelGGL(rainy_features_kernel, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0,
                     s, u, nodes, x, y);
}

}  // namespace atmo
