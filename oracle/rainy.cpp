// CPU ORACLE (test infrastructure only): compressible rainy Euler equations in
// 2D (CompressibleRainyEulerEquations2D) on a uniform Cartesian StructuredMesh,
// the per-node saturation Newton solve of NonlinearSolveDG, the DGSEM rhs!
// (weak form or flux differencing with flux_ec_rain, flux_lax_friedrichs on
// the surface), and CarpenterKennedy2N54 with the Newton solve as stage
// limiter.  FP64, restated from
//   src/equations/compressible_rainy_euler_2d.jl
//   src/callbacks_stage/nonlinear_solve_dg2d.jl, nonlinear_solve_dg.jl
//   src/callbacks_step/stepsize_dg2d.jl:5-43 (max_dt pattern)
//   examples/euler/precipitation/tests/convergence_test.jl (manufactured source)
// and Trixi's StructuredMesh{2} DGSEM kernels [ext] (weak-form and
// flux-differencing volume integrals, calc_interface_flux!, surface integral,
// apply_jacobian!, calc_sources!).
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int NV = 9;

// CompressibleRainyEulerEquations2D() constants (compressible_rainy_euler_2d.jl:77-113)
struct C {
  static constexpr double c_l = 4186, c_pd = 1004, c_vd = 717, c_pv = 1885, c_vv = 1424;
  static constexpr double R_d = c_pd - c_vd, R_v = c_pv - c_vv;
  static constexpr double ref_sp = 610.7, ref_T = 273.15, L_ref = 2500000, ref_p = 100000;
  static constexpr double g = 9.80616, N_0 = 8000000, v_0 = 130;
};

// :265-274, :276-279
struct Dens {
  double dry, moist, rain, rho, inv;
};
inline Dens densities(const double* u) {
  Dens d{u[0], u[1], u[2], u[0] + u[1] + u[2], 0.0};
  d.inv = 1.0 / d.rho;
  return d;
}

// :219-260 (the dry branch recovers T from the energy)
void nonlinear_sol(const double* u, double* vct) {
  const Dens d = densities(u);
  if (d.moist == 0.0 && d.rain == 0.0) {
    const double v1 = u[3] * d.inv, v2 = u[4] * d.inv;
    const double ek = 0.5 * (v1 * v1 + v2 * v2) * d.rho;
    vct[0] = 0.0;
    vct[1] = 0.0;
    vct[2] = (u[5] - ek) / (C::c_vd * d.rho);
    return;
  }
  double rv = u[6], rc = u[7];
  if (rv < 0.0 && rv >= -1e-15) rv = 0.0;
  if (rc < 0.0 && rc >= -1e-15) rc = 0.0;
  vct[0] = rv;
  vct[1] = rc;
  vct[2] = u[8];
}

// :285-297
double pressure(const double* u) {
  double s[3];
  nonlinear_sol(u, s);
  return (C::R_d * u[0] + C::R_v * s[0]) * s[2];
}

// :299-334 (returns sound speed; gamma_m through *gamma)
double speed_of_sound(const double* u, double* gamma) {
  const Dens d = densities(u);
  double s[3];
  nonlinear_sol(u, s);
  const double p = (C::R_d * u[0] + C::R_v * s[0]) * s[2];
  const double q_v = s[0] / d.dry;
  const double q_l = (s[1] + d.rain) / d.dry;
  const double gm = 1 + (C::R_d + C::R_v * q_v) / (C::c_vd + C::c_vv * q_v + C::c_l * q_l);
  if (gamma) *gamma = gm;
  return std::sqrt(gm * p * d.inv);
}

// :336-351
double terminal_velocity_rain(double rho_moist, double rho_rain) {
  if (rho_rain > 0.0)
    return C::v_0 * 1.9386213994279082 *
           std::pow(rho_rain / (M_PI * (rho_moist + rho_rain) * C::N_0), 0.125);
  return 0.0;
}

}  // namespace

// :353-373 (Clausius-Clapeyron) and :375-398 (its derivative)
extern "C" void ora_rainy_saturation_vapour_pressure(double T, double* svp, double* dsvp) {
  const double k1 = (C::c_pv - C::c_l) / C::R_v;
  const double k2 = (C::L_ref - (C::c_pv - C::c_l) * C::ref_T) / C::R_v;
  if (svp) *svp = C::ref_sp * std::pow(T / C::ref_T, k1) * std::exp(k2 * (1 / C::ref_T - 1 / T));
  if (dsvp) {
    double d = C::ref_sp / std::pow(C::ref_T, k1);
    d *= k1 * std::pow(T, k1 - 1) + k2 * std::pow(T, k1 - 2);
    d *= std::exp(k2 * (1 / C::ref_T - 1 / T));
    *dsvp = d;
  }
}

// :738-772
extern "C" void ora_rainy_residual(const double* u, const double* x, double* r) {
  const Dens d = densities(u);
  const double v1 = u[3] * d.inv, v2 = u[4] * d.inv;
  double r1 = (C::c_vd * d.dry + C::c_vv * x[0] + C::c_l * (x[1] + d.rain)) * x[2];
  r1 += x[0] * C::L_ref;
  r1 -= (u[5] - d.rho * 0.5 * (v1 * v1 + v2 * v2));
  double svp;
  ora_rainy_saturation_vapour_pressure(x[2], &svp, nullptr);
  double r2 = std::fmin(svp / (C::R_v * x[2]), d.moist);
  r2 -= x[0];
  r2 *= 1.0e7;
  double r3 = d.moist;
  r3 -= x[0] + x[1];
  r3 *= 1.0e7;
  r[0] = r1;
  r[1] = r2;
  r[2] = r3;
}

// :774-809, J[row + 3 * col]
extern "C" void ora_rainy_residual_jacobian(const double* u, const double* x, double* J) {
  const Dens d = densities(u);
  double svp, svp_t;
  ora_rainy_saturation_vapour_pressure(x[2], &svp, &svp_t);
  J[0] = C::c_vv * x[2] + C::L_ref;
  J[3] = C::c_l * x[2];
  J[6] = C::c_vd * d.dry + C::c_vv * x[0] + C::c_l * (x[1] + d.rain);
  J[1] = -10000000.0;
  J[4] = 0.0;
  J[7] = (svp / (C::R_v * x[2]) < d.moist)
             ? (svp_t * x[2] - svp) / (C::R_v * x[2] * x[2]) * 1.0e7
             : 0.0;
  J[2] = -10000000.0;
  J[5] = -10000000.0;
  J[8] = 0.0;
}

namespace {

// 3x3 solve by Gaussian elimination with partial pivoting (the `\` of a
// StaticArrays SMatrix{3,3} solves the same system)
void solve3(const double* J, const double* r, double* dx) {
  double a[3][4];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) a[i][j] = J[i + 3 * j];
    a[i][3] = r[i];
  }
  for (int c = 0; c < 3; ++c) {
    int p = c;
    for (int i = c + 1; i < 3; ++i)
      if (std::fabs(a[i][c]) > std::fabs(a[p][c])) p = i;
    if (p != c)
      for (int j = 0; j < 4; ++j) std::swap(a[c][j], a[p][j]);
    for (int i = c + 1; i < 3; ++i) {
      const double f = a[i][c] / a[c][c];
      for (int j = c; j < 4; ++j) a[i][j] -= f * a[c][j];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double s = a[i][3];
    for (int j = i + 1; j < 3; ++j) s -= a[i][j] * dx[j];
    dx[i] = s / a[i][i];
  }
}

}  // namespace

// one node of nonlinear_solve_dg_2d! (nonlinear_solve_dg2d.jl:10-43): the guess
// is u[6..8]; a negative rain density is clamped to zero in u (the residual
// still sees the value read before the clamp); returns the number of Newton
// updates, or -1 if the tolerance was not met within max_iter updates (the
// last iterate is kept, as the reference does after its warning).
extern "C" int ora_rainy_newton_node(double* u, const double* guess_in, double tol, int max_iter) {
  double un[NV];
  std::memcpy(un, u, sizeof(un));
  double x[3] = {u[6], u[7], u[8]};
  if (guess_in) {
    x[0] = guess_in[0];
    x[1] = guess_in[1];
    x[2] = guess_in[2];
  }
  if (u[2] < 0.0) u[2] = 0.0;
  int updates = -1;
  for (int it = 1; it <= max_iter; ++it) {
    double r[3];
    ora_rainy_residual(un, x, r);
    if (std::fmax(std::fabs(r[0]), std::fmax(std::fabs(r[1]), std::fabs(r[2]))) < tol) {
      updates = it - 1;
      break;
    }
    double J[9], dx[3];
    ora_rainy_residual_jacobian(un, x, J);
    solve3(J, r, dx);
    for (int k = 0; k < 3; ++k) x[k] += -dx[k];
  }
  u[6] = x[0];
  u[7] = x[1];
  u[8] = x[2];
  return updates;
}

namespace {

// pointwise flux (:448-500 orientation form, :502-545 normal form)
void flux(const double* u, double n1, double n2, double* f) {
  const Dens d = densities(u);
  double s[3];
  nonlinear_sol(u, s);
  const double T = s[2];
  const double v1 = u[3] * d.inv, v2 = u[4] * d.inv;
  const double vr = terminal_velocity_rain(d.moist, d.rain);
  const double vn = v1 * n1 + v2 * n2;
  const double vrn = vr * n2;
  const double p = pressure(u);
  const double E = u[5];
  f[0] = d.dry * vn;
  f[1] = d.moist * vn;
  f[2] = d.rain * (vn - vrn);
  f[3] = d.rho * vn * v1 + p * n1 - d.rain * vrn * v1;
  f[4] = d.rho * vn * v2 + p * n2 - d.rain * vrn * v2;
  f[5] = (E + p) * vn - (C::c_l * T + 0.5 * (v1 * v1 + v2 * v2)) * d.rain * vrn;
  f[6] = f[7] = f[8] = 0.0;
}

// the orientation form of flux (:448-500): the same expressions with a unit
// direction except that the orientation-1 flux has no rain terms at all
void flux_orientation(const double* u, int orientation, double* f) {
  const Dens d = densities(u);
  double s[3];
  nonlinear_sol(u, s);
  const double T = s[2];
  const double v1 = u[3] * d.inv, v2 = u[4] * d.inv;
  const double vr = terminal_velocity_rain(d.moist, d.rain);
  const double p = pressure(u);
  const double E = u[5];
  if (orientation == 1) {
    f[0] = d.dry * v1;
    f[1] = d.moist * v1;
    f[2] = d.rain * v1;
    f[3] = d.rho * v1 * v1 + p;
    f[4] = d.rho * v1 * v2;
    f[5] = (E + p) * v1;
  } else {
    f[0] = d.dry * v2;
    f[1] = d.moist * v2;
    f[2] = d.rain * (v2 - vr);
    f[3] = d.rho * v1 * v2 - d.rain * vr * v1;
    f[4] = d.rho * v2 * v2 + p - d.rain * vr * v2;
    f[5] = (E + p) * v2 - (C::c_l * T + 0.5 * (v1 * v1 + v2 * v2)) * d.rain * vr;
  }
  f[6] = f[7] = f[8] = 0.0;
}

// :607-624
double max_abs_speed_naive(const double* ul, const double* ur, double n1, double n2) {
  const Dens dl = densities(ul), dr = densities(ur);
  const double cl = speed_of_sound(ul, nullptr), cr = speed_of_sound(ur, nullptr);
  const double vrl = terminal_velocity_rain(dl.moist, dl.rain);
  const double vrr = terminal_velocity_rain(dr.moist, dr.rain);
  double vl = std::fabs(ul[3] * dl.inv * n1 + ul[4] * dl.inv * n2);
  vl += std::fabs(vrl * n2);
  double vr = std::fabs(ur[3] * dr.inv * n1 + ur[4] * dr.inv * n2);
  vr += std::fabs(vrr * n2);
  return std::fmax(vl, vr) + std::fmax(cl, cr) * std::sqrt(n1 * n1 + n2 * n2);
}

// flux_lax_friedrichs: central flux plus -lambda/2 [u] on all nine variables
void flux_llf(const double* ul, const double* ur, double n1, double n2, double* f) {
  double fl[NV], fr[NV];
  flux(ul, n1, n2, fl);
  flux(ur, n1, n2, fr);
  const double lam = max_abs_speed_naive(ul, ur, n1, n2);
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * lam * (ur[v] - ul[v]);
}

}  // namespace

extern "C" void ora_rainy_flux_ec(const double* ul, const double* ur, const double* n, double* f);

namespace {

// the surface flux of the configuration: 0 flux_lax_friedrichs, 1 flux_ec_rain
int g_surface = 0;
void surface(const double* ul, const double* ur, double n1, double n2, double* f) {
  if (g_surface == 1) {
    const double nn[2] = {n1, n2};
    ora_rainy_flux_ec(ul, ur, nn, f);
  } else {
    flux_llf(ul, ur, n1, n2, f);
  }
}

}  // namespace

// :872-960 flux_ec_rain
extern "C" void ora_rainy_flux_ec(const double* ul, const double* ur, const double* n, double* f) {
  const Dens dl = densities(ul), dr = densities(ur);
  double sl[3], sr[3];
  nonlinear_sol(ul, sl);
  nonlinear_sol(ur, sr);
  const double itl = 1.0 / sl[2], itr = 1.0 / sr[2];
  const double v1l = ul[3] * dl.inv, v2l = ul[4] * dl.inv;
  const double v1r = ur[3] * dr.inv, v2r = ur[4] * dr.inv;
  const double vrl = terminal_velocity_rain(sl[0] + sl[1], dl.rain);
  const double vrr = terminal_velocity_rain(sr[0] + sr[1], dr.rain);
  const double v1a = 0.5 * (v1l + v1r), v2a = 0.5 * (v2l + v2r);
  const double v1sq = 0.5 * (v1l * v1l + v1r * v1r), v2sq = 0.5 * (v2l * v2l + v2r * v2r);
  const double K = 0.5 * (v1sq + v2sq);
  const double vdn = n[0] * v1a + n[1] * v2a;
  const double vrn = 0.5 * n[1] * (vrl + vrr);
  const double dry_a = 0.5 * (dl.dry + dr.dry);
  const double vap_a = 0.5 * (sl[0] + sr[0]);
  const double cld_a = 0.5 * (sl[1] + sr[1]);
  const double rain_a = 0.5 * (dl.rain + dr.rain);
  double dry_log = 0, vap_log = 0;
  if (dl.dry != 0.0 && dr.dry != 0.0) dry_log = ora_ln_mean(dl.dry, dr.dry);
  if (sl[0] != 0.0 && sr[0] != 0.0) vap_log = ora_ln_mean(sl[0], sr[0]);
  const double it_avg = 0.5 * (itl + itr);
  const double it_log = ora_inv_ln_mean(itl, itr);
  const double p_int = (1.0 / it_avg) * (C::R_d * dry_a + C::R_v * vap_a);
  const double f_vap = vap_log * vdn;
  const double f_cld = cld_a * vdn;
  const double f_dry = dry_log * vdn;
  const double f_rain = rain_a * (vdn - vrn);
  const double f_moist = f_vap + f_cld;
  const double f_rho = f_dry + f_moist + f_rain;
  const double f_m1 = f_rho * v1a + p_int * n[0];
  const double f_m2 = f_rho * v2a + p_int * n[1];
  const double f_E = (C::c_vd * it_log - K) * f_dry + (C::c_vv * it_log - K + C::L_ref) * f_vap +
                     (C::c_l * it_log - K) * (f_cld + f_rain) + (v1a * f_m1 + v2a * f_m2);
  f[0] = f_dry;
  f[1] = f_moist;
  f[2] = f_rain;
  f[3] = f_m1;
  f[4] = f_m2;
  f[5] = f_E;
  f[6] = f[7] = f[8] = 0.0;
}

extern "C" void ora_rainy_flux(const double* u, const double* n, double* f) {
  flux(u, n[0], n[1], f);
}

extern "C" void ora_rainy_flux_llf(const double* ul, const double* ur, const double* n,
                                   double* f) {
  flux_llf(ul, ur, n[0], n[1], f);
}

extern "C" double ora_rainy_max_abs_speed_naive(const double* ul, const double* ur,
                                                const double* n) {
  return max_abs_speed_naive(ul, ur, n[0], n[1]);
}

// :547-577 source_terms_rainy (s[9])
extern "C" void ora_rainy_source(const double* u, double* s) {
  const Dens d = densities(u);
  double sol[3];
  nonlinear_sol(u, sol);
  const double T = sol[2];
  double svp;
  ora_rainy_saturation_vapour_pressure(T, &svp, nullptr);
  const double rho_vs = svp / (C::R_v * T);
  double S_ev = (3.86e-3 - 9.41e-5 * (T - C::ref_T)) * (1 + 9.1 * std::pow(d.rain, 0.1875));
  S_ev *= (rho_vs - sol[0]) * std::pow(d.rain, 0.5);
  const double S_auto = 0.001 * sol[1];
  const double S_acc = 1.72 * sol[1] * std::pow(d.rain, 0.875);
  const double S = S_auto + S_acc - S_ev;
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  s[1] = -S;
  s[2] = S;
  s[4] = -d.rho * C::g;
  s[5] = -u[4] * C::g;
}

namespace {

// examples/euler/precipitation/tests/convergence_test.jl:45-139 (manufactured source)
void source_convergence(double x, double y, double t, double* s) {
  const double c = 2, A = 0.1, L = 2, f = 1 / L, w = 2 * M_PI * f;
  const double c_help = C::v_0 * 1.9386213994279082 * std::pow(M_PI * C::N_0, -0.125);
  const double si = std::sin(w * (x + y - t)), co = std::cos(w * (x + y - t));
  const double rho = c + A * si;
  const double rho_x = w * A * co;
  const double rho_t = -rho_x;
  const double T = rho + 250.0;
  double svp, dsvp;
  ora_rainy_saturation_vapour_pressure(T, &svp, &dsvp);
  const double rv = svp / (C::R_v * T);
  const double rc = rho / C::c_l * 3000;
  const double rr = rho / C::c_l * 1000;
  const double rm = rv + rc;
  const double rd = rho - rm - rr;
  const double vr = terminal_velocity_rain(rm, rr);
  const double svp_t = rho_t * dsvp, svp_x = rho_x * dsvp;
  const double rv_t = (svp_t * T - rho_t * svp) / (C::R_v * T * T);
  const double rv_x = (svp_x * T - rho_x * svp) / (C::R_v * T * T);
  const double rc_t = rho_t / C::c_l * 3000, rc_x = rho_x / C::c_l * 3000;
  const double rr_t = rho_t / C::c_l * 1000, rr_x = rho_x / C::c_l * 1000;
  const double rm_t = rv_t + rc_t, rm_x = rv_x + rc_x;
  const double rd_t = rho_t - rm_t - rr_t, rd_x = rho_x - rm_x - rr_x;
  double E_t = (C::c_vd * rd_t + C::c_vv * rv_t + C::c_l * (rc_t + rr_t)) * T;
  E_t += (C::c_vd * rd + C::c_vv * rv + C::c_l * (rc + rr)) * rho_t;
  E_t += rv_t * C::L_ref + rho_t;
  double E_x = (C::c_vd * rd_x + C::c_vv * rv_x + C::c_l * (rc_x + rr_x)) * T;
  E_x += (C::c_vd * rd + C::c_vv * rv + C::c_l * (rc + rr)) * rho_x;
  E_x += rv_x * C::L_ref + rho_x;
  double p_x = (rd_x * C::R_d + rv_x * C::R_v) * T;
  p_x += (rd * C::R_d + rv * C::R_v) * rho_x;
  double vr_x = c_help * 0.125 * ((rr_x * rm - rr * rm_x) / ((rm + rr) * (rm + rr)));
  vr_x *= std::pow(rr / (rm + rr), -0.875);
  const double rrvr_x = rr_x * vr + rr * vr_x;
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  s[0] = rd_t + 2 * rd_x;
  s[1] = rm_t + 2 * rm_x;
  s[2] = rr_t + 2 * rr_x - rrvr_x;
  s[3] = rho_x + p_x - rrvr_x;
  s[4] = rho_x + p_x - rrvr_x;
  double sE = E_t + 2 * (E_x + p_x) - (C::c_l * rho_x * rr * vr);
  sE -= (C::c_l * T + 1) * rrvr_x;
  s[5] = sE;
}

struct Mesh {
  int n, nx, ny, E;
  double dx, dy, Ja11, Ja22, invJ;
  std::vector<double> xi, w, D, Ds, dhat;
  explicit Mesh(const ora_rainy_params* p) : n(p->n), nx(p->nx), ny(p->ny), E(p->nx * p->ny) {
    dx = (p->x1 - p->x0) / nx;
    dy = (p->y1 - p->y0) / ny;
    // uniform Cartesian cells: Ja^1 = (dy/2, 0), Ja^2 = (0, dx/2), J = dx dy / 4
    Ja11 = 0.5 * dy;
    Ja22 = 0.5 * dx;
    invJ = 1.0 / (0.25 * dx * dy);
    xi.resize(n);
    w.resize(n);
    D.resize(n * n);
    Ds.resize(n * n);
    dhat.resize(n * n);
    ora_lgl_nodes_weights(n, xi.data(), w.data());
    ora_derivative_matrix(n, xi.data(), D.data());
    ora_derivative_split(n, xi.data(), w.data(), Ds.data());
    // Trixi's calc_dhat: dhat[j, m] = -D[m, j] w_m / w_j
    for (int j = 0; j < n; ++j)
      for (int m = 0; m < n; ++m) dhat[j * n + m] = -D[m * n + j] * w[m] / w[j];
  }
};

}  // namespace

extern "C" void ora_rainy_node_coordinates(const ora_rainy_params* p, double* xy) {
  const Mesh M(p);
  for (int ey = 0; ey < p->ny; ++ey)
    for (int ex = 0; ex < p->nx; ++ex)
      for (int j = 0; j < M.n; ++j)
        for (int i = 0; i < M.n; ++i) {
          const size_t q = i + M.n * (j + (size_t)M.n * (ex + p->nx * ey));
          xy[2 * q] = p->x0 + M.dx * (ex + 0.5 * (M.xi[i] + 1.0));
          xy[2 * q + 1] = p->y0 + M.dy * (ey + 0.5 * (M.xi[j] + 1.0));
        }
}

extern "C" int ora_rainy_rhs(const ora_rainy_params* p, const double* u, double t, double* du) {
  const Mesh M(p);
  g_surface = p->surface_flux;
  const int n = M.n, n2 = n * n, E = M.E;
  std::vector<double> xy((size_t)2 * n2 * E);
  ora_rainy_node_coordinates(p, xy.data());
  auto U = [&](int e, int i, int j) { return u + (size_t)NV * (i + n * (j + (size_t)n * e)); };
  auto DU = [&](int e, int i, int j) { return du + (size_t)NV * (i + n * (j + (size_t)n * e)); };
  // surface_flux_values [v, l, direction 0..3, e]
  std::vector<double> sfv((size_t)NV * n * 4 * E, 0.0);

#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    std::memset(DU(e, 0, 0), 0, sizeof(double) * NV * n2);
    if (p->volume_integral == 0) {
      // weak form (Trixi weak_form_kernel! for curved 2D meshes)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double f1[NV], f2[NV], cf[NV];
          flux_orientation(U(e, i, j), 1, f1);
          flux_orientation(U(e, i, j), 2, f2);
          for (int v = 0; v < NV; ++v) cf[v] = M.Ja11 * f1[v] + 0.0 * f2[v];
          for (int ii = 0; ii < n; ++ii)
            for (int v = 0; v < NV; ++v) DU(e, ii, j)[v] += M.dhat[ii * n + i] * cf[v];
          for (int v = 0; v < NV; ++v) cf[v] = 0.0 * f1[v] + M.Ja22 * f2[v];
          for (int jj = 0; jj < n; ++jj)
            for (int v = 0; v < NV; ++v) DU(e, i, jj)[v] += M.dhat[jj * n + j] * cf[v];
        }
    } else {
      // flux differencing with flux_ec_rain and the averaged contravariant vectors
      const double n1[2] = {M.Ja11, 0.0}, n2v[2] = {0.0, M.Ja22};
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          for (int ii = i + 1; ii < n; ++ii) {
            double f[NV];
            ora_rainy_flux_ec(U(e, i, j), U(e, ii, j), n1, f);
            for (int v = 0; v < NV; ++v) {
              DU(e, i, j)[v] += M.Ds[i * n + ii] * f[v];
              DU(e, ii, j)[v] += M.Ds[ii * n + i] * f[v];
            }
          }
          for (int jj = j + 1; jj < n; ++jj) {
            double f[NV];
            ora_rainy_flux_ec(U(e, i, j), U(e, i, jj), n2v, f);
            for (int v = 0; v < NV; ++v) {
              DU(e, i, j)[v] += M.Ds[j * n + jj] * f[v];
              DU(e, i, jj)[v] += M.Ds[jj * n + j] * f[v];
            }
          }
        }
    }
  }

  // interfaces (Trixi calc_interface_flux! for StructuredMesh{2}): every
  // element computes the flux on its -x and -y faces with its own Ja
  for (int e = 0; e < E; ++e) {
    const int ex = e % M.nx, ey = e / M.nx;
    // -x face: left neighbour (periodic in x)
    {
      const int el = ((ex + M.nx - 1) % M.nx) + M.nx * ey;
      const double nrm[2] = {M.Ja11, 0.0};
      for (int l = 0; l < n; ++l) {
        double f[NV];
        surface(U(el, n - 1, l), U(e, 0, l), nrm[0], nrm[1], f);
        for (int v = 0; v < NV; ++v) {
          sfv[(size_t)NV * (l + n * (1 + 4 * (size_t)el)) + v] = f[v];
          sfv[(size_t)NV * (l + n * (0 + 4 * (size_t)e)) + v] = f[v];
        }
      }
    }
    // -y face
    const double nrm[2] = {0.0, M.Ja22};
    if (ey > 0 || p->periodic_y) {
      const int eb = ex + M.nx * ((ey + M.ny - 1) % M.ny);
      for (int l = 0; l < n; ++l) {
        double f[NV];
        surface(U(eb, l, n - 1), U(e, l, 0), nrm[0], nrm[1], f);
        for (int v = 0; v < NV; ++v) {
          sfv[(size_t)NV * (l + n * (3 + 4 * (size_t)eb)) + v] = f[v];
          sfv[(size_t)NV * (l + n * (2 + 4 * (size_t)e)) + v] = f[v];
        }
      }
    }
  }
  if (!p->periodic_y) {
    // boundary_condition_simple_slip_wall (compressible_rainy_euler_2d.jl:963-983):
    // the mirrored state with -rho_v2, the surface flux with the outer state
    // on the right at the top (direction 4) and on the left at the bottom
    const double nrm[2] = {0.0, M.Ja22};
    for (int ex = 0; ex < M.nx; ++ex) {
      const int eb = ex, et = ex + M.nx * (M.ny - 1);
      for (int l = 0; l < n; ++l) {
        double ub[NV], f[NV];
        std::memcpy(ub, U(eb, l, 0), sizeof(ub));
        ub[4] = -ub[4];
        surface(ub, U(eb, l, 0), nrm[0], nrm[1], f);
        for (int v = 0; v < NV; ++v) sfv[(size_t)NV * (l + n * (2 + 4 * (size_t)eb)) + v] = f[v];
        std::memcpy(ub, U(et, l, n - 1), sizeof(ub));
        ub[4] = -ub[4];
        surface(U(et, l, n - 1), ub, nrm[0], nrm[1], f);
        for (int v = 0; v < NV; ++v) sfv[(size_t)NV * (l + n * (3 + 4 * (size_t)et)) + v] = f[v];
      }
    }
  }

  // surface integral, Jacobian, sources
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    const double* s = sfv.data() + (size_t)NV * n * 4 * e;
    for (int l = 0; l < n; ++l)
      for (int v = 0; v < NV; ++v) {
        DU(e, 0, l)[v] -= s[NV * (l + n * 0) + v] * (1.0 / M.w[0]);
        DU(e, n - 1, l)[v] += s[NV * (l + n * 1) + v] * (1.0 / M.w[n - 1]);
        DU(e, l, 0)[v] -= s[NV * (l + n * 2) + v] * (1.0 / M.w[0]);
        DU(e, l, n - 1)[v] += s[NV * (l + n * 3) + v] * (1.0 / M.w[n - 1]);
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double* d = DU(e, i, j);
        for (int v = 0; v < NV; ++v) d[v] *= -M.invJ;
        if (p->source) {
          double sv[NV];
          const size_t q = i + n * (j + (size_t)n * e);
          if (p->source == 1) ora_rainy_source(U(e, i, j), sv);
          else if (p->source == 2) source_convergence(xy[2 * q], xy[2 * q + 1], t, sv);
          else {  // source_terms_no_phase_change (:579-593)
            const Dens dd = densities(U(e, i, j));
            for (int v = 0; v < NV; ++v) sv[v] = 0.0;
            sv[4] = -C::g * dd.rho;
            sv[5] = -C::g * U(e, i, j)[4];
          }
          for (int v = 0; v < NV; ++v) d[v] += sv[v];
        }
      }
  }
  return 0;
}

// max_dt (stepsize_dg2d.jl:5-43 pattern for StructuredMesh{2}) with
// max_abs_speeds (:595-600): (|v1| + c, |v2| + c + |v_r|)
extern "C" double ora_rainy_max_dt(const ora_rainy_params* p, const double* u) {
  const Mesh M(p);
  const int n2 = M.n * M.n;
  double max_scaled = 5e-324;
  for (int e = 0; e < M.E; ++e) {
    double l1m = 0, l2m = 0;
    for (int q = 0; q < n2; ++q) {
      const double* un = u + (size_t)NV * (q + (size_t)n2 * e);
      const Dens d = densities(un);
      const double c = speed_of_sound(un, nullptr);
      const double vr = terminal_velocity_rain(d.moist, d.rain);
      const double lam1 = std::fabs(un[3] * d.inv) + c;
      const double lam2 = std::fabs(un[4] * d.inv) + c + std::fabs(vr);
      l1m = std::fmax(l1m, std::fabs(M.Ja11 * lam1 + 0.0 * lam2) * M.invJ);
      l2m = std::fmax(l2m, std::fabs(0.0 * lam1 + M.Ja22 * lam2) * M.invJ);
    }
    max_scaled = std::fmax(max_scaled, l1m + l2m);
  }
  return 2.0 / (M.n * max_scaled);
}

// NonlinearSolveDG over all nodes; returns the number of nodes that did not
// converge; iterations[node] (may be NULL) receives the update count
extern "C" int ora_rainy_newton(const ora_rainy_params* p, double* u, double tol, int max_iter,
                                int* iterations) {
  const size_t nodes = (size_t)p->n * p->n * p->nx * p->ny;
  int fails = 0;
#pragma omp parallel for schedule(static) reduction(+ : fails)
  for (size_t q = 0; q < nodes; ++q) {
    const int it = ora_rainy_newton_node(u + NV * q, nullptr, tol, max_iter);
    if (iterations) iterations[q] = it;
    if (it < 0) ++fails;
  }
  return fails;
}

// CarpenterKennedy2N54(williamson_condition = false, stage_limiter!) steps:
// after every stage update (the last one included) the Newton solve runs on
// the updated state, so every rhs sees a solved state (OrdinaryDiffEq's 2N
// low-storage perform_step! calls stage_limiter! after each stage)
extern "C" int ora_rainy_ck2n54(const ora_rainy_params* p, double* u, double t, double dt,
                                int n_steps, double tol, int max_iter) {
  double A[5], B[5], c[5];
  ora_ck2n54_coefficients(A, B, c);
  const size_t N = (size_t)NV * p->n * p->n * p->nx * p->ny;
  std::vector<double> du(N), k(N, 0.0);
  for (int s = 0; s < n_steps; ++s) {
    for (int st = 0; st < 5; ++st) {
      ora_rainy_rhs(p, u, t + c[st] * dt, du.data());
      for (size_t q = 0; q < N; ++q) {
        k[q] = A[st] * k[q] + dt * du[q];
        u[q] += B[st] * k[q];
      }
      ora_rainy_newton(p, u, tol, max_iter, nullptr);
    }
    t += dt;
  }
  return 0;
}

// cons2entropy (:131-170): entropy variables of the first six entries
extern "C" void ora_rainy_cons2entropy(const double* u, double* w) {
  const Dens d = densities(u);
  double s[3];
  nonlinear_sol(u, s);
  const double lnT = std::log(s[2]), iT = 1.0 / s[2];
  const double v1 = u[3] * d.inv, v2 = u[4] * d.inv;
  const double vsq = 0.5 * (v1 * v1 + v2 * v2) * iT;
  const double lv = s[0] > 0.0 ? std::log(s[0]) : 0.0;
  const double om_d = C::c_vd * lnT - C::R_d * std::log(d.dry) + vsq - C::c_vd - C::R_d;
  const double om_v = C::c_vv * lnT - C::R_v * lv + vsq - C::c_vv - C::R_v - C::L_ref * iT;
  const double om_l = C::c_l * lnT + vsq - C::c_l;
  w[0] = om_d;
  w[1] = om_v + om_l;
  w[2] = om_l;
  w[3] = -v1 * iT;
  w[4] = -v2 * iT;
  w[5] = iT;
  w[6] = w[7] = w[8] = 0.0;
}

// out[0..8]: integral of every variable; out[9]: int w(u) . du (du may be NULL)
extern "C" void ora_rainy_integrals(const ora_rainy_params* p, const double* u, const double* du,
                                    double* out) {
  const Mesh M(p);
  const int n = M.n;
  for (int k = 0; k < 10; ++k) out[k] = 0.0;
  const double J = 1.0 / M.invJ;
  for (int e = 0; e < M.E; ++e)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t q = i + n * (j + (size_t)n * e);
        const double wJ = M.w[i] * M.w[j] * J;
        for (int v = 0; v < NV; ++v) out[v] += wJ * u[NV * q + v];
        if (du) {
          double w[NV];
          ora_rainy_cons2entropy(u + NV * q, w);
          for (int v = 0; v < 6; ++v) out[9] += wJ * w[v] * du[NV * q + v];
        }
      }
}
