// CPU ORACLE -- test infrastructure only (never linked by the product).
//
// FP64 restatement of the covariant `rhs_hyperbolic!` of TrixiAtmo for
// SplitCovariantShallowWaterEquations2D, and for the standard
// CovariantShallowWaterEquations2D with the weak form (covariant_shallow_water.jl:166-238),
// on 2D P4est meshes of the sphere:
//   src/solvers/dgsem_p4est/dg_2d_manifold_in_3d_covariant.jl:8-71 (phases),
//   :100-130 (weak form), :134-234 (flux differencing, symmetric + nonconservative),
//   :239-432 (interface flux in the local contravariant bases of both sides),
//   :435-458 (sources with the auxiliary variables), :461-478 (exact J),
// with the pointwise physics of src/equations/covariant_shallow_water_split.jl
// and the covariant flux/dissipation wrappers of src/equations/equations.jl:256-290.
// Auxiliary variables follow equations.jl:178-250 (26 per node).
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int NV = 3, NA = 26;

struct Aux {
  const double* a;
  double A(int r, int c) const { return a[r + 3 * c]; }      // basis_covariant 3x2
  double B(int r, int c) const { return a[6 + r + 2 * c]; }  // basis_contravariant 2x3
  double J() const { return a[12]; }
  double Gcov(int i, int j) const { return (i == 0 && j == 0) ? a[13] : (i == 1 && j == 1) ? a[15] : a[14]; }
  double Gcon(int i, int j) const { return (i == 0 && j == 0) ? a[16] : (i == 1 && j == 1) ? a[18] : a[17]; }
  double hs() const { return a[19]; }
  double Gam(int k, int i, int j) const {  // Gamma^k_ij
    const int o = 20 + 3 * k;
    return (i == 0 && j == 0) ? a[o] : (i == 1 && j == 1) ? a[o + 2] : a[o + 1];
  }
};

// covariant_shallow_water_split.jl:66-80 (pointwise flux without pressure)
void flux_split(const double* u, const Aux& x, int o, double* f) {
  const double mass = x.J() * u[1 + o];
  f[0] = mass;
  f[1] = mass * (u[1] / u[0]);
  f[2] = mass * (u[2] / u[0]);
}

// :100-122 (flux_ec, orientation)
void flux_ec(const double* ul, const double* ur, const Aux& xl, const Aux& xr, int o,
             double* f) {
  const double mass = 0.5 * (xl.J() * ul[1 + o] + xr.J() * ur[1 + o]);
  f[0] = mass;
  f[1] = 0.5 * (ul[1] / ul[0] + ur[1] / ur[0]) * mass;
  f[2] = 0.5 * (ul[2] / ul[0] + ur[2] / ur[0]) * mass;
}

// equations.jl:273-279
void flux_central(const double* ul, const double* ur, const Aux& xl, const Aux& xr, int o,
                  double* f) {
  double fl[NV], fr[NV];
  flux_split(ul, xl, o, fl);
  flux_split(ur, xr, o, fr);
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]);
}

// :157-187 (flux_nonconservative_ec, orientation)
void noncons_ec(double g, const double* ul, const double* ur, const Aux& xl, const Aux& xr, int o,
                double* f) {
  const double hl = ul[0], hr = ur[0];
  const double vr[2] = {ur[1] / hr, ur[2] / hr};
  const double vcovr[2] = {xr.Gcov(0, 0) * vr[0] + xr.Gcov(0, 1) * vr[1],
                           xr.Gcov(1, 0) * vr[0] + xr.Gcov(1, 1) * vr[1]};
  const double hs_jump = xr.hs() - xl.hs();
  double geo[2], pres[2];
  for (int k = 0; k < 2; ++k) {
    geo[k] = 0.5 * ul[1 + o] * ((xl.Gcon(k, 0) * vcovr[0] + xl.Gcon(k, 1) * vcovr[1]) - vr[k]);
    pres[k] = g * xl.Gcon(k, o) * hl * (hr + hs_jump);
  }
  f[0] = 0.0;
  f[1] = xl.J() * (geo[0] + pres[0]);
  f[2] = xl.J() * (geo[1] + pres[1]);
}

// :229-250 (flux_nonconservative_surface_simplified, orientation)
void noncons_surface_simplified(double g, const double* ul, const double* ur, const Aux& xl, int o,
                                double* f) {
  f[0] = 0.0;
  f[1] = xl.J() * (g * xl.Gcon(0, o) * ul[0] * ur[0]);
  f[2] = xl.J() * (g * xl.Gcon(1, o) * ul[0] * ur[0]);
}

// covariant_shallow_water.jl:246-265 (max_abs_speed, orientation)
double max_abs_speed(double g, const double* ul, const double* ur, const Aux& xl, const Aux& xr,
                     int o) {
  return std::fmax(std::fabs(ul[1 + o] / ul[0]) + std::sqrt(xl.Gcon(o, o) * ul[0] * g),
                   std::fabs(ur[1 + o] / ur[0]) + std::sqrt(xr.Gcon(o, o) * ur[0] * g));
}

// CovariantShallowWaterEquations2D (the standard, non-split form):
// covariant_shallow_water.jl:166-187, flux in orientation o with the pressure
// term G^{io} g h^2 / 2 inside the momentum flux
void flux_standard(double g, const double* u, const Aux& x, int o, double* f) {
  const double h = u[0];
  const double vcon = u[1 + o] / h;
  const double grav = 0.5 * g * h * h;
  const double m1 = u[1] * vcon + x.Gcon(0, o) * grav;
  const double m2 = u[2] * vcon + x.Gcon(1, o) * grav;
  f[0] = x.J() * u[1 + o];
  f[1] = x.J() * m1;
  f[2] = x.J() * m2;
}

// :215-238, source_terms_geometric_coriolis of the standard form
void source_standard(double g, double omega, const double* u, const double* x, const Aux& a,
                     double* s) {
  const double h = u[0];
  const double hv[2] = {u[1], u[2]};
  const double v[2] = {u[1] / h, u[2] / h};
  const double grav = 0.5 * g * h * h;
  double tau[2][2];  // h v^i v^j + G^{ij} g h^2 / 2
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) tau[i][j] = hv[i] * v[j] + grav * a.Gcon(i, j);
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double f = 2 * omega * x[2] / r;
  double src[2];
  for (int k = 0; k < 2; ++k) {
    double sg = 0.0;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) sg += a.Gam(k, i, j) * tau[i][j];
    src[k] = sg + f * a.J() * (a.Gcon(k, 1) * hv[0] - a.Gcon(k, 0) * hv[1]);
  }
  s[0] = 0.0;
  s[1] = -src[0];
  s[2] = -src[1];
}

struct Fluxes {
  double g;
  int volume_flux;      // 0 flux_ec, 1 flux_central, 2 weak form of the standard form
  int surface_flux;     // 0 FluxPlusDissipation(flux_ec, LLF), 1 flux_lax_friedrichs, 2 flux_ec,
                        // 3 flux_lax_friedrichs of the standard form
  int surface_noncons;  // 0 flux_nonconservative_surface_simplified, 1 flux_nonconservative_ec
  bool standard() const { return volume_flux == 2; }
};

void volume_flux(const Fluxes& F, const double* ul, const double* ur, const Aux& xl,
                 const Aux& xr, int o, double* f) {
  if (F.volume_flux == 0) flux_ec(ul, ur, xl, xr, o, f);
  else flux_central(ul, ur, xl, xr, o, f);
}

// the surface flux called with (u_ll, u_rr, aux, aux): equations.jl:256-290
void surface_flux(const Fluxes& F, const double* ul, const double* ur, const Aux& xl,
                  const Aux& xr, int o, double* f) {
  if (F.surface_flux == 3) {  // flux_central of the standard flux (equations.jl:273-280)
    double fl[NV], fr[NV];
    flux_standard(F.g, ul, xl, o, fl);
    flux_standard(F.g, ur, xr, o, fr);
    for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]);
  } else if (F.surface_flux == 1) flux_central(ul, ur, xl, xr, o, f);
  else flux_ec(ul, ur, xl, xr, o, f);
  if (F.surface_flux != 2) {
    const double lam = max_abs_speed(F.g, ul, ur, xl, xr, o);
    for (int v = 0; v < NV; ++v) f[v] += -0.5 * xl.J() * lam * (ur[v] - ul[v]);
  }
}

void surface_noncons(const Fluxes& F, const double* ul, const double* ur, const Aux& xl,
                     const Aux& xr, int o, double* f) {
  if (F.standard()) {  // the standard form has no nonconservative terms
    for (int v = 0; v < NV; ++v) f[v] = 0.0;
  } else if (F.surface_noncons == 0) noncons_surface_simplified(F.g, ul, ur, xl, o, f);
  else noncons_ec(F.g, ul, ur, xl, xr, o, f);
}

// covariant_shallow_water_split.jl:273-302 (source_terms_geometric_coriolis)
void source_split(double omega, const double* u, const double* x, const Aux& a, double* s) {
  const double hv[2] = {u[1], u[2]};
  const double v[2] = {u[1] / u[0], u[2] / u[0]};
  double hvv[2][2], hcv[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) hvv[i][j] = hv[i] * v[j];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) hcv[i][j] = a.Gcov(i, 0) * hvv[0][j] + a.Gcov(i, 1) * hvv[1][j];
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double f = 2 * omega * x[2] / r;
  double t1[2], t2[2];  // sum(Gamma^k .* hvv)
  t1[0] = a.Gam(0, 0, 0) * hvv[0][0] + a.Gam(0, 1, 0) * hvv[1][0] + a.Gam(0, 0, 1) * hvv[0][1] +
          a.Gam(0, 1, 1) * hvv[1][1];
  t1[1] = a.Gam(1, 0, 0) * hvv[0][0] + a.Gam(1, 1, 0) * hvv[1][0] + a.Gam(1, 0, 1) * hvv[0][1] +
          a.Gam(1, 1, 1) * hvv[1][1];
  // Gamma1 * hcv[0, :] + Gamma2 * hcv[1, :]
  for (int i = 0; i < 2; ++i)
    t2[i] = a.Gam(0, i, 0) * hcv[0][0] + a.Gam(0, i, 1) * hcv[0][1] + a.Gam(1, i, 0) * hcv[1][0] +
            a.Gam(1, i, 1) * hcv[1][1];
  double sgeo[2];
  for (int k = 0; k < 2; ++k)
    sgeo[k] = 0.5 * (t1[k] - (a.Gcon(k, 0) * t2[0] + a.Gcon(k, 1) * t2[1]));
  const double s1 = sgeo[0] + f * a.J() * (a.Gcon(0, 1) * hv[0] - a.Gcon(0, 0) * hv[1]);
  const double s2 = sgeo[1] + f * a.J() * (a.Gcon(1, 1) * hv[0] - a.Gcon(1, 0) * hv[1]);
  s[0] = 0.0;
  s[1] = -s1;
  s[2] = -s2;
}

// contravariant2global then global2contravariant (covariant_shallow_water.jl:140-153)
void transform(const double* u, const Aux& from, const Aux& to, double* out) {
  double gl[3];
  for (int r = 0; r < 3; ++r) gl[r] = from.A(r, 0) * u[1] + from.A(r, 1) * u[2];
  out[0] = u[0];
  out[1] = to.B(0, 0) * gl[0] + to.B(0, 1) * gl[1] + to.B(0, 2) * gl[2];
  out[2] = to.B(1, 0) * gl[0] + to.B(1, 1) * gl[1] + to.B(1, 2) * gl[2];
}

inline int code2d(int code, int i, int n) {
  switch (code) {
    case 0: return 0;
    case 1: return n - 1;
    case 2: return i;
    default: return n - 1 - i;
  }
}

inline int dir2d(const int* idx) {
  if (idx[0] <= 1) return idx[0];         // :begin -> 0 (-x), :end -> 1 (+x)
  return 2 + idx[1];                       // :begin -> 2 (-y), :end -> 3 (+y)
}

}  // namespace

// mag = true: every contribution enters with its absolute value, giving the
// scale that bounds the rounding error of each tendency entry
static int swe_rhs_impl(const ora_swe_params* p, const ora_connectivity2d* c, const double* aux,
                        const double* xyz, const double* u, double* du, bool mag) {
  auto M = [mag](double v) { return mag ? std::fabs(v) : v; };
  const int n = p->n, n2 = n * n, E = p->n_elements;
  const Fluxes F{p->gravity, p->volume_flux, p->surface_flux, p->surface_noncons};
  std::vector<double> x(n), w(n), Ds(n * n), D(n * n), dhat(n * n);
  ora_lgl_nodes_weights(n, x.data(), w.data());
  ora_derivative_split(n, x.data(), w.data(), Ds.data());
  ora_derivative_matrix(n, x.data(), D.data());
  // Trixi's derivative_hat: dhat[j, m] = -D[m, j] w_m / w_j
  for (int j = 0; j < n; ++j)
    for (int m = 0; m < n; ++m) dhat[j * n + m] = -D[m * n + j] * w[m] / w[j];
  std::vector<double> sfv((size_t)NV * n * 4 * E, 0.0);
  auto U = [&](int e, int i, int j) { return u + (size_t)NV * (i + n * (j + (size_t)n * e)); };
  auto AX = [&](int e, int i, int j) { return Aux{aux + (size_t)NA * (i + n * (j + (size_t)n * e))}; };

#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    double* due = du + (size_t)NV * n2 * e;
    std::memset(due, 0, sizeof(double) * NV * n2);
    if (F.standard()) {
      // weak form (dg_2d_manifold_in_3d_covariant.jl:100-130): node (i, j)'s
      // contravariant fluxes enter node (ii, j) with dhat[ii, i], (i, jj) with dhat[jj, j]
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double f1[NV], f2[NV];
          flux_standard(F.g, U(e, i, j), AX(e, i, j), 0, f1);
          flux_standard(F.g, U(e, i, j), AX(e, i, j), 1, f2);
          for (int ii = 0; ii < n; ++ii)
            for (int v = 0; v < NV; ++v) due[NV * (ii + n * j) + v] += M(dhat[ii * n + i] * f1[v]);
          for (int jj = 0; jj < n; ++jj)
            for (int v = 0; v < NV; ++v) due[NV * (i + n * jj) + v] += M(dhat[jj * n + j] * f2[v]);
        }
      continue;
    }
    // symmetric part (dg_2d_manifold_in_3d_covariant.jl:134-183)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        for (int ii = i + 1; ii < n; ++ii) {
          double f[NV];
          volume_flux(F, U(e, i, j), U(e, ii, j), AX(e, i, j), AX(e, ii, j), 0, f);
          for (int v = 0; v < NV; ++v) {
            due[NV * (i + n * j) + v] += M(Ds[i * n + ii] * f[v]);
            due[NV * (ii + n * j) + v] += M(Ds[ii * n + i] * f[v]);
          }
        }
        for (int jj = j + 1; jj < n; ++jj) {
          double f[NV];
          volume_flux(F, U(e, i, j), U(e, i, jj), AX(e, i, j), AX(e, i, jj), 1, f);
          for (int v = 0; v < NV; ++v) {
            due[NV * (i + n * j) + v] += M(Ds[j * n + jj] * f[v]);
            due[NV * (i + n * jj) + v] += M(Ds[jj * n + j] * f[v]);
          }
        }
      }
    // nonconservative part (:188-234)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double acc[NV] = {0, 0, 0};
        for (int ii = 0; ii < n; ++ii) {
          if (ii == i) continue;  // Dsplit has a vanishing diagonal
          double f[NV];
          noncons_ec(F.g, U(e, i, j), U(e, ii, j), AX(e, i, j), AX(e, ii, j), 0, f);
          for (int v = 0; v < NV; ++v) acc[v] += M(Ds[i * n + ii] * f[v]);
        }
        for (int jj = 0; jj < n; ++jj) {
          if (jj == j) continue;
          double f[NV];
          noncons_ec(F.g, U(e, i, j), U(e, i, jj), AX(e, i, j), AX(e, i, jj), 1, f);
          for (int v = 0; v < NV; ++v) acc[v] += M(Ds[j * n + jj] * f[v]);
        }
        for (int v = 0; v < NV; ++v) due[NV * (i + n * j) + v] += 0.5 * acc[v];
      }
  }

  // interfaces (:239-432): each side's flux in its own contravariant basis
#pragma omp parallel for schedule(static)
  for (int I = 0; I < c->n_interfaces; ++I) {
    const int ep = c->neighbor_ids[2 * I], es = c->neighbor_ids[2 * I + 1];
    const int* ip = c->node_indices + 4 * I;
    const int* is = ip + 2;
    const int dp = dir2d(ip), ds = dir2d(is);
    const int sp_code = ip[0] >= 2 ? ip[0] : ip[1];
    const int ss_code = is[0] >= 2 ? is[0] : is[1];
    for (int k = 0; k < n; ++k) {
      const int pi = code2d(ip[0], k, n), pj = code2d(ip[1], k, n);
      const int si = code2d(is[0], k, n), sj = code2d(is[1], k, n);
      const double* ull = U(ep, pi, pj);
      const double* urr = U(es, si, sj);
      const Aux all = AX(ep, pi, pj), arr = AX(es, si, sj);
      double fp[NV], fs[NV], t[NV], f[NV], g[NV];
      // primary frame
      transform(urr, arr, all, t);
      const int op = dp / 2;
      if (dp % 2 == 0) {
        surface_flux(F, t, ull, all, all, op, f);
        surface_noncons(F, t, ull, all, all, op, g);
        for (int v = 0; v < NV; ++v) fp[v] = -(f[v] + 0.5 * g[v]);
      } else {
        surface_flux(F, ull, t, all, all, op, f);
        surface_noncons(F, ull, t, all, all, op, g);
        for (int v = 0; v < NV; ++v) fp[v] = f[v] + 0.5 * g[v];
      }
      // secondary frame
      transform(ull, all, arr, t);
      const int os = ds / 2;
      if (ds % 2 == 0) {
        surface_flux(F, t, urr, arr, arr, os, f);
        surface_noncons(F, t, urr, arr, arr, os, g);
        for (int v = 0; v < NV; ++v) fs[v] = -(f[v] + 0.5 * g[v]);
      } else {
        surface_flux(F, urr, t, arr, arr, os, f);
        surface_noncons(F, urr, t, arr, arr, os, g);
        for (int v = 0; v < NV; ++v) fs[v] = f[v] + 0.5 * g[v];
      }
      const int kp = code2d(sp_code, k, n), ks = code2d(ss_code, k, n);
      for (int v = 0; v < NV; ++v) {
        sfv[(size_t)NV * (kp + n * (dp + 4 * (size_t)ep)) + v] = fp[v];
        sfv[(size_t)NV * (ks + n * (ds + 4 * (size_t)es)) + v] = fs[v];
      }
    }
  }

  // surface integral (outward fluxes added with 1/w_end), exact Jacobian (-1/J),
  // then the sources
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    double* due = du + (size_t)NV * n2 * e;
    const double* s = sfv.data() + (size_t)NV * n * 4 * e;
    for (int l = 0; l < n; ++l)
      for (int v = 0; v < NV; ++v) {
        due[NV * (0 + n * l) + v] += M(s[NV * (l + n * 0) + v] / w[0]);
        due[NV * ((n - 1) + n * l) + v] += M(s[NV * (l + n * 1) + v] / w[n - 1]);
        due[NV * (l + n * 0) + v] += M(s[NV * (l + n * 2) + v] / w[0]);
        due[NV * (l + n * (n - 1)) + v] += M(s[NV * (l + n * 3) + v] / w[n - 1]);
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const Aux a = AX(e, i, j);
        const double fac = -1.0 / a.J();
        double* d = due + NV * (i + n * j);
        for (int v = 0; v < NV; ++v) d[v] *= mag ? std::fabs(fac) : fac;
        if (p->source_terms) {
          double sv[NV];
          const double* xn = xyz + 3 * (i + n * (j + (size_t)n * e));
          if (p->source_terms == 2) source_standard(p->gravity, p->rotation_rate, U(e, i, j), xn, a, sv);
          else source_split(p->rotation_rate, U(e, i, j), xn, a, sv);
          for (int v = 0; v < NV; ++v) d[v] += M(sv[v]);
        }
      }
  }
  return 0;
}

extern "C" int ora_swe_rhs(const ora_swe_params* p, const ora_connectivity2d* c,
                           const double* aux, const double* xyz, const double* u, double* du) {
  return swe_rhs_impl(p, c, aux, xyz, u, du, false);
}

extern "C" int ora_swe_rhs_magnitude(const ora_swe_params* p, const ora_connectivity2d* c,
                                     const double* aux, const double* xyz, const double* u,
                                     double* du_abs) {
  return swe_rhs_impl(p, c, aux, xyz, u, du_abs, true);
}

// StepsizeCallback max_dt, covariant form (src/callbacks_step/stepsize_dg2d.jl:47-71)
extern "C" double ora_swe_max_dt(const ora_swe_params* p, const double* aux, const double* u) {
  const int n = p->n, n2 = n * n;
  double max_scaled = 5e-324;
  for (int e = 0; e < p->n_elements; ++e) {
    double l1m = 0, l2m = 0;
    for (int q = 0; q < n2; ++q) {
      const double* un = u + (size_t)NV * (q + (size_t)n2 * e);
      const Aux a{aux + (size_t)NA * (q + (size_t)n2 * e)};
      const double v1 = un[1] / un[0], v2 = un[2] / un[0];
      l1m = std::fmax(l1m, std::fabs(v1) + std::sqrt(a.Gcon(0, 0) * un[0] * p->gravity));
      l2m = std::fmax(l2m, std::fabs(v2) + std::sqrt(a.Gcon(1, 1) * un[0] * p->gravity));
    }
    max_scaled = std::fmax(max_scaled, l1m + l2m);
  }
  return 2.0 / (n * max_scaled);
}

extern "C" int ora_swe_ck2n54(const ora_swe_params* p, const ora_connectivity2d* c,
                              const double* aux, const double* xyz, double* u, double dt,
                              int n_steps) {
  double A[5], Bc[5], cc[5];
  ora_ck2n54_coefficients(A, Bc, cc);
  const size_t N = (size_t)NV * p->n * p->n * p->n_elements;
  std::vector<double> du(N), k(N, 0.0);
  for (int step = 0; step < n_steps; ++step)
    for (int s = 0; s < 5; ++s) {
      ora_swe_rhs(p, c, aux, xyz, u, du.data());
      for (size_t q = 0; q < N; ++q) {
        k[q] = A[s] * k[q] + dt * du[q];
        u[q] += Bc[s] * k[q];
      }
    }
  return 0;
}

// entropy (total energy) and its time derivative: covariant_shallow_water.jl:129-164,
// analysis_covariant.jl:53-105 (out[0] mass, out[1] entropy, out[2] dS/dt)
extern "C" void ora_swe_integrals(const ora_swe_params* p, const double* aux, const double* u,
                                  const double* du, double* out) {
  const int n = p->n;
  std::vector<double> x(n), w(n);
  ora_lgl_nodes_weights(n, x.data(), w.data());
  out[0] = out[1] = out[2] = 0.0;
  for (int e = 0; e < p->n_elements; ++e)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t q = i + n * (j + (size_t)n * e);
        const Aux a{aux + NA * q};
        const double* un = u + NV * q;
        const double h = un[0], v1 = un[1] / h, v2 = un[2] / h;
        const double vc1 = a.Gcov(0, 0) * v1 + a.Gcov(0, 1) * v2;
        const double vc2 = a.Gcov(1, 0) * v1 + a.Gcov(1, 1) * v2;
        const double wJ = w[i] * w[j] * a.J();
        out[0] += wJ * h;
        out[1] += wJ * (0.5 * (h * (vc1 * v1 + vc2 * v2) + p->gravity * h * h) + p->gravity * h * a.hs());
        if (du) {
          const double w1 = p->gravity * (h + a.hs()) - 0.5 * (vc1 * v1 + vc2 * v2);
          out[2] += wJ * (w1 * du[NV * q] + vc1 * du[NV * q + 1] + vc2 * du[NV * q + 2]);
        }
      }
}

// SaveSolutionCallback's solution_variables = cons2prim_and_vorticity
// (src/callbacks_step/save_solution_covariant.jl:65-76): per node the total
// height H = h + h_s (cons2prim, covariant_shallow_water.jl:113-118), the
// contravariant velocity mapped to global components by the covariant basis
// (contravariant2global, :140-145), h_s, and the relative vorticity of
// calc_vorticity_node (save_solution_covariant.jl:92-117): the collocation
// derivatives d1 of v_2 along i and d2 of v_1 along j, v_i = G_ij vcon^j,
// divided by J.  out[6][nodes], variable-major.
extern "C" void ora_swe_prim_vorticity(const ora_swe_params* p, const double* aux, const double* u,
                                       double* out) {
  const int n = p->n;
  const size_t n_nodes = (size_t)n * n * p->n_elements;
  std::vector<double> x(n), w(n), D(n * n);
  ora_lgl_nodes_weights(n, x.data(), w.data());
  ora_derivative_matrix(n, x.data(), D.data());
  auto vcov = [&](size_t q, int c) {
    const Aux a{aux + NA * q};
    const double* un = u + NV * q;
    const double vcon[2] = {un[1] / un[0], un[2] / un[0]};
    return a.Gcov(c, 0) * vcon[0] + a.Gcov(c, 1) * vcon[1];
  };
  for (int e = 0; e < p->n_elements; ++e)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t base = (size_t)n * n * e, q = base + i + n * j;
        const Aux a{aux + NA * q};
        const double* un = u + NV * q;
        const double h = un[0], v1 = un[1] / h, v2 = un[2] / h;
        double dv2dxi1 = 0.0, dv1dxi2 = 0.0;
        for (int m = 0; m < n; ++m) {
          dv2dxi1 += D[i * n + m] * vcov(base + m + n * j, 1);
          dv1dxi2 += D[j * n + m] * vcov(base + i + n * m, 0);
        }
        const double w[6] = {h + a.hs(), a.A(0, 0) * v1 + a.A(0, 1) * v2,
                             a.A(1, 0) * v1 + a.A(1, 1) * v2, a.A(2, 0) * v1 + a.A(2, 1) * v2,
                             a.hs(), (dv2dxi1 - dv1dxi2) / a.J()};
        for (int v = 0; v < 6; ++v) out[v * n_nodes + q] = w[v];
      }
}
