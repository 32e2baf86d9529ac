// CPU ORACLE -- test infrastructure only (never linked by the product).
//
// FP64 restatement of the DGSEM `rhs!` for
// CompressibleEulerEnergyEquationsWithGravity3D on curved P4est hexahedral
// meshes, following the structure TrixiAtmo overrides for its covariant solver
// (src/solvers/dgsem_p4est/dg_2d_manifold_in_3d_covariant.jl:8-71) and the
// 3D curvilinear Trixi kernels it mirrors [ext]:
//   set_zero -> volume flux differencing (symmetric + nonconservative)
//   -> prolong2interfaces -> interface flux -> boundary flux
//   -> surface integral -> apply_jacobian (times -1/J).
// Pointwise physics restates
// src/equations/compressible_euler_energy_with_gravity_3d.jl.
#include <cmath>
#include <cstring>
#include <vector>

#include <algorithm>

#include "counted.h"
#include "oracle.h"

OraCounts& ora_counts() {
  static OraCounts c;
  return c;
}

namespace {

// OpenMP for the FP64 oracle; counting runs are serial (plain counters)
template <typename R>
constexpr bool kParallel = true;
template <>
constexpr bool kParallel<Counted> = false;

constexpr int NV = 6;

struct Eq {
  double gamma, gm1, inv_gm1;
  explicit Eq(const ora_e3_params* p) {
    gamma = p->c_p / p->c_v;
    gm1 = gamma - 1.0;
    inv_gm1 = 1.0 / gm1;
  }
};

// compressible_euler_energy_with_gravity_3d.jl:568-578
template <typename R>
inline void cons2prim(const Eq& eq, const R* u, R& rho, R& v1,
                      R& v2, R& v3, R& p) {
  rho = u[0];
  v1 = u[1] / rho;
  v2 = u[2] / rho;
  v3 = u[3] / rho;
  p = eq.gm1 * (u[4] - 0.5 * (u[1] * v1 + u[2] * v2 + u[3] * v3));
}

// compressible_euler_energy_with_gravity_3d.jl:77-91.  The reference reads
// `rho_e = last(u)`, i.e. the sixth state entry (the geopotential phi), for the
// energy flux; that behaviour is kept as is.
template <typename R>
void flux(const Eq& eq, const R* u, const R* n, R* f) {
  R rho, v1, v2, v3, p;
  cons2prim(eq, u, rho, v1, v2, v3, p);
  const R rho_e = u[5];
  const R vn = v1 * n[0] + v2 * n[1] + v3 * n[2];
  const R rvn = rho * vn;
  f[0] = rvn;
  f[1] = rvn * v1 + p * n[0];
  f[2] = rvn * v2 + p * n[1];
  f[3] = rvn * v3 + p * n[2];
  f[4] = (rho_e + p) * vn;
  f[5] = 0.0;
}

// :382-415
template <typename R>
void flux_ranocha(const Eq& eq, const R* ul, const R* ur,
                  const R* n, R* f) {
  R rl, v1l, v2l, v3l, pl, rr, v1r, v2r, v3r, pr;
  cons2prim(eq, ul, rl, v1l, v2l, v3l, pl);
  cons2prim(eq, ur, rr, v1r, v2r, v3r, pr);
  const R vnl = v1l * n[0] + v2l * n[1] + v3l * n[2];
  const R vnr = v1r * n[0] + v2r * n[1] + v3r * n[2];
  const R rho_mean = ora_ln_mean_t(rl, rr);
  const R inv_rho_p_mean = pl * pr * ora_inv_ln_mean_t(rl * pr, rr * pl);
  const R v1a = 0.5 * (v1l + v1r), v2a = 0.5 * (v2l + v2r),
               v3a = 0.5 * (v3l + v3r);
  const R pa = 0.5 * (pl + pr);
  const R vsq = 0.5 * (v1l * v1r + v2l * v2r + v3l * v3r);
  const R f1 = rho_mean * 0.5 * (vnl + vnr);
  f[0] = f1;
  f[1] = f1 * v1a + pa * n[0];
  f[2] = f1 * v2a + pa * n[1];
  f[3] = f1 * v3a + pa * n[2];
  f[4] = f1 * (vsq + inv_rho_p_mean * eq.inv_gm1) + 0.5 * (pl * vnr + pr * vnl);
  f[5] = 0.0;
}

// :289-318
template <typename R>
void flux_kennedy_gruber(const Eq& eq, const R* ul, const R* ur,
                         const R* n, R* f) {
  R rl, v1l, v2l, v3l, pl, rr, v1r, v2r, v3r, pr;
  cons2prim(eq, ul, rl, v1l, v2l, v3l, pl);
  cons2prim(eq, ur, rr, v1r, v2r, v3r, pr);
  const R ra = 0.5 * (rl + rr);
  const R v1a = 0.5 * (v1l + v1r), v2a = 0.5 * (v2l + v2r),
               v3a = 0.5 * (v3l + v3r);
  const R pa = 0.5 * (pl + pr);
  const R ea = 0.5 * (ul[4] / rl + ur[4] / rr);
  const R vna = v1a * n[0] + v2a * n[1] + v3a * n[2];
  const R f1 = ra * vna;
  f[0] = f1;
  f[1] = f1 * v1a + pa * n[0];
  f[2] = f1 * v2a + pa * n[1];
  f[3] = f1 * v3a + pa * n[2];
  f[4] = f1 * ea + pa * vna;
  f[5] = 0.0;
}

// :248-276
template <typename R>
void flux_shima(const Eq& eq, const R* ul, const R* ur, const R* n,
                R* f) {
  R rl, v1l, v2l, v3l, pl, rr, v1r, v2r, v3r, pr;
  cons2prim(eq, ul, rl, v1l, v2l, v3l, pl);
  cons2prim(eq, ur, rr, v1r, v2r, v3r, pr);
  const R vnl = v1l * n[0] + v2l * n[1] + v3l * n[2];
  const R vnr = v1r * n[0] + v2r * n[1] + v3r * n[2];
  const R ra = 0.5 * (rl + rr);
  const R v1a = 0.5 * (v1l + v1r), v2a = 0.5 * (v2l + v2r),
               v3a = 0.5 * (v3l + v3r);
  const R vna = 0.5 * (vnl + vnr);
  const R pa = 0.5 * (pl + pr);
  const R vsq = 0.5 * (v1l * v1r + v2l * v2r + v3l * v3r);
  const R f1 = ra * vna;
  f[0] = f1;
  f[1] = f1 * v1a + pa * n[0];
  f[2] = f1 * v2a + pa * n[1];
  f[3] = f1 * v3a + pa * n[2];
  f[4] = f1 * vsq + pa * vna * eq.inv_gm1 + 0.5 * (pl * vnr + pr * vnl);
  f[5] = 0.0;
}

// Trixi flux_central = average of the pointwise fluxes (equations.jl:273-279
// restates the same for the covariant form).
template <typename R>
void flux_central(const Eq& eq, const R* ul, const R* ur,
                  const R* n, R* f) {
  R fl[NV], fr[NV];
  flux(eq, ul, n, fl);
  flux(eq, ur, n, fr);
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]);
}

// :538-557 (max_abs_speed with normal direction, the non-naive variant)
template <typename R>
R max_abs_speed(const Eq& eq, const R* ul, const R* ur,
                     const R* n) {
  R rl, v1l, v2l, v3l, pl, rr, v1r, v2r, v3r, pr;
  cons2prim(eq, ul, rl, v1l, v2l, v3l, pl);
  cons2prim(eq, ur, rr, v1r, v2r, v3r, pr);
  const R vl = v1l * n[0] + v2l * n[1] + v3l * n[2];
  const R vr = v1r * n[0] + v2r * n[1] + v3r * n[2];
  const R cl = sqrt(eq.gamma * pl / rl);
  const R cr = sqrt(eq.gamma * pr / rr);
  const R nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  return fmax(fabs(vl) + cl * nn, fabs(vr) + cr * nn);
}

// FluxLaxFriedrichs = FluxPlusDissipation(flux_central,
// DissipationLocalLaxFriedrichs(max_abs_speed)); equations.jl:256-290 restates
// the same construction for the covariant equations.
template <typename R>
void flux_llf(const Eq& eq, const R* ul, const R* ur, const R* n,
              R* f) {
  flux_central(eq, ul, ur, n, f);
  const R lam = max_abs_speed(eq, ul, ur, n);
  for (int v = 0; v < NV; ++v) f[v] += -0.5 * lam * (ur[v] - ul[v]);
}

// :430-466 (FluxLMARS)
template <typename R>
void flux_lmars(const Eq& eq, R c, const R* ul, const R* ur,
                const R* n, R* f) {
  R rl, v1l, v2l, v3l, pl, rr, v1r, v2r, v3r, pr;
  cons2prim(eq, ul, rl, v1l, v2l, v3l, pl);
  cons2prim(eq, ur, rr, v1r, v2r, v3r, pr);
  const R vl = v1l * n[0] + v2l * n[1] + v3l * n[2];
  const R vr = v1r * n[0] + v2r * n[1] + v3r * n[2];
  const R nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const R rho = 0.5 * (rl + rr);
  const R p = 0.5 * (pl + pr) - 0.5 * c * rho * (vr - vl) / nn;
  const R v = 0.5 * (vl + vr) - 1.0 / (2.0 * c * rho) * (pr - pl) * nn;
  const R* us = v >= 0 ? ul : ur;
  const R ps = v >= 0 ? pl : pr;
  f[0] = v * us[0];
  f[1] = v * us[1] + p * n[0];
  f[2] = v * us[2] + p * n[1];
  f[3] = v * us[3] + p * n[2];
  f[4] = v * us[4] + ps * v;
  f[5] = 0.0;
}

// :135-159 (waruszewski: ln_mean density), :206-230 (souza: arithmetic mean)
template <typename R>
void noncons_gravity(int which, const R* ul, const R* ur,
                     const R* n, R* f) {
  for (int v = 0; v < NV; ++v) f[v] = 0.0;
  if (which == ORA_NC_ZERO) return;
  const R v1l = ul[1] / ul[0], v2l = ul[2] / ul[0], v3l = ul[3] / ul[0];
  const R v1r = ur[1] / ur[0], v2r = ur[2] / ur[0], v3r = ur[3] / ur[0];
  const R v1a = 0.5 * (v1l + v1r), v2a = 0.5 * (v2l + v2r),
               v3a = 0.5 * (v3l + v3r);
  const R ra = which == ORA_NC_WARUSZEWSKI ? ora_ln_mean_t(ul[0], ur[0])
                                                 : 0.5 * (ul[0] + ur[0]);
  const R jump = ur[5] - ul[5];
  f[1] = n[0] * ra * jump;
  f[2] = n[1] * ra * jump;
  f[3] = n[2] * ra * jump;
  f[4] = n[0] * ra * v1a * jump + n[1] * ra * v2a * jump + n[2] * ra * v3a * jump;
}

template <typename R>
void volume_flux(const Eq& eq, int which, const R* ul, const R* ur,
                 const R* n, R* f) {
  switch (which) {
    case ORA_VOL_RANOCHA: flux_ranocha(eq, ul, ur, n, f); break;
    case ORA_VOL_KENNEDY_GRUBER: flux_kennedy_gruber(eq, ul, ur, n, f); break;
    case ORA_VOL_SHIMA: flux_shima(eq, ul, ur, n, f); break;
    default: flux_central(eq, ul, ur, n, f); break;
  }
}

template <typename R>
void surface_flux(const Eq& eq, const ora_e3_params* p, const R* ul,
                  const R* ur, const R* n, R* f) {
  switch (p->surface_flux) {
    case ORA_SURF_LAX_FRIEDRICHS: flux_llf(eq, ul, ur, n, f); break;
    case ORA_SURF_RANOCHA: flux_ranocha(eq, ul, ur, n, f); break;
    case ORA_SURF_KENNEDY_GRUBER: flux_kennedy_gruber(eq, ul, ur, n, f); break;
    case ORA_SURF_LMARS: flux_lmars(eq, R(p->lmars_c), ul, ur, n, f); break;
    default: {
      // FluxPlusDissipation(f#, DissipationLocalLaxFriedrichs(max_abs_speed)):
      // equations.jl:256-290 builds the same for the covariant equations
      if (p->surface_flux == ORA_SURF_KG_PLUS_LLF) flux_kennedy_gruber(eq, ul, ur, n, f);
      else flux_ranocha(eq, ul, ur, n, f);
      const R lam = max_abs_speed(eq, ul, ur, n);
      for (int v = 0; v < NV; ++v) f[v] += -0.5 * lam * (ur[v] - ul[v]);
      break;
    }
  }
}

struct Basis {
  int n;
  std::vector<double> x, w, Ds;
  explicit Basis(int n_) : n(n_), x(n_), w(n_), Ds(n_ * n_) {
    ora_lgl_nodes_weights(n, x.data(), w.data());
    ora_derivative_split(n, x.data(), w.data(), Ds.data());
  }
};

// Trixi index_to_start_step_3d semantics: element index of surface node (a, b)
inline int code_index(int code, int a, int b, int n) {
  switch (code) {
    case 0: return 0;
    case 1: return n - 1;
    case 2: return a;
    case 3: return n - 1 - a;
    case 4: return b;
    default: return n - 1 - b;
  }
}

// Trixi indices2direction: (begin,_,_) -> 1, (end,_,_) -> 2, (_,begin,_) -> 3 ...
inline int indices2direction(const int* idx) {
  for (int d = 0; d < 3; ++d) {
    if (idx[d] == 0) return 2 * d;
    if (idx[d] == 1) return 2 * d + 1;
  }
  return -1;
}

// Trixi get_normal_direction: Ja^{orientation}, negated on odd (Julia) i.e.
// negative-side directions.
template <typename R>
inline void normal_direction(int dir0, const R* cv, int node, R* nrm) {
  const int orientation = dir0 / 2;
  const R* ja = cv + 9 * node + 3 * orientation;
  const double s = (dir0 % 2 == 0) ? -1.0 : 1.0;
  nrm[0] = s * ja[0];
  nrm[1] = s * ja[1];
  nrm[2] = s * ja[2];
}

}  // namespace

extern "C" {

void ora_e3_flux(const ora_e3_params* p, const double* u, const double* n,
                 double* f) {
  Eq eq(p);
  flux(eq, u, n, f);
}
void ora_e3_volume_flux(const ora_e3_params* p, int which, const double* ul,
                        const double* ur, const double* n, double* f) {
  Eq eq(p);
  volume_flux(eq, which, ul, ur, n, f);
}
void ora_e3_noncons_flux(const ora_e3_params*, int which, const double* ul,
                         const double* ur, const double* n, double* f) {
  noncons_gravity(which, ul, ur, n, f);
}
void ora_e3_surface_flux(const ora_e3_params* p, int which, const double* ul,
                         const double* ur, const double* n, double* f) {
  Eq eq(p);
  ora_e3_params q = *p;
  q.surface_flux = which;
  surface_flux(eq, &q, ul, ur, n, f);
}
double ora_e3_max_abs_speed(const ora_e3_params* p, const double* ul,
                            const double* ur, const double* n) {
  Eq eq(p);
  return max_abs_speed(eq, ul, ur, n);
}

// :581-601
void ora_e3_cons2entropy(const ora_e3_params* p, const double* u, double* w) {
  Eq eq(p);
  const double rho = u[0];
  const double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
  const double vsq = v1 * v1 + v2 * v2 + v3 * v3;
  const double pr = eq.gm1 * (u[4] - 0.5 * rho * vsq);
  const double s = std::log(pr) - eq.gamma * std::log(rho);
  const double rho_p = rho / pr;
  w[0] = (eq.gamma - s) * eq.inv_gm1 - 0.5 * rho_p * vsq;
  w[1] = rho_p * v1;
  w[2] = rho_p * v2;
  w[3] = rho_p * v3;
  w[4] = -rho_p;
  w[5] = 0.0;
}

// :649-673 (entropy_math)
double ora_e3_entropy(const ora_e3_params* p, const double* u) {
  Eq eq(p);
  const double rho = u[0];
  const double pr =
      eq.gm1 * (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / rho);
  const double s = std::log(pr) - eq.gamma * std::log(rho);
  return -s * rho * eq.inv_gm1;
}

}  // extern "C"

// `mag` = true accumulates the absolute value of every contribution instead,
// giving the magnitude sum_terms |term| that bounds the rounding error of the
// tendency (used by the tests as the error scale of cancelling sums).
template <typename R>
static int e3_rhs_impl(const ora_e3_params* p, const ora_connectivity* c,
                       const R* cv, const R* invJ, const R* u,
                       R* du, bool mag) {
  auto M = [mag](R x) { return mag ? fabs(x) : x; };
  const Eq eq(p);
  const int n = p->n, n3 = n * n * n, E = p->n_elements;
  const Basis B(n);
  const double* Ds = B.Ds.data();
  std::vector<R> sfv((size_t)NV * n * n * 6 * E, 0.0);

  // --- volume integral: flux_differencing_kernel! (symmetric part, then the
  // nonconservative part with the factor 1/2), per element.
#pragma omp parallel for schedule(static) if (kParallel<R>)
  for (int e = 0; e < E; ++e) {
    R* due = du + (size_t)NV * n3 * e;
    std::fill(due, due + NV * n3, R(0.0));
    const R* ue = u + (size_t)NV * n3 * e;
    const R* cve = cv + (size_t)9 * n3 * e;
    auto nid = [n](int i, int j, int k) { return i + n * (j + n * k); };
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int a = nid(i, j, k);
          for (int d = 0; d < 3; ++d) {
            const int ia = d == 0 ? i : (d == 1 ? j : k);
            for (int m = ia + 1; m < n; ++m) {
              const int b = d == 0 ? nid(m, j, k) : (d == 1 ? nid(i, m, k) : nid(i, j, m));
              R Ja[3], f[NV];
              for (int q = 0; q < 3; ++q)
                Ja[q] = 0.5 * (cve[9 * a + 3 * d + q] + cve[9 * b + 3 * d + q]);
              volume_flux(eq, p->volume_flux, ue + NV * a, ue + NV * b, Ja, f);
              for (int v = 0; v < NV; ++v) {
                due[NV * a + v] += M(Ds[ia * n + m] * f[v]);
                due[NV * b + v] += M(Ds[m * n + ia] * f[v]);
              }
            }
          }
        }
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int a = nid(i, j, k);
          R acc[NV] = {0, 0, 0, 0, 0, 0};
          for (int d = 0; d < 3; ++d) {
            const int ia = d == 0 ? i : (d == 1 ? j : k);
            for (int m = 0; m < n; ++m) {
              const int b = d == 0 ? nid(m, j, k) : (d == 1 ? nid(i, m, k) : nid(i, j, m));
              R Ja[3], g[NV];
              for (int q = 0; q < 3; ++q)
                Ja[q] = 0.5 * (cve[9 * a + 3 * d + q] + cve[9 * b + 3 * d + q]);
              noncons_gravity(p->volume_noncons, ue + NV * a, ue + NV * b, Ja, g);
              for (int v = 0; v < NV; ++v) acc[v] += M(Ds[ia * n + m] * g[v]);
            }
          }
          for (int v = 0; v < NV; ++v) due[NV * a + v] += 0.5 * acc[v];
        }
  }

  // --- prolong2interfaces! + calc_interface_flux!: the primary element's
  // outward normal; primary gets f + g(ll,rr)/2, secondary -(f + g(rr,ll)/2).
#pragma omp parallel for schedule(static) if (kParallel<R>)
  for (int I = 0; I < c->n_interfaces; ++I) {
    const int ep = c->neighbor_ids[2 * I], es = c->neighbor_ids[2 * I + 1];
    const int* ip = c->node_indices + 6 * I;
    const int* is = c->node_indices + 6 * I + 3;
    const int dp = indices2direction(ip), ds = indices2direction(is);
    int sp[2], ss[2], np_ = 0, ns_ = 0;  // surface codes (dropping begin/end)
    for (int d = 0; d < 3; ++d) {
      if (ip[d] >= 2) sp[np_++] = ip[d];
      if (is[d] >= 2) ss[ns_++] = is[d];
    }
    for (int bb = 0; bb < n; ++bb)
      for (int aa = 0; aa < n; ++aa) {
        const int pi = code_index(ip[0], aa, bb, n), pj = code_index(ip[1], aa, bb, n),
                  pk = code_index(ip[2], aa, bb, n);
        const int si = code_index(is[0], aa, bb, n), sj = code_index(is[1], aa, bb, n),
                  sk = code_index(is[2], aa, bb, n);
        const int pnode = pi + n * (pj + n * pk), snode = si + n * (sj + n * sk);
        const R* ul = u + (size_t)NV * (pnode + (size_t)n3 * ep);
        const R* ur = u + (size_t)NV * (snode + (size_t)n3 * es);
        R nrm[3];
        normal_direction(dp, cv, pnode + n3 * ep, nrm);
        R f[NV], gp[NV], gs[NV];
        surface_flux(eq, p, ul, ur, nrm, f);
        noncons_gravity(p->surface_noncons, ul, ur, nrm, gp);
        noncons_gravity(p->surface_noncons, ur, ul, nrm, gs);
        // secondary surface-node index
        const int sa = code_index(ss[0], aa, bb, n), sb = code_index(ss[1], aa, bb, n);
        (void)sp;
        R* fp = sfv.data() + (size_t)NV * (aa + n * (bb + n * (dp + 6 * (size_t)ep)));
        R* fs = sfv.data() + (size_t)NV * (sa + n * (sb + n * (ds + 6 * (size_t)es)));
        for (int v = 0; v < NV; ++v) {
          fp[v] = mag ? fabs(f[v]) + 0.5 * fabs(gp[v]) : f[v] + 0.5 * gp[v];
          fs[v] = mag ? fabs(f[v]) + 0.5 * fabs(gs[v]) : -(f[v] + 0.5 * gs[v]);
        }
      }
  }

  // --- prolong2boundaries! + calc_boundary_flux! (slip wall:
  // compressible_euler_energy_with_gravity_3d.jl:93-116)
#pragma omp parallel for schedule(static) if (kParallel<R>)
  for (int Bi = 0; Bi < c->n_boundaries; ++Bi) {
    const int e = c->bnd_element[Bi];
    const int* ib = c->bnd_node_indices + 3 * Bi;
    const int dir = indices2direction(ib);
    for (int bb = 0; bb < n; ++bb)
      for (int aa = 0; aa < n; ++aa) {
        const int ii = code_index(ib[0], aa, bb, n), jj = code_index(ib[1], aa, bb, n),
                  kk = code_index(ib[2], aa, bb, n);
        const int node = ii + n * (jj + n * kk);
        const R* ui = u + (size_t)NV * (node + (size_t)n3 * e);
        R nrm[3];
        normal_direction(dir, cv, node + n3 * e, nrm);
        const R nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        const R nh[3] = {nrm[0] / nn, nrm[1] / nn, nrm[2] / nn};
        const R un = nh[0] * ui[1] + nh[1] * ui[2] + nh[2] * ui[3];
        const R ub[NV] = {ui[0], ui[1] - 2 * un * nh[0], ui[2] - 2 * un * nh[1],
                               ui[3] - 2 * un * nh[2], ui[4], ui[5]};
        R f[NV], g[NV];
        surface_flux(eq, p, ui, ub, nrm, f);
        noncons_gravity(p->surface_noncons, ui, ub, nrm, g);
        R* fo = sfv.data() + (size_t)NV * (aa + n * (bb + n * (dir + 6 * (size_t)e)));
        for (int v = 0; v < NV; ++v)
          fo[v] = mag ? fabs(f[v]) + 0.5 * fabs(g[v]) : f[v] + 0.5 * g[v];
      }
  }

  // --- calc_surface_integral! (outward normals: every face adds f / w_end),
  // then apply_jacobian! (du *= -1/J).
  const double f1 = 1.0 / B.w[0], f2 = 1.0 / B.w[n - 1];
#pragma omp parallel for schedule(static) if (kParallel<R>)
  for (int e = 0; e < E; ++e) {
    R* due = du + (size_t)NV * n3 * e;
    const R* s = sfv.data() + (size_t)NV * n * n * 6 * e;
    auto S = [&](int a, int b, int dir, int v) { return s[NV * (a + n * (b + n * dir)) + v]; };
    for (int m = 0; m < n; ++m)
      for (int l = 0; l < n; ++l)
        for (int v = 0; v < NV; ++v) {
          due[NV * (0 + n * (l + n * m)) + v] += S(l, m, 0, v) * f1;
          due[NV * ((n - 1) + n * (l + n * m)) + v] += S(l, m, 1, v) * f2;
          due[NV * (l + n * (0 + n * m)) + v] += S(l, m, 2, v) * f1;
          due[NV * (l + n * ((n - 1) + n * m)) + v] += S(l, m, 3, v) * f2;
          due[NV * (l + n * (m + n * 0)) + v] += S(l, m, 4, v) * f1;
          due[NV * (l + n * (m + n * (n - 1))) + v] += S(l, m, 5, v) * f2;
        }
    for (int a = 0; a < n3; ++a) {
      const R fac = mag ? fabs(invJ[(size_t)n3 * e + a]) : -invJ[(size_t)n3 * e + a];
      for (int v = 0; v < NV; ++v) due[NV * a + v] *= fac;
    }
    // calc_sources!: source_terms_coriolis, -2 Omega e_z x rho v
    // (src/equations/initial_condition_compressible_euler_3d.jl:245-270)
    if (p->coriolis_rate > 0) {
      const R* ue = u + (size_t)NV * n3 * e;
      for (int a = 0; a < n3; ++a) {
        due[NV * a + 1] += M(2 * p->coriolis_rate * ue[NV * a + 2]);
        due[NV * a + 2] += M(-2 * p->coriolis_rate * ue[NV * a + 1]);
      }
    }
  }
  return 0;
}

extern "C" {

// operation counts of one rhs! evaluation (the same code as ora_e3_rhs run
// with the counting scalar; du receives the same values); counts[4] = +/-,
// *, /, sqrt/log/exp/pow
int ora_e3_count_flops(const ora_e3_params* p, const ora_connectivity* c, const double* cv,
                       const double* invJ, const double* u, double* du, long long* counts) {
  const size_t n3 = (size_t)p->n * p->n * p->n, E = p->n_elements;
  std::vector<Counted> cvc(cv, cv + 9 * n3 * E), ijc(invJ, invJ + n3 * E),
      uc(u, u + 6 * n3 * E), duc(6 * n3 * E);
  ora_counts() = OraCounts{};
  const int rc = e3_rhs_impl(p, c, cvc.data(), ijc.data(), uc.data(), duc.data(), false);
  const OraCounts k = ora_counts();
  counts[0] = k.add;
  counts[1] = k.mul;
  counts[2] = k.div;
  counts[3] = k.special;
  for (size_t q = 0; q < duc.size(); ++q) du[q] = duc[q].v;
  return rc;
}

int ora_e3_rhs(const ora_e3_params* p, const ora_connectivity* c, const double* cv,
               const double* invJ, const double* u, double* du) {
  return e3_rhs_impl(p, c, cv, invJ, u, du, false);
}

int ora_e3_rhs_magnitude(const ora_e3_params* p, const ora_connectivity* c,
                         const double* cv, const double* invJ, const double* u,
                         double* du_abs) {
  return e3_rhs_impl(p, c, cv, invJ, u, du_abs, true);
}

// Carpenter & Kennedy (1994), NASA TM-109112, five-stage fourth-order 2N
// scheme ("solution 3"), as used through
// examples/shallow_water/covariant/elixir_barotropic_instability.jl:79.
void ora_ck2n54_coefficients(double* A, double* B, double* c) {
  const double a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                       -2404267990393.0 / 2016746695238.0,
                       -3550918686646.0 / 2091501179385.0,
                       -1275806237668.0 / 842570457699.0};
  const double b[5] = {1432997174477.0 / 9575080441755.0,
                       5161836677717.0 / 13612068292357.0,
                       1720146321549.0 / 2090206949498.0,
                       3134564353537.0 / 4481467310338.0,
                       2277821191437.0 / 14882151754819.0};
  const double cc[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                        2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0,
                        2802321613138.0 / 2924317926251.0};
  for (int s = 0; s < 5; ++s) { A[s] = a[s]; B[s] = b[s]; c[s] = cc[s]; }
}

int ora_e3_ck2n54(const ora_e3_params* p, const ora_connectivity* c,
                  const double* cv, const double* invJ, double* u, double dt,
                  int n_steps) {
  double A[5], Bc[5], cc[5];
  ora_ck2n54_coefficients(A, Bc, cc);
  const size_t N = (size_t)NV * p->n * p->n * p->n * p->n_elements;
  std::vector<double> du(N), k(N, 0.0);
  for (int step = 0; step < n_steps; ++step) {
    for (int s = 0; s < 5; ++s) {
      ora_e3_rhs(p, c, cv, invJ, u, du.data());
#pragma omp parallel for schedule(static)
      for (size_t q = 0; q < N; ++q) {
        k[q] = A[s] * k[q] + dt * du[q];
        u[q] += Bc[s] * k[q];
      }
    }
  }
  return 0;
}

double ora_e3_max_dt(const ora_e3_params* p, const double* cv,
                     const double* invJ, const double* u) {
  const Eq eq(p);
  const int n = p->n, n3 = n * n * n;
  double max_scaled = 5e-324;  // nextfloat(0.0)
  for (int e = 0; e < p->n_elements; ++e) {
    double l1m = 0, l2m = 0, l3m = 0;
    for (int a = 0; a < n3; ++a) {
      const size_t node = (size_t)n3 * e + a;
      const double* un = u + NV * node;
      double rho, v1, v2, v3, pr;
      cons2prim(eq, un, rho, v1, v2, v3, pr);
      const double cs = std::sqrt(eq.gamma * pr / rho);
      const double l1 = std::fabs(v1) + cs, l2 = std::fabs(v2) + cs,
                   l3 = std::fabs(v3) + cs;
      const double* ja = cv + 9 * node;
      const double t1 = std::fabs(ja[0] * l1 + ja[1] * l2 + ja[2] * l3);
      const double t2 = std::fabs(ja[3] * l1 + ja[4] * l2 + ja[5] * l3);
      const double t3 = std::fabs(ja[6] * l1 + ja[7] * l2 + ja[8] * l3);
      const double ij = std::fabs(invJ[node]);
      l1m = std::fmax(l1m, ij * t1);
      l2m = std::fmax(l2m, ij * t2);
      l3m = std::fmax(l3m, ij * t3);
    }
    max_scaled = std::fmax(max_scaled, l1m + l2m + l3m);
  }
  return 2.0 / (n * max_scaled);
}

void ora_e3_integrals(const ora_e3_params* p, const double* invJ,
                      const double* u, const double* du, double* out) {
  const int n = p->n, n3 = n * n * n;
  const Basis B(n);
  for (int q = 0; q < 9; ++q) out[q] = 0.0;
  for (int e = 0; e < p->n_elements; ++e)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const size_t node = (size_t)n3 * e + i + n * (j + n * k);
          const double wJ = B.w[i] * B.w[j] * B.w[k] / invJ[node];
          const double* un = u + NV * node;
          out[0] += wJ * un[0];
          out[1] += wJ * ora_e3_entropy(p, un);
          if (du) {
            double w[NV];
            ora_e3_cons2entropy(p, un, w);
            double s = 0;
            for (int v = 0; v < NV; ++v) s += w[v] * du[NV * node + v];
            out[2] += wJ * s;
          }
          for (int v = 0; v < NV; ++v) out[3 + v] += wJ * un[v];
        }
}

}  // extern "C"
