// gfx950 kernels of the covariant shallow-water right-hand side
// (SplitCovariantShallowWaterEquations2D on the cubed sphere) and its RK stage.
//
// Layout: u [3, n, n, E] (h, h v^1, h v^2), aux [26, n, n, E] as Trixi's
// aux_node_vars (src/equations/equations.jl:178-250), node coordinates
// [3, n, n, E].  The stage kernel reads a compact structure-of-arrays copy of
// the 14 auxiliary rows it needs plus z / r (Trixi's node-major aux rows span
// 208 bytes, so reading a few of them fetches the whole row), and one 2 x 2
// basis map per face node instead of the 12 basis entries of both sides.
// A workgroup of 64, 128 or 256 lanes holds EPB = lanes / n^2
// elements (one lane per node); the per-node fields the two-point fluxes need
// (state, velocity v = hv / h and covariant velocity G v formed once per node,
// J, G, G^-1, h_s) are staged in LDS.  Volume terms are gathered per node:
// sum_m Dsplit[own, m] (f#(own, m) + g(own, m) / 2) along both directions (the
// symmetric f# is evaluated for each ordering, the nonconservative g is
// non-symmetric anyway).  Face nodes are
// then mapped to lanes: each element evaluates the interface flux of its own
// faces in its own contravariant basis with its own auxiliary variables,
// exactly as the reference does for each side
// (src/solvers/dgsem_p4est/dg_2d_manifold_in_3d_covariant.jl:301-432), lifts it
// with 1/w_end, and every node adds its face terms, applies -1/J and the
// geometric + Coriolis source, and stores du or applies the 2N RK update
// (per-element time step: ensemble members step with their own dt).
#include <hip/hip_runtime.h>

#include "swe_kernels.hpp"
#include "speed_max.hpp"

namespace atmo {

namespace {

constexpr int NV = 3, NA = 26;

struct AuxV {
  const double* a;  // 26 entries with stride `s`
  int s;
  __device__ __forceinline__ double at(int k) const { return a[k * s]; }
  __device__ __forceinline__ double A(int r, int c) const { return at(r + 3 * c); }
  __device__ __forceinline__ double B(int r, int c) const { return at(6 + r + 2 * c); }
  __device__ __forceinline__ double J() const { return at(12); }
  __device__ __forceinline__ double Gcov(int i, int j) const {
    return (i == 0 && j == 0) ? at(13) : (i == 1 && j == 1) ? at(15) : at(14);
  }
  __device__ __forceinline__ double Gcon(int i, int j) const {
    return (i == 0 && j == 0) ? at(16) : (i == 1 && j == 1) ? at(18) : at(17);
  }
  __device__ __forceinline__ double hs() const { return at(19); }
  __device__ __forceinline__ double Gam(int k, int i, int j) const {
    const int o = 20 + 3 * k;
    return (i == 0 && j == 0) ? at(o) : (i == 1 && j == 1) ? at(o + 2) : at(o + 1);
  }
};

// covariant_shallow_water_split.jl:66-80
__device__ __forceinline__ void flux_split(const double* u, const AuxV& x, int o, double* f) {
  const double mass = x.J() * u[1 + o];
  f[0] = mass;
  f[1] = mass * (u[1] / u[0]);
  f[2] = mass * (u[2] / u[0]);
}

// :100-122 flux_ec
__device__ __forceinline__ void flux_ec(const double* ul, const double* ur, const AuxV& xl,
                                        const AuxV& xr, int o, double* f) {
  const double mass = 0.5 * (xl.J() * ul[1 + o] + xr.J() * ur[1 + o]);
  f[0] = mass;
  f[1] = 0.5 * (ul[1] / ul[0] + ur[1] / ur[0]) * mass;
  f[2] = 0.5 * (ul[2] / ul[0] + ur[2] / ur[0]) * mass;
}

__device__ __forceinline__ void flux_central(const double* ul, const double* ur, const AuxV& xl,
                                             const AuxV& xr, int o, double* f) {
  double fl[NV], fr[NV];
  flux_split(ul, xl, o, fl);
  flux_split(ur, xr, o, fr);
#pragma unroll
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]);
}

// :157-187 flux_nonconservative_ec
__device__ __forceinline__ void noncons_ec(double g, const double* ul, const double* ur,
                                           const AuxV& xl, const AuxV& xr, int o, double* f) {
  const double hl = ul[0], hr = ur[0];
  const double vr0 = ur[1] / hr, vr1 = ur[2] / hr;
  const double vc0 = xr.Gcov(0, 0) * vr0 + xr.Gcov(0, 1) * vr1;
  const double vc1 = xr.Gcov(1, 0) * vr0 + xr.Gcov(1, 1) * vr1;
  const double hs_jump = xr.hs() - xl.hs();
  const double geo0 = 0.5 * ul[1 + o] * ((xl.Gcon(0, 0) * vc0 + xl.Gcon(0, 1) * vc1) - vr0);
  const double geo1 = 0.5 * ul[1 + o] * ((xl.Gcon(1, 0) * vc0 + xl.Gcon(1, 1) * vc1) - vr1);
  const double pr0 = g * xl.Gcon(0, o) * hl * (hr + hs_jump);
  const double pr1 = g * xl.Gcon(1, o) * hl * (hr + hs_jump);
  f[0] = 0.0;
  f[1] = xl.J() * (geo0 + pr0);
  f[2] = xl.J() * (geo1 + pr1);
}

__device__ __forceinline__ double max_abs_speed(double g, const double* ul, const double* ur,
                                                const AuxV& xl, const AuxV& xr, int o) {
  return fmax(fabs(ul[1 + o] / ul[0]) + sqrt(xl.Gcon(o, o) * ul[0] * g),
              fabs(ur[1 + o] / ur[0]) + sqrt(xr.Gcon(o, o) * ur[0] * g));
}

// CovariantShallowWaterEquations2D, the standard form (covariant_shallow_water.jl:166-187):
// contravariant flux in orientation o with the pressure term G^{io} g h^2 / 2
__device__ __forceinline__ void flux_standard(double g, double h, double hv1, double hv2,
                                              double J, double Gi_0o, double Gi_1o, double hvo,
                                              double* f) {
  const double vcon = hvo / h;
  const double grav = 0.5 * g * h * h;
  f[0] = J * hvo;
  f[1] = J * (hv1 * vcon + Gi_0o * grav);
  f[2] = J * (hv2 * vcon + Gi_1o * grav);
}

__device__ __forceinline__ void flux_standard(double g, const double* u, const AuxV& x, int o,
                                              double* f) {
  flux_standard(g, u[0], u[1], u[2], x.J(), x.Gcon(0, o), x.Gcon(1, o), u[1 + o], f);
}

// source_terms_geometric_coriolis of the standard form (:215-238)
__device__ __forceinline__ void source_standard(double g, double omega, const double* u,
                                                double sin_lat, const AuxV& a, double* s) {
  const double h = u[0], hv[2] = {u[1], u[2]};
  const double v[2] = {u[1] / h, u[2] / h};
  const double grav = 0.5 * g * h * h;
  double tau[2][2];  // h v^i v^j + G^{ij} g h^2 / 2
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) tau[i][j] = hv[i] * v[j] + grav * a.Gcon(i, j);
  const double f = 2 * omega * sin_lat;
  double src[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    double sg = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) sg += a.Gam(k, i, j) * tau[i][j];
    src[k] = sg + f * a.J() * (a.Gcon(k, 1) * hv[0] - a.Gcon(k, 0) * hv[1]);
  }
  s[0] = 0.0;
  s[1] = -src[0];
  s[2] = -src[1];
}

// surface flux (u_ll, u_rr, aux, aux) + 1/2 the nonconservative surface term
__device__ __forceinline__ void surface_terms(const SweLaunch& L, const double* ul,
                                              const double* ur, const AuxV& x, int o,
                                              double* f) {
  if (L.surface_flux == 3) {
    // flux_lax_friedrichs of the standard form: flux_central of the standard
    // flux plus the local Lax-Friedrichs dissipation; no nonconservative term
    double fl[NV], fr[NV];
    flux_standard(L.gravity, ul, x, o, fl);
    flux_standard(L.gravity, ur, x, o, fr);
    const double lam = max_abs_speed(L.gravity, ul, ur, x, x, o);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]) + -0.5 * x.J() * lam * (ur[v] - ul[v]);
    return;
  }
  if (L.surface_flux == 1) flux_central(ul, ur, x, x, o, f);
  else flux_ec(ul, ur, x, x, o, f);
  if (L.surface_flux != 2) {  // DissipationLocalLaxFriedrichs, equations.jl:283-290
    const double lam = max_abs_speed(L.gravity, ul, ur, x, x, o);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] += -0.5 * x.J() * lam * (ur[v] - ul[v]);
  }
  double g[NV];
  if (L.surface_noncons == 0) {  // flux_nonconservative_surface_simplified, :229-250
    g[0] = 0.0;
    g[1] = x.J() * (L.gravity * x.Gcon(0, o) * ul[0] * ur[0]);
    g[2] = x.J() * (L.gravity * x.Gcon(1, o) * ul[0] * ur[0]);
  } else {
    noncons_ec(L.gravity, ul, ur, x, x, o, g);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) f[v] += 0.5 * g[v];
}

// covariant_shallow_water_split.jl:273-302 (sin_lat = z / r of the node)
__device__ __forceinline__ void source_split(double omega, const double* u, double sin_lat,
                                             const AuxV& a, double* s) {
  const double hv0 = u[1], hv1 = u[2];
  const double v0 = u[1] / u[0], v1 = u[2] / u[0];
  const double hvv[2][2] = {{hv0 * v0, hv0 * v1}, {hv1 * v0, hv1 * v1}};
  double hcv[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) hcv[i][j] = a.Gcov(i, 0) * hvv[0][j] + a.Gcov(i, 1) * hvv[1][j];
  const double f = 2 * omega * sin_lat;
  double t1[2], t2[2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
    t1[k] = a.Gam(k, 0, 0) * hvv[0][0] + a.Gam(k, 1, 0) * hvv[1][0] + a.Gam(k, 0, 1) * hvv[0][1] +
            a.Gam(k, 1, 1) * hvv[1][1];
#pragma unroll
  for (int i = 0; i < 2; ++i)
    t2[i] = a.Gam(0, i, 0) * hcv[0][0] + a.Gam(0, i, 1) * hcv[0][1] + a.Gam(1, i, 0) * hcv[1][0] +
            a.Gam(1, i, 1) * hcv[1][1];
  const double sg0 = 0.5 * (t1[0] - (a.Gcon(0, 0) * t2[0] + a.Gcon(0, 1) * t2[1]));
  const double sg1 = 0.5 * (t1[1] - (a.Gcon(1, 0) * t2[0] + a.Gcon(1, 1) * t2[1]));
  s[0] = 0.0;
  s[1] = -(sg0 + f * a.J() * (a.Gcon(0, 1) * hv0 - a.Gcon(0, 0) * hv1));
  s[2] = -(sg1 + f * a.J() * (a.Gcon(1, 1) * hv0 - a.Gcon(1, 0) * hv1));
}

// Workgroup shape: WAVES waves of 64 lanes, one lane per node, EPB = WAVES *
// 64 / n^2 elements (n^2 <= 64); with PERSIST the grid is capped and every
// workgroup loops over element groups with a grid stride.
template <int N, int WAVES>
struct SweShape {
  static constexpr int N2 = N * N;
  static constexpr int BLOCK = 64 * WAVES;
  static constexpr int EPB = BLOCK / N2;
  static constexpr int NF = 4 * N;  // face nodes per element
};

// per-node fields staged in LDS: h, hv^1, hv^2, v^1, v^2 (v = hv / h), J,
// G_11, G_12, G_22, G^11, G^12, G^22, h_s (13 doubles per node: with the
// face buffer 8 KB per wave at polydeg 3, which leaves room for five waves
// per SIMD)
constexpr int kSweFields = 13;

// the stage's geometry rows and face maps in their storage type GT (double, or
// float under ATMO_PREC_FP32_GEOMETRY); a float operand widens to double in
// every expression it enters
template <typename GT>
__device__ __forceinline__ const GT* geo_rows(const SweLaunch& L) {
  if constexpr (sizeof(GT) == 4) return L.geo32; else return L.geo;
}
template <typename GT>
__device__ __forceinline__ const GT* face_maps(const SweLaunch& L) {
  if constexpr (sizeof(GT) == 4) return L.fxf32; else return L.fxf;
}

}  // namespace

// StepsizeCallback's scaled wave speeds at one node (stepsize_dg2d.jl:47-71):
// |v^i| + sqrt(G^ii h g); the max-speed reduction and the last stage of a
// resident step both call this one function, so their speeds agree bit for bit
__device__ __forceinline__ void swe_node_speeds(const double* u, double gcon11, double gcon22,
                                                double gravity, double l[2]) {
  l[0] = fabs(u[1] / u[0]) + sqrt(gcon11 * u[0] * gravity);
  l[1] = fabs(u[2] / u[0]) + sqrt(gcon22 * u[0] * gravity);
}

// SPEED: the RK stage also writes L.speed_partial (the last stage of a
// resident step); a variant of its own, so the other stages keep their
// register budget
// polydeg 3 fits five waves per SIMD (90-92 VGPRs, 8 KB of LDS per wave, no
// spills); the other degrees keep the compiler's register choice (a five-wave
// budget spills there)
template <int N, int WAVES, bool PERSIST, bool SPEED, typename GT>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(N == 4 ? 5 : 1)))
void swe_stage(SweLaunch L) {
  using S = SweShape<N, WAVES>;
  constexpr int N2 = S::N2, EPB = S::EPB, B = S::BLOCK;
  __shared__ double s_n[EPB][kSweFields][N2];
  __shared__ double s_sf[EPB][NV][S::NF];

  const int t = threadIdx.x;
  const int es = t / N2, q = t - es * N2;
  const int ti = q % N, tj = q / N;
  const int groups = (L.n_elements + EPB - 1) / EPB;
  for (int grp = blockIdx.x; grp < groups; grp += PERSIST ? gridDim.x : groups) {
    const int e0 = grp * EPB;
    const int e = e0 + es;
    const bool active = es < EPB && e < L.n_elements;
    if (PERSIST) __syncthreads();  // the previous group's LDS reads are done
    if (active) {
      const size_t g = (size_t)e * N2 + q;
      const double* un = L.u + g * NV;
      const GT* gg = geo_rows<GT>(L) + g;
      const long long NN = L.n_nodes;
      const double h = un[0], hv1 = un[1], hv2 = un[2];
      const double v1 = hv1 / h, v2 = hv2 / h;
      const double G11 = gg[0], G12 = gg[NN], G22 = gg[2 * NN];
      double J, c11, c12, c22;
      metric_from_gcov(G11, G12, G22, J, c11, c12, c22);
      // FP32 storage: J is the stored area element rounded once; it is the
      // same number on both sides of a face (G is not: it is expressed in each
      // element's own basis), so the two mass fluxes there agree
      if constexpr (sizeof(GT) == 4) J = gg[11 * NN];
      s_n[es][0][q] = h;
      s_n[es][1][q] = hv1;
      s_n[es][2][q] = hv2;
      s_n[es][3][q] = v1;
      s_n[es][4][q] = v2;
      s_n[es][5][q] = J;
      s_n[es][6][q] = G11;
      s_n[es][7][q] = G12;
      s_n[es][8][q] = G22;
      s_n[es][9][q] = c11;
      s_n[es][10][q] = c12;
      s_n[es][11][q] = c22;
      if constexpr (sizeof(GT) == 4) s_n[es][12][q] = L.topography ? L.hs64[g] : 0.0;
      else s_n[es][12][q] = L.topography ? gg[3 * NN] : 0.0;
    }
    __syncthreads();

    double du[NV] = {0, 0, 0};
    if (active && L.volume_flux == 2) {
      // weak form (dg_2d_manifold_in_3d_covariant.jl:100-130): node (i, j)
      // gathers derivative_hat[i, m] F^1(m, j) + derivative_hat[j, m] F^2(i, m),
      // the standard flux of every node on its two lines evaluated in place
#pragma unroll
      for (int m = 0; m < N; ++m) {
        const int p1 = m + N * tj, p2 = ti + N * m;
        double f1[NV], f2[NV];
        flux_standard(L.gravity, s_n[es][0][p1], s_n[es][1][p1], s_n[es][2][p1], s_n[es][5][p1],
                      s_n[es][9][p1], s_n[es][10][p1], s_n[es][1][p1], f1);
        flux_standard(L.gravity, s_n[es][0][p2], s_n[es][1][p2], s_n[es][2][p2], s_n[es][5][p2],
                      s_n[es][10][p2], s_n[es][11][p2], s_n[es][2][p2], f2);
        const double d1 = L.dhat[ti * N + m], d2 = L.dhat[tj * N + m];
#pragma unroll
        for (int v = 0; v < NV; ++v) du[v] += d1 * f1[v] + d2 * f2[v];
      }
    } else if (active) {
      const double hl = s_n[es][0][q], hvl[2] = {s_n[es][1][q], s_n[es][2][q]};
      const double vl[2] = {s_n[es][3][q], s_n[es][4][q]};
      const double Jl = s_n[es][5][q], hsl = s_n[es][12][q];
      const double Gi[3] = {s_n[es][9][q], s_n[es][10][q], s_n[es][11][q]};
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int own = o == 0 ? ti : tj;
        // the partners in rotated order own + 1, ..., own + n - 1 (mod n):
        // every lane has a partner in each step
#pragma unroll 1
        for (int j = 1; j < N; ++j) {
          const int m = own + j >= N ? own + j - N : own + j;
          const int pq = o == 0 ? m + N * tj : ti + N * m;
          const double hr = s_n[es][0][pq], hvr_o = s_n[es][1 + o][pq];
          const double vr[2] = {s_n[es][3][pq], s_n[es][4][pq]};
          // the partner's covariant velocity G v from its staged G and v
          const double vc0 = s_n[es][6][pq] * vr[0] + s_n[es][7][pq] * vr[1];
          const double vc1 = s_n[es][7][pq] * vr[0] + s_n[es][8][pq] * vr[1];
          const double Jr = s_n[es][5][pq];
          double f[NV];
          if (L.volume_flux == 0) {  // flux_ec (:100-122)
            const double mass = 0.5 * (Jl * hvl[o] + Jr * hvr_o);
            f[0] = mass;
            f[1] = 0.5 * (vl[0] + vr[0]) * mass;
            f[2] = 0.5 * (vl[1] + vr[1]) * mass;
          } else {  // flux_central of the split flux (:66-80)
            const double ml = Jl * hvl[o], mr = Jr * hvr_o;
            f[0] = 0.5 * (ml + mr);
            f[1] = 0.5 * (ml * vl[0] + mr * vr[0]);
            f[2] = 0.5 * (ml * vl[1] + mr * vr[1]);
          }
          // flux_nonconservative_ec (:157-187) of (own, partner)
          const double hs_jump = s_n[es][12][pq] - hsl;
          const double geo0 = 0.5 * hvl[o] * ((Gi[0] * vc0 + Gi[1] * vc1) - vr[0]);
          const double geo1 = 0.5 * hvl[o] * ((Gi[1] * vc0 + Gi[2] * vc1) - vr[1]);
          const double Gio0 = o == 0 ? Gi[0] : Gi[1], Gio1 = o == 0 ? Gi[1] : Gi[2];
          const double pr0 = L.gravity * Gio0 * hl * (hr + hs_jump);
          const double pr1 = L.gravity * Gio1 * hl * (hr + hs_jump);
          const double g1 = Jl * (geo0 + pr0), g2 = Jl * (geo1 + pr1);
          const double ds = L.dsplit[own * N + m];
          du[0] += ds * f[0];
          du[1] += ds * f[1] + 0.5 * ds * g1;
          du[2] += ds * f[2] + 0.5 * ds * g2;
        }
      }
    }
    // face nodes of the group's elements, one lane each
    const double inv_w = L.inv_w;
    for (int fid = t; fid < EPB * S::NF; fid += B) {
      const int fe = fid / S::NF, r = fid - fe * S::NF;
      const int f = r / N, k = r - f * N;
      if (e0 + fe >= L.n_elements) continue;
      const int i = f == 0 ? 0 : (f == 1 ? N - 1 : k);
      const int j = f == 2 ? 0 : (f == 3 ? N - 1 : k);
      const int qn = i + N * j;
      const double ul[NV] = {s_n[fe][0][qn], s_n[fe][1][qn], s_n[fe][2][qn]};
      // the own node's J, G, G^-1, h_s (aux rows 12-19) from LDS
      double xa[NA] = {};
      xa[12] = s_n[fe][5][qn];
#pragma unroll
      for (int c = 0; c < 6; ++c) xa[13 + c] = s_n[fe][6 + c][qn];
      xa[19] = s_n[fe][12][qn];
      const AuxV ax{xa, 1};
      const size_t fnode = ((size_t)(e0 + fe) * 4 + f) * N + k;
      const int link = L.nbr[fnode];
      const double* ur = L.u + (size_t)link * NV;
      // neighbour state in this element's contravariant basis (contravariant2global,
      // then global2contravariant: covariant_shallow_water.jl:140-153)
      const GT* M = face_maps<GT>(L) + fnode * 4;
      const double tr[NV] = {ur[0], M[0] * ur[1] + M[1] * ur[2], M[2] * ur[1] + M[3] * ur[2]};
      double fs[NV];
      if constexpr (N == 4) {
        // polydeg 3 keeps everything in registers with the orientation as a
        // run-time value, and runs faster so (measured)
        // the two states selected per lane and one call: the lanes of a wave
        // hold faces of both signs, and a branch per sign runs both calls
        const int o = f / 2;
        const bool neg = f % 2 == 0;  // negative-side face: -(f*(u_nbr, u_own) + g/2)
        double sl[NV], sr[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          sl[v] = neg ? tr[v] : ul[v];
          sr[v] = neg ? ul[v] : tr[v];
        }
        surface_terms(L, sl, sr, ax, o, fs);
#pragma unroll
        for (int v = 0; v < NV; ++v) fs[v] = neg ? -fs[v] : fs[v];
      } else {
        // the orientation as a compile-time constant in each branch: every
        // index into the node arrays (u[1 + o], G^{io}) is then fixed and they
        // stay in registers (with a run-time orientation they went to scratch);
        // the sign's two states selected per lane, so a wave runs at most the
        // two orientations' calls
        const bool neg = f % 2 == 0;  // negative-side faces: -(f*(u_nbr, u_own) + g/2)
        double sl[NV], sr[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          sl[v] = neg ? tr[v] : ul[v];
          sr[v] = neg ? ul[v] : tr[v];
        }
        if (f < 2) surface_terms(L, sl, sr, ax, 0, fs);
        else surface_terms(L, sl, sr, ax, 1, fs);
#pragma unroll
        for (int v = 0; v < NV; ++v) fs[v] = neg ? -fs[v] : fs[v];
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) s_sf[fe][v][r] = fs[v] * inv_w;
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const bool on = f == 0 ? ti == 0 : f == 1 ? ti == N - 1 : f == 2 ? tj == 0 : tj == N - 1;
        if (!on) continue;
        const int k = f < 2 ? tj : ti;
#pragma unroll
        for (int v = 0; v < NV; ++v) du[v] += s_sf[es][v][f * N + k];
      }
      const size_t g = (size_t)e * N2 + q;
      const double fac = -1.0 / s_n[es][5][q];
      const double uo[NV] = {s_n[es][0][q], s_n[es][1][q], s_n[es][2][q]};
#pragma unroll
      for (int v = 0; v < NV; ++v) du[v] *= fac;
      if (L.source_terms) {
        const GT* gg = geo_rows<GT>(L) + g;
        const long long NN = L.n_nodes;
        double xa[NA] = {};
        xa[12] = s_n[es][5][q];
#pragma unroll
        for (int c = 0; c < 6; ++c) xa[13 + c] = s_n[es][6 + c][q];
#pragma unroll
        for (int c = 0; c < 6; ++c) xa[20 + c] = gg[(4 + c) * NN];
        double sv[NV];
        if (L.source_terms == 2)
          source_standard(L.gravity, L.rotation_rate, uo, gg[10 * NN], AuxV{xa, 1}, sv);
        else
          source_split(L.rotation_rate, uo, gg[10 * NN], AuxV{xa, 1}, sv);
#pragma unroll
        for (int v = 0; v < NV; ++v) du[v] += sv[v];
      }
      if (L.mode == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) L.out[g * NV + v] = du[v];
      } else {
        // the member's step size, else the launch's; selected as values (one
        // conditional form with both reads merges them into a read through a
        // pointer to global or argument memory, a flat load)
        const double dt_m = L.dt_member ? L.dt_member[L.member[e]] : 0.0;
        const double dt = L.dt_member ? dt_m : L.rk_dt;
        double un[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const double kn = L.rk_a * L.k[g * NV + v] + dt * du[v];
          L.k[g * NV + v] = kn;
          un[v] = uo[v] + L.rk_b * kn;
          L.out[g * NV + v] = un[v];
        }
        if constexpr (SPEED) {
          // the node's speeds go to its first two LDS fields, which only this
          // lane reads after the face phase and which it has read already
          double l[2];
          swe_node_speeds(un, s_n[es][9][q], s_n[es][11][q], L.gravity, l);
          s_n[es][0][q] = l[0];
          s_n[es][1][q] = l[1];
        }
      }
    }
    if constexpr (SPEED) {
      // per element: the maximum over its nodes of each speed (exact in any order),
      // then the sum over the directions, by the element's first lane
      __syncthreads();
      if (active && q == 0) {
        double m1 = s_n[es][0][0], m2 = s_n[es][1][0];
        for (int r = 1; r < N2; ++r) {
          m1 = speed_max(m1, s_n[es][0][r]);
          m2 = speed_max(m2, s_n[es][1][r]);
        }
        L.speed_partial[e] = m1 + m2;
      }
    }
  }
}

// per-element maximum of the scaled wave speeds (stepsize_dg2d.jl:47-71)
// G^11 and G^22 are recomputed from the stage's G rows in their storage type,
// as the fused last stage does
template <int N, typename GT>
__global__ __launch_bounds__(64) void swe_max_speed(SweLaunch L, double* partial) {
  constexpr int N2 = N * N;
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double s1[64], s2[64];
  double l1 = 0.0, l2 = 0.0;
  const GT* gr = geo_rows<GT>(L);
  for (int q = t; q < N2; q += 64) {
    const size_t g = (size_t)e * N2 + q;
    const double* u = L.u + g * NV;
    double J, c11, c12, c22, l[2];
    metric_from_gcov(gr[g], gr[L.n_nodes + g], gr[2 * L.n_nodes + g], J, c11, c12, c22);
    swe_node_speeds(u, c11, c22, L.gravity, l);
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

// per-member maximum over its elements
__global__ void swe_member_max(const double* partial, const int* member_first, int n_members,
                               double* out) {
  const int m = blockIdx.x, t = threadIdx.x;
  __shared__ double s[256];
  double acc = 5e-324;
  for (int e = member_first[m] + t; e < member_first[m + 1]; e += 256) acc = speed_max(acc, partial[e]);
  s[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) s[t] = speed_max(s[t], s[t + w]);
    __syncthreads();
  }
  if (t == 0) out[m] = s[0];
}

namespace {

template <int N, int WAVES, bool PERSIST>
void launch_variant(const SweLaunch& L, hipStream_t s) {
  using S = SweShape<N, WAVES>;
  const int groups = (L.n_elements + S::EPB - 1) / S::EPB;
  // persistent: a few workgroups per CU (256 CUs), each looping over groups
  const int grid = PERSIST ? (groups < 256 * 8 ? groups : 256 * 8) : groups;
  const bool speed = L.speed_partial && L.mode == 1;
  if (L.geo32) {
    if (speed)
      hipLaunchKernelGGL((swe_stage<N, WAVES, PERSIST, true, float>), dim3(grid), dim3(S::BLOCK), 0,
                         s, L);
    else
      hipLaunchKernelGGL((swe_stage<N, WAVES, PERSIST, false, float>), dim3(grid), dim3(S::BLOCK),
                         0, s, L);
  } else if (speed) {
    hipLaunchKernelGGL((swe_stage<N, WAVES, PERSIST, true, double>), dim3(grid), dim3(S::BLOCK), 0,
                       s, L);
  } else {
    hipLaunchKernelGGL((swe_stage<N, WAVES, PERSIST, false, double>), dim3(grid), dim3(S::BLOCK), 0,
                       s, L);
  }
}

template <int N>
bool launch_n(const SweLaunch& L, int variant, hipStream_t s) {
  // variants: 0..2 = 1, 2, 4 waves per workgroup; +3 = persistent grid
  if constexpr (N * N > 64) {
    return false;
  } else {
    switch (variant) {
      case 0: launch_variant<N, 1, false>(L, s); return true;
      case 1: launch_variant<N, 2, false>(L, s); return true;
      case 2: launch_variant<N, 4, false>(L, s); return true;
      case 3: launch_variant<N, 1, true>(L, s); return true;
      case 4: launch_variant<N, 2, true>(L, s); return true;
      case 5: launch_variant<N, 4, true>(L, s); return true;
    }
    return false;
  }
}

}  // namespace

bool swe_launch_stage(const SweLaunch& L, hipStream_t s) {
  const int variant = L.variant;
  switch (L.n) {
    case 2: return launch_n<2>(L, variant, s);
    case 3: return launch_n<3>(L, variant, s);
    case 4: return launch_n<4>(L, variant, s);
    case 5: return launch_n<5>(L, variant, s);
    case 6: return launch_n<6>(L, variant, s);
    case 7: return launch_n<7>(L, variant, s);
    case 8: return launch_n<8>(L, variant, s);
  }
  return false;
}

bool swe_launch_member_max_speed(const SweLaunch& L, double* partial, const int* member_first,
                                 int n_members, double* out, hipStream_t s) {
#define ATMO_SWE_CASE(NN)                                                                       \
  case NN:                                                                                      \
    if (L.geo32)                                                                                \
      hipLaunchKernelGGL((swe_max_speed<NN, float>), dim3(L.n_elements), dim3(64), 0, s, L,     \
                         partial);                                                              \
    else                                                                                        \
      hipLaunchKernelGGL((swe_max_speed<NN, double>), dim3(L.n_elements), dim3(64), 0, s, L,    \
                         partial);                                                              \
    break;
  switch (L.n) {
    ATMO_SWE_CASE(2)
    ATMO_SWE_CASE(3)
    ATMO_SWE_CASE(4)
    ATMO_SWE_CASE(5)
    ATMO_SWE_CASE(6)
    ATMO_SWE_CASE(7)
    ATMO_SWE_CASE(8)
    default:
      return false;
  }
#undef ATMO_SWE_CASE
  hipLaunchKernelGGL(swe_member_max, dim3(n_members), dim3(256), 0, s, partial, member_first,
                     n_members, out);
  return true;
}

}  // namespace atmo

namespace atmo {

// per element: mass, total energy (entropy) and int w(u).du with the exact
// area element (analysis_covariant.jl:53-105; covariant_shallow_water.jl:129-164)
template <int N>
__global__ __launch_bounds__(64) void swe_integrals_kernel(SweLaunch L, const double* du,
                                                           const double* w, double* partial) {
  constexpr int N2 = N * N;
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double s[3][64];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int q = t; q < N2; q += 64) {
    const size_t g = (size_t)e * N2 + q;
    const double* u = L.u + g * NV;
    const AuxV a{L.aux + g * NA, 1};
    const double h = u[0], v1 = u[1] / h, v2 = u[2] / h;
    const double vc1 = a.Gcov(0, 0) * v1 + a.Gcov(0, 1) * v2;
    const double vc2 = a.Gcov(1, 0) * v1 + a.Gcov(1, 1) * v2;
    const double wJ = w[q % N] * w[q / N] * a.J();
    a0 += wJ * h;
    a1 += wJ * (0.5 * (h * (vc1 * v1 + vc2 * v2) + L.gravity * h * h) + L.gravity * h * a.hs());
    if (du) {
      const double w1 = L.gravity * (h + a.hs()) - 0.5 * (vc1 * v1 + vc2 * v2);
      a2 += wJ * (w1 * du[g * NV] + vc1 * du[g * NV + 1] + vc2 * du[g * NV + 2]);
    }
  }
  s[0][t] = a0;
  s[1][t] = a1;
  s[2][t] = a2;
  __syncthreads();
  for (int k = 32; k > 0; k >>= 1) {
    if (t < k)
      for (int c = 0; c < 3; ++c) s[c][t] += s[c][t + k];
    __syncthreads();
  }
  if (t < 3) partial[(size_t)e * 3 + t] = s[t][0];
}

__global__ void swe_sum_columns(const double* partial, int n_items, double* out) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int i = t; i < n_items; i += 256) acc += partial[(size_t)i * 3 + c];
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

bool swe_launch_integrals(const SweLaunch& L, const double* du, const double* w, double* partial,
                          double* out, hipStream_t s) {
#define ATMO_SWE_CASE(NN)                                                                     \
  case NN:                                                                                    \
    hipLaunchKernelGGL(swe_integrals_kernel<NN>, dim3(L.n_elements), dim3(64), 0, s, L, du, w, \
                       partial);                                                              \
    break;
  switch (L.n) {
    ATMO_SWE_CASE(2)
    ATMO_SWE_CASE(3)
    ATMO_SWE_CASE(4)
    ATMO_SWE_CASE(5)
    ATMO_SWE_CASE(6)
    ATMO_SWE_CASE(7)
    ATMO_SWE_CASE(8)
    default:
      return false;
  }
#undef ATMO_SWE_CASE
  hipLaunchKernelGGL(swe_sum_columns, dim3(1), dim3(256), 0, s, partial, L.n_elements, out);
  return true;
}

}  // namespace atmo
