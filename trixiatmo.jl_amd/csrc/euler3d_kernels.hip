// gfx950 kernels of the 3D DGSEM right-hand side for
// CompressibleEulerEnergyEquationsWithGravity3D on curved hexahedra.
//
// One workgroup owns one element and one lane one volume node.  A single
// launch per RK stage computes, for its element:
//   volume integral  sum_d sum_m Dsplit[i_d, m] f#(u_i, u_m, avg Ja^d)
//                    + 1/2 sum_d sum_m Dsplit[i_d, m] g(u_i, u_m, avg Ja^d)
//   surface integral sum over the element's own faces of
//                    (f*(u_own, u_nbr, n_out) + 1/2 g(u_own, u_nbr, n_out)) / w_end
//   Jacobian         times -1/J
//   and either stores du (rhs!) or applies the 2N low-storage RK update.
// Every face flux is evaluated by each of its two elements.  Both use the
// primary element's outward normal n_p (Trixi's get_normal_direction on the
// primary side); the secondary evaluates f*(u_s, u_p, -n_p) + g(u_s, u_p, -n_p)/2,
// which for these fluxes equals Trixi's -(f*(u_p, u_s, n_p) + g(u_s, u_p, n_p)/2)
// (f*(b, a, -n) = -f*(a, b, n), g linear in n) up to the operand order of a
// few commutative operations.  This removes the surface_flux_values round trip
// through HBM and needs no atomics.  The RK update writes to a second state buffer (ping-pong)
// because neighbours still read face states of the old one.
#include <hip/hip_runtime.h>

#include "euler3d_kernels.hpp"
#include "euler3d_physics.hpp"

namespace atmo {

__constant__ double c_dsplit[9][64];  // [n][i*n+m], n = 2..8
__constant__ double c_weights[9][8];

template <int N>
struct Shape {
  static constexpr int N3 = N * N * N;
  static constexpr int BLOCK = ((N3 + 63) / 64) * 64;
};

template <typename T>
__device__ __forceinline__ void surface_flux(const E3Params& q, int surf, const T* ul, const T* ur,
                                             const Prim<T>& l, const Prim<T>& r, T n0, T n1,
                                             T n2, T f[6]) {
  if (surf == 0) {  // flux_lax_friedrichs
    flux_central(l, r, n0, n1, n2, f);
    const T lam = max_abs_speed(q, l, r, n0, n1, n2);
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] += T(-0.5) * lam * (ur[v] - ul[v]);
    f[5] = T(-0.5) * lam * (ur[5] - ul[5]);
  } else if (surf == 1) {
    flux_ranocha(q, l, r, n0, n1, n2, f);
    f[5] = T(0);
  } else if (surf == 2) {
    flux_kennedy_gruber(l, r, n0, n1, n2, f);
    f[5] = T(0);
  } else if (surf == 3) {
    flux_lmars(q, ul, ur, l, r, n0, n1, n2, f);
  } else {
    // FluxPlusDissipation(f#, DissipationLocalLaxFriedrichs(max_abs_speed))
    if (surf == 4) flux_kennedy_gruber(l, r, n0, n1, n2, f);
    else flux_ranocha(q, l, r, n0, n1, n2, f);
    const T lam = max_abs_speed(q, l, r, n0, n1, n2);
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] += T(-0.5) * lam * (ur[v] - ul[v]);
    f[5] = T(-0.5) * lam * (ur[5] - ul[5]);
  }
}

template <typename T>
__device__ __forceinline__ void surface_noncons(int snc, const Prim<T>& l, const Prim<T>& r, T n0,
                                                T n1, T n2, T g[4]) {
  if (snc == 1) noncons_gravity<1>(l, r, n0, n1, n2, g);
  else noncons_gravity<2>(l, r, n0, n1, n2, g);
}

template <typename T, int VOL>
__device__ __forceinline__ void volume_flux(const E3Params& q, const Prim<T>& l, const Prim<T>& r,
                                            T n0, T n1, T n2, T f[5]) {
  if constexpr (VOL == 0) flux_ranocha(q, l, r, n0, n1, n2, f);
  else if constexpr (VOL == 1) flux_kennedy_gruber(l, r, n0, n1, n2, f);
  else if constexpr (VOL == 2) flux_shima(q, l, r, n0, n1, n2, f);
  else flux_central(l, r, n0, n1, n2, f);
}

// mode 0: du = rhs(u);  mode 1: k = a k + dt rhs(u), u_out = u + b k
template <int N, int VOL, int NC>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_stage_fp64(
    E3Params q, const double* __restrict__ u, double* __restrict__ out,
    double* __restrict__ kreg, const double* __restrict__ cv, const double* __restrict__ invJ,
    const int* __restrict__ nbr, const int* __restrict__ face_info, int surf, int snc, int mode,
    double rk_a, double rk_dt,
    double rk_b, double coriolis) {
  constexpr int N3 = Shape<N>::N3;
  constexpr int NV = 6;
  __shared__ double s_rho[N3], s_v1[N3], s_v2[N3], s_v3[N3], s_p[N3], s_re[N3], s_phi[N3];
  __shared__ double s_ja[3][N3];
  __shared__ double s_ds[N * N];

  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const bool active = t < N3;
  const int ti = t % N, tj = (t / N) % N, tk = t / (N * N);
  const size_t gnode = (size_t)e * N3 + t;

  if (t < N * N) s_ds[t] = c_dsplit[N][t];

  double uo[NV];
  Prim<double> own;
  double ja_own[9];
  if (active) {
#pragma unroll
    for (int v = 0; v < NV; ++v) uo[v] = u[gnode * NV + v];
    own = cons2prim(q, uo);
    s_rho[t] = own.rho; s_v1[t] = own.v1; s_v2[t] = own.v2; s_v3[t] = own.v3;
    s_p[t] = own.p; s_re[t] = own.rho_e; s_phi[t] = own.phi;
#pragma unroll
    for (int c = 0; c < 9; ++c) ja_own[c] = cv[gnode * 9 + c];
  }

  double acc[5] = {0, 0, 0, 0, 0};
  double acc_nc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    __syncthreads();
    if (active) {
      s_ja[0][t] = ja_own[3 * d + 0];
      s_ja[1][t] = ja_own[3 * d + 1];
      s_ja[2][t] = ja_own[3 * d + 2];
    }
    __syncthreads();
    if (active) {
      const int own_d = d == 0 ? ti : (d == 1 ? tj : tk);
      const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
      const int base = t - own_d * stride;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        const int b = base + m * stride;
        const double ds = s_ds[own_d * N + m];
        Prim<double> pr;
        pr.rho = s_rho[b]; pr.v1 = s_v1[b]; pr.v2 = s_v2[b]; pr.v3 = s_v3[b];
        pr.p = s_p[b]; pr.rho_e = s_re[b]; pr.phi = s_phi[b];
        const double n0 = 0.5 * (ja_own[3 * d + 0] + s_ja[0][b]);
        const double n1 = 0.5 * (ja_own[3 * d + 1] + s_ja[1][b]);
        const double n2 = 0.5 * (ja_own[3 * d + 2] + s_ja[2][b]);
        double f[5];
        volume_flux<double, VOL>(q, own, pr, n0, n1, n2, f);
#pragma unroll
        for (int v = 0; v < 5; ++v) acc[v] += ds * f[v];
        if constexpr (NC != 0) {
          double g[4];
          noncons_gravity<NC>(own, pr, n0, n1, n2, g);
#pragma unroll
          for (int v = 0; v < 4; ++v) acc_nc[v] += ds * g[v];
        }
      }
    }
  }
  if (!active) return;

  double du[NV];
  du[0] = acc[0];
#pragma unroll
  for (int v = 1; v < 5; ++v) du[v] = acc[v] + 0.5 * acc_nc[v - 1];
  du[5] = 0.0;

  // surface terms of the faces this node lies on
  const double inv_w = 1.0 / c_weights[N][0];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2;
    const int own_d = d == 0 ? ti : (d == 1 ? tj : tk);
    const bool on_face = (f % 2 == 0) ? (own_d == 0) : (own_d == N - 1);
    if (!on_face) continue;
    const int a = d == 0 ? tj : ti;
    const int bb = d == 2 ? tj : tk;
    const int link = nbr[((size_t)e * 6 + f) * (N * N) + a + N * bb];
    // face_info: 0 = primary side or boundary (own outward normal);
    // 8 | dp = secondary side of an interface whose primary face direction is dp
    const int info = face_info[(size_t)e * 6 + f];
    double n0, n1, n2;
    if (info == 0) {
      const double sgn = (f % 2 == 0) ? -1.0 : 1.0;
      n0 = sgn * ja_own[3 * d + 0];
      n1 = sgn * ja_own[3 * d + 1];
      n2 = sgn * ja_own[3 * d + 2];
    } else {
      const int dp = info & 7;
      const double sgn = (dp % 2 == 0) ? 1.0 : -1.0;  // minus the primary's outward normal
      const double* jp = cv + (size_t)link * 9 + 3 * (dp / 2);
      n0 = sgn * jp[0];
      n1 = sgn * jp[1];
      n2 = sgn * jp[2];
    }
    double ur[NV];
    if (link >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) ur[v] = u[(size_t)link * NV + v];
    } else {
      // boundary_condition_slip_wall (:93-116): mirror the normal momentum
      const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      const double h0 = n0 / nn, h1 = n1 / nn, h2 = n2 / nn;
      const double un = h0 * uo[1] + h1 * uo[2] + h2 * uo[3];
      ur[0] = uo[0];
      ur[1] = uo[1] - 2.0 * un * h0;
      ur[2] = uo[2] - 2.0 * un * h1;
      ur[3] = uo[3] - 2.0 * un * h2;
      ur[4] = uo[4];
      ur[5] = uo[5];
    }
    const Prim<double> pr = cons2prim(q, ur);
    double fs[6];
    surface_flux<double>(q, surf, uo, ur, own, pr, n0, n1, n2, fs);
    if (snc != 0) {
      double g[4];
      surface_noncons<double>(snc, own, pr, n0, n1, n2, g);
#pragma unroll
      for (int v = 0; v < 4; ++v) fs[v + 1] += 0.5 * g[v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) du[v] += fs[v] * inv_w;
  }

  const double fac = -invJ[gnode];
#pragma unroll
  for (int v = 0; v < NV; ++v) du[v] *= fac;
  // calc_sources!: source_terms_coriolis (-2 Omega e_z x rho v), after the Jacobian
  du[1] += coriolis * uo[2];
  du[2] += -coriolis * uo[1];
  if (mode == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[gnode * NV + v] = du[v];
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double kn = rk_a * kreg[gnode * NV + v] + rk_dt * du[v];
      kreg[gnode * NV + v] = kn;
      out[gnode * NV + v] = uo[v] + rk_b * kn;
    }
  }
}

// ---------------------------------------------------------------- reductions
// per element: max over nodes of |Ja^d . lambda| / J summed over d
template <int N>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_max_speed(E3Params q, const double* __restrict__ u,
                                                               const double* __restrict__ cv,
                                                               const double* __restrict__ invJ,
                                                               double* __restrict__ partial) {
  constexpr int N3 = Shape<N>::N3;
  constexpr int B = Shape<N>::BLOCK;
  __shared__ double s[3][B];
  const int e = blockIdx.x, t = threadIdx.x;
  double l1 = 0, l2 = 0, l3 = 0;
  if (t < N3) {
    const size_t g = (size_t)e * N3 + t;
    double uu[6];
    for (int v = 0; v < 6; ++v) uu[v] = u[g * 6 + v];
    const Prim<double> w = cons2prim(q, uu);
    const double c = sqrt(q.gamma * w.p / w.rho);
    const double a1 = fabs(w.v1) + c, a2 = fabs(w.v2) + c, a3 = fabs(w.v3) + c;
    const double* ja = cv + g * 9;
    const double ij = fabs(invJ[g]);
    l1 = ij * fabs(ja[0] * a1 + ja[1] * a2 + ja[2] * a3);
    l2 = ij * fabs(ja[3] * a1 + ja[4] * a2 + ja[5] * a3);
    l3 = ij * fabs(ja[6] * a1 + ja[7] * a2 + ja[8] * a3);
  }
  s[0][t] = l1; s[1][t] = l2; s[2][t] = l3;
  __syncthreads();
  for (int w = B / 2; w > 0; w >>= 1) {
    if (t < w) {
      s[0][t] = fmax(s[0][t], s[0][t + w]);
      s[1][t] = fmax(s[1][t], s[1][t + w]);
      s[2][t] = fmax(s[2][t], s[2][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) partial[e] = s[0][0] + s[1][0] + s[2][0];
}

// per element: integrals of mass, entropy, w(u).du and every conserved variable
template <int N>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_integrals(E3Params q, const double* __restrict__ u,
                                                               const double* __restrict__ du,
                                                               const double* __restrict__ invJ,
                                                               double* __restrict__ partial) {
  constexpr int N3 = Shape<N>::N3;
  constexpr int B = Shape<N>::BLOCK;
  constexpr int K = 9;
  __shared__ double s[K][B];
  const int e = blockIdx.x, t = threadIdx.x;
  double val[K] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (t < N3) {
    const size_t g = (size_t)e * N3 + t;
    const int i = t % N, j = (t / N) % N, k = t / (N * N);
    const double wJ = c_weights[N][i] * c_weights[N][j] * c_weights[N][k] / invJ[g];
    double uu[6];
    for (int v = 0; v < 6; ++v) uu[v] = u[g * 6 + v];
    const double rho = uu[0];
    const double v1 = uu[1] / rho, v2 = uu[2] / rho, v3 = uu[3] / rho;
    const double vsq = v1 * v1 + v2 * v2 + v3 * v3;
    const double p = q.gm1 * (uu[4] - 0.5 * rho * vsq);
    const double sth = log(p) - q.gamma * log(rho);
    val[0] = wJ * rho;
    val[1] = wJ * (-sth * rho * q.inv_gm1);
    if (du) {
      const double rp = rho / p;
      const double w0 = (q.gamma - sth) * q.inv_gm1 - 0.5 * rp * vsq;
      const double* d = du + g * 6;
      val[2] = wJ * (w0 * d[0] + rp * v1 * d[1] + rp * v2 * d[2] + rp * v3 * d[3] - rp * d[4]);
    }
    for (int v = 0; v < 6; ++v) val[3 + v] = wJ * uu[v];
  }
  for (int c = 0; c < K; ++c) s[c][t] = val[c];
  __syncthreads();
  for (int w = B / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int c = 0; c < K; ++c) s[c][t] += s[c][t + w];
    __syncthreads();
  }
  if (t < K) partial[(size_t)e * K + t] = s[t][0];
}

// single-block reduction of partial[n_items * width] (op 0: sum, 1: max)
__global__ __launch_bounds__(256) void reduce_columns(const double* __restrict__ partial,
                                                      int n_items, int width, int op,
                                                      double* __restrict__ result) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  for (int c = 0; c < width; ++c) {
    double acc = op == 0 ? 0.0 : -1e300;
    for (int i = t; i < n_items; i += 256) {
      const double x = partial[(size_t)i * width + c];
      acc = op == 0 ? acc + x : fmax(acc, x);
    }
    s[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) s[t] = op == 0 ? s[t] + s[t + w] : fmax(s[t], s[t + w]);
      __syncthreads();
    }
    if (t == 0) result[c] = s[0];
    __syncthreads();
  }
}

// ------------------------------------------------------------ host launchers
hipError_t e3_upload_basis_tables(const double* dsplit_all, const double* weights_all) {
  hipError_t err = hipMemcpyToSymbol(HIP_SYMBOL(c_dsplit), dsplit_all, sizeof(double) * 9 * 64);
  if (err != hipSuccess) return err;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_weights), weights_all, sizeof(double) * 9 * 8);
}

template <int N, int VOL, int NC>
static void launch_stage_t(const E3Launch& L, hipStream_t s) {
  hipLaunchKernelGGL((e3_stage_fp64<N, VOL, NC>), dim3(L.n_elements), dim3(Shape<N>::BLOCK), 0, s,
                     L.q, L.u, L.out, L.k, L.cv, L.invJ, L.nbr, L.face_info, L.surface_flux,
                     L.surface_noncons, L.mode, L.rk_a, L.rk_dt, L.rk_b, L.coriolis);
}

template <int N, int VOL>
static bool dispatch_nc(const E3Launch& L, hipStream_t s) {
  switch (L.volume_noncons) {
    case 0: launch_stage_t<N, VOL, 0>(L, s); return true;
    case 1: launch_stage_t<N, VOL, 1>(L, s); return true;
    case 2: launch_stage_t<N, VOL, 2>(L, s); return true;
  }
  return false;
}

template <int N>
static bool dispatch_vol(const E3Launch& L, hipStream_t s) {
  switch (L.volume_flux) {
    case 0: return dispatch_nc<N, 0>(L, s);
    case 1: return dispatch_nc<N, 1>(L, s);
    case 2: return dispatch_nc<N, 2>(L, s);
  }
  return false;
}

bool e3_launch_stage_fp64(const E3Launch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return dispatch_vol<2>(L, s);
    case 3: return dispatch_vol<3>(L, s);
    case 4: return dispatch_vol<4>(L, s);
    case 5: return dispatch_vol<5>(L, s);
    case 6: return dispatch_vol<6>(L, s);
    case 7: return dispatch_vol<7>(L, s);
    case 8: return dispatch_vol<8>(L, s);
  }
  return false;
}

template <int N>
static void launch_red(const E3Launch& L, int which, const double* du, double* partial,
                       hipStream_t s) {
  if (which == 0)
    hipLaunchKernelGGL((e3_max_speed<N>), dim3(L.n_elements), dim3(Shape<N>::BLOCK), 0, s, L.q,
                       L.u, L.cv, L.invJ, partial);
  else
    hipLaunchKernelGGL((e3_integrals<N>), dim3(L.n_elements), dim3(Shape<N>::BLOCK), 0, s, L.q,
                       L.u, du, L.invJ, partial);
}

bool e3_launch_reduction(const E3Launch& L, int which, const double* du, double* partial,
                         double* result, hipStream_t s) {
  switch (L.n) {
    case 2: launch_red<2>(L, which, du, partial, s); break;
    case 3: launch_red<3>(L, which, du, partial, s); break;
    case 4: launch_red<4>(L, which, du, partial, s); break;
    case 5: launch_red<5>(L, which, du, partial, s); break;
    case 6: launch_red<6>(L, which, du, partial, s); break;
    case 7: launch_red<7>(L, which, du, partial, s); break;
    case 8: launch_red<8>(L, which, du, partial, s); break;
    default: return false;
  }
  const int width = which == 0 ? 1 : 9;
  hipLaunchKernelGGL(reduce_columns, dim3(1), dim3(256), 0, s, partial, L.n_elements, width,
                     which == 0 ? 1 : 0, result);
  return true;
}

}  // namespace atmo
