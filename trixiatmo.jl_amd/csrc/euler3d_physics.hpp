// Pointwise physics of CompressibleEulerEnergyEquationsWithGravity3D
// (src/equations/compressible_euler_energy_with_gravity_3d.jl) as device
// functions.  State order (rho, rho_v1, rho_v2, rho_v3, rho_e, phi); rho_e is
// total energy without the geopotential, p = (gamma-1)(rho_e - rho |v|^2 / 2).
#ifndef ATMO_EULER3D_PHYSICS_HPP
#define ATMO_EULER3D_PHYSICS_HPP

#include <hip/hip_runtime.h>

namespace atmo {

struct E3Params {
  double gamma, gm1, inv_gm1, lmars_c;
};

// Arithmetic helpers.  FP32 uses the native single-instruction reciprocal,
// square root and base-2 logarithm (about 1 ulp).  FP64 division stays IEEE
// in div_; the hot per-node and per-face-node quantities use div_nr, sqrt_ and
// log_ below, which stay within one unit in the last place for the positive
// normal operands the fluxes give them (the FP64 path is compared with the
// oracle to 1e-12 of its term scale).
__device__ __forceinline__ double div_(double a, double b) { return a / b; }
__device__ __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// 1 / x from the hardware reciprocal and one Newton step
__device__ __forceinline__ double rcp_refined(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return r * fma(-x, r, 2.0);
}
__device__ __forceinline__ float rcp_refined(float x) { return __builtin_amdgcn_rcpf(x); }

// a / b from a reciprocal refined by one Newton step and one correction of
// the quotient by its residual: six VALU operations instead of the eleven of
// the IEEE division sequence, within one unit in the last place of the
// IEEE quotient (no special operands reach it: b is a nonzero difference
// of two values or of their logarithms, a positive series denominator or a
// sum of positive terms)
__device__ __forceinline__ double div_nr(double a, double b) {
  const double r = rcp_refined(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ float div_nr(float a, float b) { return div_(a, b); }

// square root of x >= 0 from the hardware reciprocal square root and the
// coupled Newton iterations for sqrt(x) and 1 / (2 sqrt(x)) of Goldschmidt
// (Markstein, IA-64 and Elementary Functions, 2000, ch. 9), without the
// exponent scaling that keeps a library square root exact near the ends of
// the exponent range (the speeds and normals here are far from them);
// sqrt(0) = 0
__device__ __forceinline__ double sqrt_(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  g = fma(d, h, g);
  return x == 0.0 ? 0.0 : g;
}
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }

// natural logarithm: x = 2^k m with m in [sqrt(1/2), sqrt(2)), f = m - 1
// exact, log(1 + f) = 2 atanh(s) with s = f / (2 + f), |s| < 0.172, written
// as f - (f^2/2 - s (f^2/2 + R)) with R = sum_{j=1..10} 2 s^(2j) / (2j + 1)
// (the truncated Taylor series; its remainder is below 2^-57 of the result),
// and k log 2 added in two parts (the upper part has 21 trailing zero bits,
// so k times it is exact).  Within 0.85 ulp of the exact value over
// [1e-300, 1e300] (measured against a long-double logarithm); zero,
// negative, subnormal and non-finite operands take the library logarithm.
__device__ __forceinline__ double log_(double x) {
  if (!(x >= 0x1p-1022 && x <= 0x1.fffffffffffffp+1023)) return log(x);
  int k = __builtin_amdgcn_frexp_exp(x);
  double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
  const bool lo = m < 0x1.6a09e667f3bcdp-1;   // sqrt(1/2)
  m = lo ? m + m : m;
  k = lo ? k - 1 : k;
  const double f = m - 1.0;
  const double s = div_nr(f, 2.0 + f);
  const double z = s * s;
  double R = 0x1.8618618618618p-4;  // 2/21
  R = fma(R, z, 0x1.af286bca1af28p-4);  // 2/19
  R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
  R = fma(R, z, 0x1.1111111111111p-3);  // 2/15
  R = fma(R, z, 0x1.3b13b13b13b14p-3);  // 2/13
  R = fma(R, z, 0x1.745d1745d1746p-3);  // 2/11
  R = fma(R, z, 0x1.c71c71c71c71cp-3);  // 2/9
  R = fma(R, z, 0x1.2492492492492p-2);  // 2/7
  R = fma(R, z, 0x1.999999999999ap-2);  // 2/5
  R = fma(R, z, 0x1.5555555555555p-1);  // 2/3
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double dk = double(k);
  return dk * 0x1.62e42fee00000p-1 -
         ((hfsq - fma(s, hfsq + R, dk * 0x1.a39ef35793c76p-33)) - f);
}
__device__ __forceinline__ float log_(float x) {
  return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

// Logarithmic mean and its inverse with the f^2 < 1e-4 series branch of
// Ranocha et al. (2023, ACM TOMS), the form of Trixi's ln_mean/inv_ln_mean
// imported at src/TrixiAtmo.jl:37.
template <typename T>
__device__ __forceinline__ T ln_mean(T x, T y) {
  const T f2 = div_nr(x * (x - T(2) * y) + y * y, x * (x + T(2) * y) + y * y);
  if (f2 < T(1.0e-4)) {
    return div_nr(x + y, T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0))));
  }
  return div_nr(y - x, log_(div_nr(y, x)));
}

template <typename T>
__device__ __forceinline__ T inv_ln_mean(T x, T y) {
  const T f2 = div_nr(x * (x - T(2) * y) + y * y, x * (x + T(2) * y) + y * y);
  if (f2 < T(1.0e-4)) {
    return div_nr(T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0))), x + y);
  }
  return div_nr(log_(div_nr(y, x)), y - x);
}

// ln_mean / inv_ln_mean from node logarithms: (y - x) / (log y - log x) away
// from x = y, the series of Ranocha et al. (2023) where (x-y)^2/(x+y)^2 < 1e-4.
// Both branches end in one division of a selected numerator and denominator,
// so lanes of a wave that take different branches do not serialise; f^2 uses
// a reciprocal refined by one Newton step (it only enters the small series
// terms).
template <typename T>
__device__ __forceinline__ T ln_mean_logs(T x, T y, T lx, T ly) {
  const T d = y - x, s = x + y;
  const bool ser = d * d < T(1.0e-4) * (s * s);
  const T f2 = ser ? (d * d) * rcp_refined(s * s) : T(0);
  const T q = T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0)));
  return div_nr(ser ? s : d, ser ? q : ly - lx);
}

// the same value with a branch instead of selects, for callers whose lanes
// mostly agree on the branch (the pairs of one line direction of an element)
template <typename T>
__device__ __forceinline__ T ln_mean_logs_branch(T x, T y, T lx, T ly) {
  const T d = y - x, s = x + y;
  if (d * d < T(1.0e-4) * (s * s)) {
    const T f2 = (d * d) * rcp_refined(s * s);
    return div_nr(s, T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0))));
  }
  return div_nr(d, ly - lx);
}

// ln_mean for the two states of a face node: they describe one physical
// point, so the series branch is the common one and the logarithms are taken
// only in the other branch (log y - log x there, which differs from log(y/x)
// by rounding only)
template <typename T>
__device__ __forceinline__ T ln_mean_face(T x, T y) {
  const T d = y - x, s = x + y;
  if (d * d < T(1.0e-4) * (s * s)) {
    const T f2 = (d * d) * rcp_refined(s * s);
    return div_nr(s, T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0))));
  }
  return div_nr(d, log_(y) - log_(x));
}

template <typename T>
__device__ __forceinline__ T inv_ln_mean_logs(T x, T y, T lx_minus_ly) {
  const T d = y - x, s = x + y;
  const bool ser = d * d < T(1.0e-4) * (s * s);
  const T f2 = ser ? (d * d) * rcp_refined(s * s) : T(0);
  const T q = T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0)));
  return div_nr(ser ? q : -lx_minus_ly, ser ? s : d);
}

// primitive quantities a two-point flux needs from one node
// (e = rho_e / rho, inv_rho = 1 / rho)
template <typename T>
struct Prim {
  T rho, v1, v2, v3, p, rho_e, phi, e, inv_rho;
};

template <typename T>
__device__ __forceinline__ Prim<T> cons2prim(const E3Params& q, const T* u) {
  Prim<T> w;
  w.rho = u[0];
  const T inv = div_nr(T(1), w.rho);
  w.v1 = u[1] * inv;
  w.v2 = u[2] * inv;
  w.v3 = u[3] * inv;
  w.p = T(q.gm1) * (u[4] - T(0.5) * (u[1] * w.v1 + u[2] * w.v2 + u[3] * w.v3));
  w.rho_e = u[4];
  w.phi = u[5];
  w.e = u[4] * inv;
  w.inv_rho = inv;
  return w;
}

// StepsizeCallback's scaled wave speeds at one node: l[d] = |1/J| |Ja^d . a|
// with a_c = |v_c| + c (src/callbacks_step/stepsize_dg2d.jl:5-43 in 3D);
// the max-speed reduction and the last stage of a resident step both call
// this one function, so their speeds agree bit for bit
__device__ __forceinline__ void node_scaled_speeds(const E3Params& q, const double* uu,
                                                   const double* ja, double ij_abs, double l[3]) {
  const Prim<double> w = cons2prim(q, uu);
  const double c = sqrt(q.gamma * w.p / w.rho);
  const double a1 = fabs(w.v1) + c, a2 = fabs(w.v2) + c, a3 = fabs(w.v3) + c;
  l[0] = ij_abs * fabs(ja[0] * a1 + ja[1] * a2 + ja[2] * a3);
  l[1] = ij_abs * fabs(ja[3] * a1 + ja[4] * a2 + ja[5] * a3);
  l[2] = ij_abs * fabs(ja[6] * a1 + ja[7] * a2 + ja[8] * a3);
}

// flux_ranocha, :382-415
template <typename T>
__device__ __forceinline__ void flux_ranocha(const E3Params& q, const Prim<T>& l, const Prim<T>& r,
                                             const T n0, const T n1, const T n2, T f[5]) {
  const T vnl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
  const T vnr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
  const T rho_mean = ln_mean(l.rho, r.rho);
  const T inv_rho_p_mean = l.p * r.p * inv_ln_mean(l.rho * r.p, r.rho * l.p);
  const T v1a = T(0.5) * (l.v1 + r.v1), v2a = T(0.5) * (l.v2 + r.v2), v3a = T(0.5) * (l.v3 + r.v3);
  const T pa = T(0.5) * (l.p + r.p);
  const T vsq = T(0.5) * (l.v1 * r.v1 + l.v2 * r.v2 + l.v3 * r.v3);
  const T f1 = rho_mean * T(0.5) * (vnl + vnr);
  f[0] = f1;
  f[1] = f1 * v1a + pa * n0;
  f[2] = f1 * v2a + pa * n1;
  f[3] = f1 * v3a + pa * n2;
  f[4] = f1 * (vsq + inv_rho_p_mean * T(q.inv_gm1)) + T(0.5) * (l.p * vnr + r.p * vnl);
}

// flux_kennedy_gruber, :289-318
template <typename T>
__device__ __forceinline__ void flux_kennedy_gruber(const Prim<T>& l, const Prim<T>& r, const T n0,
                                                    const T n1, const T n2, T f[5]) {
  const T ra = T(0.5) * (l.rho + r.rho);
  const T v1a = T(0.5) * (l.v1 + r.v1), v2a = T(0.5) * (l.v2 + r.v2), v3a = T(0.5) * (l.v3 + r.v3);
  const T pa = T(0.5) * (l.p + r.p);
  const T ea = T(0.5) * (l.e + r.e);
  const T vna = v1a * n0 + v2a * n1 + v3a * n2;
  const T f1 = ra * vna;
  f[0] = f1;
  f[1] = f1 * v1a + pa * n0;
  f[2] = f1 * v2a + pa * n1;
  f[3] = f1 * v3a + pa * n2;
  f[4] = f1 * ea + pa * vna;
}

// flux_shima_etal, :248-276
template <typename T>
__device__ __forceinline__ void flux_shima(const E3Params& q, const Prim<T>& l, const Prim<T>& r,
                                           const T n0, const T n1, const T n2, T f[5]) {
  const T vnl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
  const T vnr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
  const T ra = T(0.5) * (l.rho + r.rho);
  const T v1a = T(0.5) * (l.v1 + r.v1), v2a = T(0.5) * (l.v2 + r.v2), v3a = T(0.5) * (l.v3 + r.v3);
  const T vna = T(0.5) * (vnl + vnr);
  const T pa = T(0.5) * (l.p + r.p);
  const T vsq = T(0.5) * (l.v1 * r.v1 + l.v2 * r.v2 + l.v3 * r.v3);
  const T f1 = ra * vna;
  f[0] = f1;
  f[1] = f1 * v1a + pa * n0;
  f[2] = f1 * v2a + pa * n1;
  f[3] = f1 * v3a + pa * n2;
  f[4] = f1 * vsq + pa * vna * T(q.inv_gm1) + T(0.5) * (l.p * vnr + r.p * vnl);
}

// pointwise flux, :77-91.  The reference takes `rho_e = last(u)`, the sixth
// entry (phi), for the energy flux; kept as the reference has it.
template <typename T>
__device__ __forceinline__ void flux_point(const Prim<T>& w, const T n0, const T n1, const T n2,
                                           T f[5]) {
  const T vn = w.v1 * n0 + w.v2 * n1 + w.v3 * n2;
  const T rvn = w.rho * vn;
  f[0] = rvn;
  f[1] = rvn * w.v1 + w.p * n0;
  f[2] = rvn * w.v2 + w.p * n1;
  f[3] = rvn * w.v3 + w.p * n2;
  f[4] = (w.phi + w.p) * vn;
}

template <typename T>
__device__ __forceinline__ void flux_central(const Prim<T>& l, const Prim<T>& r, const T n0,
                                             const T n1, const T n2, T f[5]) {
  T fl[5], fr[5];
  flux_point(l, n0, n1, n2, fl);
  flux_point(r, n0, n1, n2, fr);
#pragma unroll
  for (int v = 0; v < 5; ++v) f[v] = T(0.5) * (fl[v] + fr[v]);
}

// max_abs_speed with a normal direction, :538-557
template <typename T>
__device__ __forceinline__ T max_abs_speed(const E3Params& q, const Prim<T>& l, const Prim<T>& r,
                                           const T n0, const T n1, const T n2) {
  const T vl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
  const T vr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
  const T cl = sqrt_(T(q.gamma) * l.p * l.inv_rho);
  const T cr = sqrt_(T(q.gamma) * r.p * r.inv_rho);
  const T nn = sqrt_(n0 * n0 + n1 * n1 + n2 * n2);
  return fmax(fabs(vl) + cl * nn, fabs(vr) + cr * nn);
}

// FluxLMARS, :430-466
template <typename T>
__device__ __forceinline__ void flux_lmars(const E3Params& q, const T* ul, const T* ur,
                                           const Prim<T>& l, const Prim<T>& r, const T n0,
                                           const T n1, const T n2, T f[6]) {
  const T c = T(q.lmars_c);
  const T vl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
  const T vr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
  const T nn = sqrt_(n0 * n0 + n1 * n1 + n2 * n2);
  const T rho = T(0.5) * (l.rho + r.rho);
  const T p = T(0.5) * (l.p + r.p) - div_(T(0.5) * c * rho * (vr - vl), nn);
  const T v = T(0.5) * (vl + vr) - div_(T(1), T(2) * c * rho) * (r.p - l.p) * nn;
  const T* us = v >= T(0) ? ul : ur;
  const T ps = v >= T(0) ? l.p : r.p;
  f[0] = v * us[0];
  f[1] = v * us[1] + p * n0;
  f[2] = v * us[2] + p * n1;
  f[3] = v * us[3] + p * n2;
  f[4] = v * us[4] + ps * v;
  f[5] = T(0);
}

// flux_nonconservative_waruszewski_etal (:135-159) / souza_etal (:206-230):
// g = rho_avg [phi] (0, n, n.v_avg, 0); only entries 1..4 are nonzero.
template <int NC, typename T>
__device__ __forceinline__ void noncons_gravity(const Prim<T>& l, const Prim<T>& r, const T n0,
                                                const T n1, const T n2, T g[4]) {
  const T ra = (NC == 1) ? ln_mean(l.rho, r.rho) : T(0.5) * (l.rho + r.rho);
  const T jump = r.phi - l.phi;
  const T s = ra * jump;
  const T v1a = T(0.5) * (l.v1 + r.v1), v2a = T(0.5) * (l.v2 + r.v2), v3a = T(0.5) * (l.v3 + r.v3);
  g[0] = n0 * s;
  g[1] = n1 * s;
  g[2] = n2 * s;
  g[3] = n0 * s * v1a + n1 * s * v2a + n2 * s * v3a;
}

}  // namespace atmo

#endif
