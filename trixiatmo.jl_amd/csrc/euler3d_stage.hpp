// Fused RK-stage kernel of the 3D DGSEM right-hand side for
// CompressibleEulerEnergyEquationsWithGravity3D on curved hexahedra
// (the work of Trixi's rhs! for P4estMesh{3}: volume flux differencing,
// prolong2interfaces + calc_interface_flux!, boundary fluxes, surface
// integral, apply_jacobian!, calc_sources!, then the 2N RK update).
//
// One workgroup owns one element.  Per direction, lanes evaluate every
// unordered node pair of every line once (two-point volume flux f# and the
// gravity term g with the averaged contravariant vector) into LDS; then every
// node gathers its n-1 pair terms with Dsplit[own, m] (g enters with the sign
// of (own - m), being antisymmetric).  Lines are processed in chunks so the
// pair buffer fits LDS for every polydeg.  Then one lane per face node
// evaluates the surface flux (neighbour state read from HBM, primary normal
// as Trixi's calc_interface_flux! uses it) into LDS, and every node adds its
// lifted face terms, multiplies by -1/J, adds the node-local source and
// either stores du or applies k = a k + dt du, u_out = u + b k (ping-pong
// state buffers, since neighbours still read the old state's face nodes).
//
// Compute type T: double (FP64 path, compared with the oracle to 1e-12; its
// logarithm, square root and hot quotients are the within-one-ulp helpers of
// euler3d_physics.hpp) or float (mixed path: node states, pressure recovery,
// geopotential jumps, the dissipation's state jumps, the
// hydrostatic-background split and the RK registers stay FP64;
// fluxes, geometry and volume sums are FP32, with Kahan-compensated sums when
// KAHAN is set).  With the background split the momentum rows carry only
// p' = p - p_bg and (rho_avg - rho_avg,bg) [phi]; the FP64 tendency of the
// background is added back, so the discretisation is unchanged.
#ifndef ATMO_EULER3D_STAGE_HPP
#define ATMO_EULER3D_STAGE_HPP

#include <hip/hip_runtime.h>

#include <type_traits>

#include "euler3d_kernels.hpp"
#include "euler3d_physics.hpp"
#include "speed_max.hpp"

namespace atmo {

template <int N>
struct Shape {
  static constexpr int N2 = N * N;
  static constexpr int N3 = N2 * N;
  static constexpr int BLOCK = ((N3 + 63) / 64) * 64;
  static constexpr int P = N * (N - 1) / 2;  // unordered pairs per line
};

// pair-buffer budget per workgroup (bytes); lines are chunked to fit.  A pair
// record holds the 5 volume-flux entries and, with a gravity term, its
// momentum factor sm and energy entry (the momentum entries n * sm are formed
// again from the averaged normal when the record is gathered).
template <int N, typename T, int NC = 1>
struct Chunking {
  static constexpr int BUDGET = 40 * 1024;
  static constexpr int REC = NC != 0 ? 7 : 5;
  static constexpr int PAIR_BYTES = REC * (int)sizeof(T);
  static constexpr int LINES_MAX = BUDGET / (PAIR_BYTES * Shape<N>::P);
  static constexpr int LINES = LINES_MAX >= Shape<N>::N2 ? Shape<N>::N2 : LINES_MAX;
  static constexpr int CHUNKS = (Shape<N>::N2 + LINES - 1) / LINES;
  static constexpr int NPC = LINES * Shape<N>::P;  // pairs per chunk
  static constexpr int FACE = 36 * Shape<N>::N2;   // face buffer [6 vars][6 n^2]
  static constexpr int FLEN = REC * NPC > FACE ? REC * NPC : FACE;
};

// running sum, optionally Kahan-compensated
template <typename T, bool KAHAN>
struct Acc {
  T s = T(0), c = T(0);
  __device__ __forceinline__ void add(T t) {
    if constexpr (KAHAN) {
      const T y = t - c;
      const T z = s + y;
      c = (z - s) - y;
      s = z;
    } else {
      s += t;
    }
  }
  __device__ __forceinline__ T value() const { return s; }
  // multiply by a power of two (exact)
  __device__ __forceinline__ void scale(T f) {
    s *= f;
    c *= f;
  }
};

// the dissipation's jumps are formed from the FP64 states uld, urd and then
// rounded to T (the two states of a face node differ by little, so the
// difference of the rounded states would be mostly their rounding)
template <typename T>
__device__ __forceinline__ void surface_flux(const E3Params& q, int surf, const T* ul, const T* ur,
                                             const double* uld, const double* urd,
                                             const Prim<T>& l, const Prim<T>& r, T n0, T n1, T n2,
                                             T f[6]) {
  if (surf == 0) {  // flux_lax_friedrichs
    flux_central(l, r, n0, n1, n2, f);
    const T lam = max_abs_speed(q, l, r, n0, n1, n2);
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] += T(-0.5) * lam * T(urd[v] - uld[v]);
    f[5] = T(-0.5) * lam * T(urd[5] - uld[5]);
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
    for (int v = 0; v < 5; ++v) f[v] += T(-0.5) * lam * T(urd[v] - uld[v]);
    f[5] = T(-0.5) * lam * T(urd[5] - uld[5]);
  }
}

template <typename T>
__device__ __forceinline__ void surface_noncons(int snc, const Prim<T>& l, const Prim<T>& r, T n0,
                                                T n1, T n2, T g[4]) {
  if (snc == 1) noncons_gravity<1>(l, r, n0, n1, n2, g);
  else noncons_gravity<2>(l, r, n0, n1, n2, g);
}

// Per-node quantities of the two-point fluxes, with the divisions and
// logarithms done once per node instead of once per pair.
template <typename T>
struct FastPrim {
  T rho, v1, v2, v3, p, e, phi, lrho, lrp;  // e = rho_e / rho, lrp = log(rho / p)
  // hydrostatic-background split (mixed path): p' = p - p_bg, rho' = rho - rho_bg,
  // lam = log(rho / rho_bg); the background is formed in FP64 from phi
  T pp, rb, rp, lam;
};

// Hydrostatic background of the mixed path (struct Background): the isothermal
// atmosphere p_bg = p0 exp(-phi / (R T0)), rho_bg = p_bg / (R T0), and an FP32
// copy B of the FP64 rhs of that background state (momentum rows only; mass
// and energy tendencies of a state at rest vanish).

// delta = ln_mean(x, y) - ln_mean(xb, yb) from perturbation quantities, with
// the branch of the actual state: series form where (y-x)^2 < 1e-4 (x+y)^2,
// else the logarithmic form with dLb = log yb - log xb, dlam = log(y/yb) -
// log(x/xb) (exact identities; no difference of two rounded means is formed).
template <typename T>
__device__ __forceinline__ T delta_ln_mean(T A, T s, T Ab, T sb, T Ap, T sp, T dLb, T dlam) {
  if (A * A < T(1.0e-4) * (s * s)) {
    const T f = div_(A, s), fb = div_(Ab, sb);
    const T f2 = f * f, fb2 = fb * fb;
    const T Q = T(2) + f2 * (T(2.0 / 3.0) + f2 * (T(2.0 / 5.0) + f2 * T(2.0 / 7.0)));
    const T Qb = T(2) + fb2 * (T(2.0 / 3.0) + fb2 * (T(2.0 / 5.0) + fb2 * T(2.0 / 7.0)));
    const T df = div_(Ap * sb - Ab * sp, s * sb);  // f - fb
    const T dQ = -df * (f + fb) *
                 (T(2.0 / 3.0) + T(2.0 / 5.0) * (fb2 + f2) + T(2.0 / 7.0) * (fb2 * fb2 + fb2 * f2 + f2 * f2));
    return div_(sp, Q) + div_(sb * dQ, Q * Qb);  // dQ = Qb - Q
  }
  const T dL = dLb + dlam;
  if (fabs(dLb) > T(1e-6) * fabs(dL)) return div_(Ap * dLb - Ab * dlam, dLb * dL);
  return div_(A, dL) - T(0.5) * sb;  // flat background: ln_mean(xb, yb) = (xb + yb) / 2
}

// the volume flux f (5 entries) and the gravity term g (4 entries) of one pair
template <int VOL, int NC, typename T>
__device__ __forceinline__ void pair_flux(const E3Params& q, const FastPrim<T>& l,
                                          const FastPrim<T>& r, T n0, T n1, T n2, T dphi,
                                          const Background& bg, T f[5], T g[5]) {
  const T v1a = T(0.5) * (l.v1 + r.v1), v2a = T(0.5) * (l.v2 + r.v2), v3a = T(0.5) * (l.v3 + r.v3);
  const T pa = T(0.5) * (l.p + r.p);
  // momentum pressure term: p' when the hydrostatic background is split off
  const T pm = bg.on ? T(0.5) * (l.pp + r.pp) : pa;
  if constexpr (VOL == 0) {  // flux_ranocha, :382-415
    const T vnl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
    const T vnr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
    const T rho_mean = ln_mean_logs(l.rho, r.rho, l.lrho, r.lrho);
    // inv_ln_mean(rho_l p_r, rho_r p_l): log of the ratio = lrp_r - lrp_l
    const T inv_rp = l.p * r.p * inv_ln_mean_logs(l.rho * r.p, r.rho * l.p, l.lrp - r.lrp);
    const T vsq = T(0.5) * (l.v1 * r.v1 + l.v2 * r.v2 + l.v3 * r.v3);
    const T f1 = rho_mean * T(0.5) * (vnl + vnr);
    f[0] = f1;
    f[1] = f1 * v1a + pm * n0;
    f[2] = f1 * v2a + pm * n1;
    f[3] = f1 * v3a + pm * n2;
    f[4] = f1 * (vsq + inv_rp * T(q.inv_gm1)) + T(0.5) * (l.p * vnr + r.p * vnl);
  } else if constexpr (VOL == 1) {  // flux_kennedy_gruber, :289-318
    const T ra = T(0.5) * (l.rho + r.rho);
    const T ea = T(0.5) * (l.e + r.e);
    const T vna = v1a * n0 + v2a * n1 + v3a * n2;
    const T f1 = ra * vna;
    f[0] = f1;
    f[1] = f1 * v1a + pm * n0;
    f[2] = f1 * v2a + pm * n1;
    f[3] = f1 * v3a + pm * n2;
    f[4] = f1 * ea + pa * vna;
  } else {  // flux_shima_etal, :248-276
    const T vnl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2;
    const T vnr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
    const T ra = T(0.5) * (l.rho + r.rho);
    const T vna = T(0.5) * (vnl + vnr);
    const T vsq = T(0.5) * (l.v1 * r.v1 + l.v2 * r.v2 + l.v3 * r.v3);
    const T f1 = ra * vna;
    f[0] = f1;
    f[1] = f1 * v1a + pm * n0;
    f[2] = f1 * v2a + pm * n1;
    f[3] = f1 * v3a + pm * n2;
    f[4] = f1 * vsq + pa * vna * T(q.inv_gm1) + T(0.5) * (l.p * vnr + r.p * vnl);
  }
  if constexpr (NC != 0) {  // waruszewski (:135-159) / souza (:206-230)
    const T ra = (NC == 1) ? ln_mean_logs(l.rho, r.rho, l.lrho, r.lrho) : T(0.5) * (l.rho + r.rho);
    const T sj = ra * dphi;
    T sm = sj;  // momentum part: (rho_avg - rho_avg of the background) [phi]
    if (bg.on) {
      const T Ap = r.rp - l.rp, sp = l.rp + r.rp;
      const T Ab = r.rb - l.rb, sb = l.rb + r.rb;
      const T dmean = (NC == 1) ? delta_ln_mean(Ap + Ab, sp + sb, Ab, sb, Ap, sp,
                                                -dphi * T(bg.inv_RT0), r.lam - l.lam)
                                : T(0.5) * sp;
      sm = dmean * dphi;
    }
    g[0] = n0 * sm;
    g[1] = n1 * sm;
    g[2] = n2 * sm;
    g[3] = n0 * sj * v1a + n1 * sj * v2a + n2 * sj * v3a;
    g[4] = sm;
  }
}

template <typename T>
__device__ __forceinline__ Prim<T> to_t(const Prim<double>& w) {
  Prim<T> o;
  o.rho = T(w.rho); o.v1 = T(w.v1); o.v2 = T(w.v2); o.v3 = T(w.v3);
  o.p = T(w.p); o.rho_e = T(w.rho_e); o.phi = T(w.phi);
  o.e = T(w.e); o.inv_rho = T(w.inv_rho);
  return o;
}

// Surface fluxes of all 6 n^2 face nodes of element e, one lane per face node,
// lifted by 1/w_end into LDS as s_sf[v * 6 n^2 + f n^2 + a + n b].  Node
// states and their primitives are formed in FP64 and then rounded to T.
// Momentum rows of the surface flux with the background pressure split off
// (p' for the pressure terms; wave speeds and upwinding keep the full state).
template <typename T>
__device__ __forceinline__ void surface_momentum_split(const E3Params& q, int surf, const T* ul,
                                                       const T* ur, const double* uld,
                                                       const double* urd, const Prim<T>& l,
                                                       const Prim<T>& r, T ppl, T ppr, T n0, T n1,
                                                       T n2, T m[3]) {
  const T nv[3] = {n0, n1, n2};
  const T vl[3] = {l.v1, l.v2, l.v3}, vr[3] = {r.v1, r.v2, r.v3};
  const T vnl = l.v1 * n0 + l.v2 * n1 + l.v3 * n2, vnr = r.v1 * n0 + r.v2 * n1 + r.v3 * n2;
  const T pma = T(0.5) * (ppl + ppr);
  if (surf == 0) {
    const T lam = max_abs_speed(q, l, r, n0, n1, n2);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      m[k] = T(0.5) * ((l.rho * vnl * vl[k] + ppl * nv[k]) + (r.rho * vnr * vr[k] + ppr * nv[k])) +
             T(-0.5) * lam * T(urd[k + 1] - uld[k + 1]);
  } else if (surf == 3) {  // FluxLMARS
    const T c = T(q.lmars_c);
    const T nn = sqrt_(n0 * n0 + n1 * n1 + n2 * n2);
    const T rho = T(0.5) * (l.rho + r.rho);
    const T pm = pma - div_(T(0.5) * c * rho * (vnr - vnl), nn);
    const T v = T(0.5) * (vnl + vnr) - div_(T(1), T(2) * c * rho) * (r.p - l.p) * nn;
    const T* us = v >= T(0) ? ul : ur;
#pragma unroll
    for (int k = 0; k < 3; ++k) m[k] = v * us[k + 1] + pm * nv[k];
  } else {
    T f1;
    if (surf == 1 || surf == 5) f1 = ln_mean(l.rho, r.rho) * T(0.5) * (vnl + vnr);
    else
      f1 = T(0.5) * (l.rho + r.rho) *
           (T(0.5) * (l.v1 + r.v1) * n0 + T(0.5) * (l.v2 + r.v2) * n1 + T(0.5) * (l.v3 + r.v3) * n2);
#pragma unroll
    for (int k = 0; k < 3; ++k) m[k] = f1 * T(0.5) * (vl[k] + vr[k]) + pma * nv[k];
    if (surf >= 4) {
      const T lam = max_abs_speed(q, l, r, n0, n1, n2);
#pragma unroll
      for (int k = 0; k < 3; ++k) m[k] += T(-0.5) * lam * T(urd[k + 1] - uld[k + 1]);
    }
  }
}

// face-node passes of one workgroup and the neighbour states they need,
// loaded at the start of the kernel so that the dependent loads (link, then
// the neighbour's state) overlap the volume phase; from polydeg 4 on only the
// first pass (every lane of the workgroup) is: polydeg 4's partial second
// pass then loads in place, which keeps 14 registers free during the volume
// phase (5 % faster there; polydeg 3, one wave per element, prefetches both).
// A workgroup holds EPB elements (slots es = 0..EPB-1 of Shape<N>::BLOCK
// lanes each); its EPB * 6 n^2 face nodes are spread over all its lanes,
// face node fg = t + ps * EPB * BLOCK of pass ps being node fg mod 6 n^2 of
// slot fg / (6 n^2).
template <int N, int EPB = 1>
struct FacePrefetch {
  static constexpr int GB = EPB * Shape<N>::BLOCK;
  static constexpr int NF = 6 * Shape<N>::N2;
  static constexpr int PASSES = (EPB * NF + GB - 1) / GB;
  static constexpr int PF = N >= 5 ? 1 : PASSES;
  int link[PF], info[PF];
  double ur[PF][6];
};

// link (neighbour's global node, < 0 on a boundary) and face_info of pass ps;
// elements e0 .. e0 + n_valid - 1 are the workgroup's
template <int N, int EPB = 1>
__device__ __forceinline__ void face_link(const FacePrefetch<N, EPB>& pf, int e0, int n_valid,
                                          int t, int ps, const int* __restrict__ nbr,
                                          const int* __restrict__ face_info, int& link,
                                          int& info) {
  using F = FacePrefetch<N, EPB>;
  constexpr int N2 = Shape<N>::N2, NF = F::NF;
  if (ps < F::PF) {
    link = pf.link[ps < F::PF ? ps : 0];
    info = pf.info[ps < F::PF ? ps : 0];
    return;
  }
  const int fg = t + ps * F::GB;
  link = -1;
  info = 0;
  const int es = EPB == 1 ? 0 : fg / NF;
  if (fg < EPB * NF && es < n_valid) {
    const int fid = fg - es * NF;
    const int f = fid / N2, ab = fid - f * N2;
    const size_t e = (size_t)(e0 + es);
    link = nbr[(e * 6 + f) * N2 + ab];
    info = face_info[e * 6 + f];
  }
}

template <int N, int EPB = 1>
__device__ __forceinline__ void face_prefetch(FacePrefetch<N, EPB>& pf, int e0, int n_valid, int t,
                                              const double* __restrict__ u,
                                              const int* __restrict__ nbr,
                                              const int* __restrict__ face_info) {
  using F = FacePrefetch<N, EPB>;
  constexpr int N2 = Shape<N>::N2, NF = F::NF;
#pragma unroll
  for (int ps = 0; ps < F::PF; ++ps) {
    const int fg = t + ps * F::GB;
    const int es = EPB == 1 ? 0 : fg / NF;
    pf.link[ps] = -1;
    pf.info[ps] = 0;
    if (fg < EPB * NF && es < n_valid) {
      const int fid = fg - es * NF;
      const int f = fid / N2, ab = fid - f * N2;
      const size_t e = (size_t)(e0 + es);
      pf.link[ps] = nbr[(e * 6 + f) * N2 + ab];
      pf.info[ps] = face_info[e * 6 + f];
    }
  }
#pragma unroll
  for (int ps = 0; ps < F::PF; ++ps) {
    const int link = pf.link[ps];
#pragma unroll
    for (int v = 0; v < 6; ++v) pf.ur[ps][v] = link >= 0 ? u[(size_t)link * 6 + v] : 0.0;
  }
}

// face-node geometry of pass ps: element slot es, face node fid of that
// element, face f, direction d, own node; false when the lane has no face
// node in that pass
template <int N, int EPB = 1>
__device__ __forceinline__ bool face_node_of(int t, int ps, int n_valid, int& es, int& fid, int& f,
                                             int& d, int& node) {
  using F = FacePrefetch<N, EPB>;
  constexpr int N2 = Shape<N>::N2, NF = F::NF;
  const int fg = t + ps * F::GB;
  es = EPB == 1 ? 0 : fg / NF;
  if (fg >= EPB * NF || es >= n_valid) return false;
  fid = fg - es * NF;
  f = fid / N2;
  const int ab = fid - f * N2;
  const int a = ab % N, b = ab / N;
  d = f / 2;
  const int fixed = (f % 2 == 0) ? 0 : N - 1;
  node = d == 0 ? fixed + N * (a + N * b) : (d == 1 ? a + N * (fixed + N * b) : a + N * (b + N * fixed));
  return true;
}

// normal of every face pass: the own outward contravariant vector on a primary
// or boundary face, minus the primary's on a secondary face (Trixi evaluates
// an interface with the primary element's normal).  Read before the face
// fluxes overwrite the LDS copy of the own vectors (LDS_CV: s_cv = [9][n^3]
// per slot, slot stride cv_stride; else read from memory).  A compile-time
// choice: a run-time null test leaves the compiler a pointer that may be LDS
// or global memory, which it reads with flat loads.
template <int N, typename T, int EPB = 1>
struct FaceNormals {
  T n[FacePrefetch<N, EPB>::PASSES][3];
};

template <int N, typename T, int EPB = 1, typename G = T, bool LDS_CV = true>
__device__ __forceinline__ void face_normals(FaceNormals<N, T, EPB>& fn, int e0, int n_valid, int t,
                                             const G* __restrict__ cv, const T* s_cv,
                                             int cv_stride, const FacePrefetch<N, EPB>& pf,
                                             const int* __restrict__ nbr,
                                             const int* __restrict__ face_info) {
  constexpr int N3 = Shape<N>::N3;
#pragma unroll
  for (int ps = 0; ps < FacePrefetch<N, EPB>::PASSES; ++ps) {
    int es, fid, f, d, node;
    fn.n[ps][0] = fn.n[ps][1] = fn.n[ps][2] = T(0);
    if (!face_node_of<N, EPB>(t, ps, n_valid, es, fid, f, d, node)) continue;
    // face_info: 0 = primary side or boundary; 8 | dp = secondary side of an
    // interface whose primary face direction is dp
    int link, info;
    face_link<N, EPB>(pf, e0, n_valid, t, ps, nbr, face_info, link, info);
    if (info == 0) {
      const T sgn = (f % 2 == 0) ? T(-1) : T(1);
      const size_t e = (size_t)(e0 + es);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if constexpr (LDS_CV) fn.n[ps][c] = sgn * s_cv[es * cv_stride + (3 * d + c) * N3 + node];
        else fn.n[ps][c] = sgn * T(cv[(e * N3 + node) * 9 + 3 * d + c]);
    } else {
      const int dp = info & 7;
      const T sgn = (dp % 2 == 0) ? T(1) : T(-1);  // minus the primary's outward normal
      const G* jp = cv + (size_t)link * 9 + 3 * (dp / 2);
#pragma unroll
      for (int c = 0; c < 3; ++c) fn.n[ps][c] = sgn * T(jp[c]);
    }
  }
}

// the surface fluxes of the workgroup's face nodes, lifted by 1/w_end, into
// s_sf[es * sf_stride + v * 6 n^2 + fid]; the own primitives come from the
// slot's node fields s_w[es * w_stride + ...] (and s_pb[es * N3 + ...])
template <int N, typename T, int EPB = 1>
__device__ __forceinline__ void face_fluxes_to_lds(
    const E3Params& q, int e0, int n_valid, int t, double inv_w, const double* __restrict__ u,
    const FacePrefetch<N, EPB>& pf, const FaceNormals<N, T, EPB>& fn,
    const int* __restrict__ nbr, const int* __restrict__ face_info, int surf, int snc,
    const Background& bg, const T* s_w_all, int w_stride, const double* s_pb_all,
    T* s_sf_all, int sf_stride) {
  constexpr int N2 = Shape<N>::N2, N3 = Shape<N>::N3, NF = 6 * N2;
#pragma unroll
  for (int ps = 0; ps < FacePrefetch<N, EPB>::PASSES; ++ps) {
    int es, fid, f, d, node;
    if (!face_node_of<N, EPB>(t, ps, n_valid, es, fid, f, d, node)) continue;
    const T* s_w = s_w_all + es * w_stride;
    const double* s_pb = s_pb_all + es * N3;
    T* s_sf = s_sf_all + es * sf_stride;
    const int e = e0 + es;
    double uo[6];
    if constexpr (sizeof(T) == 8) {
      // the own state from the primitives in LDS (rounding-level differences
      // only; it enters the dissipation and the wall mirror)
      const double r = s_w[0 * N3 + node];
      uo[0] = r;
      uo[1] = r * s_w[1 * N3 + node];
      uo[2] = r * s_w[2 * N3 + node];
      uo[3] = r * s_w[3 * N3 + node];
      uo[4] = r * s_w[5 * N3 + node];
      uo[5] = s_w[6 * N3 + node];
    } else {
      // the mixed path forms p' = p - p_bg from the exact FP64 state
      const size_t gn = (size_t)e * N3 + node;
#pragma unroll
      for (int v = 0; v < 6; ++v) uo[v] = u[gn * 6 + v];
    }
    int link, info;
    face_link<N, EPB>(pf, e0, n_valid, t, ps, nbr, face_info, link, info);
    const T n0 = fn.n[ps][0], n1 = fn.n[ps][1], n2 = fn.n[ps][2];
    double ur[6];
    if (link >= 0) {
#pragma unroll
      for (int v = 0; v < 6; ++v)
        ur[v] = ps < FacePrefetch<N, EPB>::PF ? pf.ur[ps < FacePrefetch<N, EPB>::PF ? ps : 0][v]
                                              : u[(size_t)link * 6 + v];
    } else {
      // boundary_condition_slip_wall (:93-116): mirror the normal momentum
      const double m0 = n0, m1 = n1, m2 = n2;
      const double nn = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
      const double h0 = m0 / nn, h1 = m1 / nn, h2 = m2 / nn;
      const double un = h0 * uo[1] + h1 * uo[2] + h2 * uo[3];
      ur[0] = uo[0];
      ur[1] = uo[1] - 2.0 * un * h0;
      ur[2] = uo[2] - 2.0 * un * h1;
      ur[3] = uo[3] - 2.0 * un * h2;
      ur[4] = uo[4];
      ur[5] = uo[5];
    }
    // own primitives as the prologue formed them (rho, v, p, e per node in LDS)
    Prim<T> own;
    own.rho = s_w[0 * N3 + node]; own.v1 = s_w[1 * N3 + node]; own.v2 = s_w[2 * N3 + node];
    own.v3 = s_w[3 * N3 + node]; own.p = s_w[4 * N3 + node]; own.e = s_w[5 * N3 + node];
    own.rho_e = T(uo[4]);
    own.phi = T(uo[5]);
    own.inv_rho = div_nr(T(1), own.rho);
    const Prim<T> pr = to_t<T>(cons2prim(q, ur));
    T ut[6], urt[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) { ut[v] = T(uo[v]); urt[v] = T(ur[v]); }
    T fs[6];
    surface_flux<T>(q, surf, ut, urt, uo, ur, own, pr, n0, n1, n2, fs);
    double pb_o = 0.0, pb_n = 0.0;
    if (bg.on) {
      // own background pressure from the prologue; across a face phi is
      // continuous up to rounding, so the neighbour's follows to first order
      pb_o = s_pb[node];
      const double dph = ur[5] - uo[5];
      pb_n = fabs(dph * bg.inv_RT0) < 1e-8 ? pb_o * (1.0 - dph * bg.inv_RT0)
                                           : bg.p0 * exp(-ur[5] * bg.inv_RT0);
      T m[3];
      const double pp_o = (q.gm1 * (uo[4] - 0.5 * (uo[1] * uo[1] + uo[2] * uo[2] + uo[3] * uo[3]) / uo[0])) - pb_o;
      const double pp_n = (q.gm1 * (ur[4] - 0.5 * (ur[1] * ur[1] + ur[2] * ur[2] + ur[3] * ur[3]) / ur[0])) - pb_n;
      surface_momentum_split<T>(q, surf, ut, urt, uo, ur, own, pr, T(pp_o), T(pp_n), n0, n1, n2, m);
      fs[1] = m[0];
      fs[2] = m[1];
      fs[3] = m[2];
    }
    if (snc != 0) {
      // gravity term with the phi jump formed in FP64
      const T dphi = T(ur[5] - uo[5]);
      const T ra = snc == 1 ? ln_mean_face(own.rho, pr.rho) : T(0.5) * (own.rho + pr.rho);
      const T sj = ra * dphi;
      T sm = sj;
      if (bg.on) {
        // the face jump of phi is rounding-sized: the background part of the
        // mean density uses the arithmetic mean of the perturbations
        const double rb_o = pb_o * bg.inv_RT0, rb_n = pb_n * bg.inv_RT0;
        sm = T(0.5 * ((uo[0] - rb_o) + (ur[0] - rb_n))) * dphi;
      }
      const T v1a = T(0.5) * (own.v1 + pr.v1), v2a = T(0.5) * (own.v2 + pr.v2),
              v3a = T(0.5) * (own.v3 + pr.v3);
      fs[1] += T(0.5) * (n0 * sm);
      fs[2] += T(0.5) * (n1 * sm);
      fs[3] += T(0.5) * (n2 * sm);
      fs[4] += T(0.5) * (n0 * sj * v1a + n1 * sj * v2a + n2 * sj * v3a);
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) s_sf[v * NF + fid] = fs[v] * T(inv_w);
  }
}

struct DsplitTable {
  double ds[64];  // Dsplit[i * n + m]
};

// element handled by workgroup b: workgroups are dispatched round-robin over
// the 8 XCDs, so XCD x receives b = x, x + 8, ...; mapping those to one
// contiguous element range keeps neighbouring elements in the same L2
__device__ __forceinline__ int xcd_element(int b, int n_elements) {
  const int per = (n_elements + 7) / 8;
  return (b & 7) * per + (b >> 3);
}

// per-node fields of the two-point fluxes, formed in FP64 from the node state:
// rho v1 v2 v3 p e phi [log(rho)] [log(rho / p)] [4 of the background split]
template <int VOL, int NC, typename T>
struct NodeFields {
  static constexpr bool MIXED = sizeof(T) == 4;
  static constexpr int NW = MIXED ? 13 : (VOL == 0 ? 9 : 8);
};

template <int N, int VOL, int NC, typename T>
__device__ __forceinline__ void node_fields_to_lds(const E3Params& q, const double* __restrict__ u,
                                                   size_t gnode, int t, T* s_w, double* s_phi,
                                                   double* s_pb, const Background& bg) {
  constexpr int N3 = Shape<N>::N3;
  constexpr bool MIXED = NodeFields<VOL, NC, T>::MIXED;
  double uo[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) uo[v] = u[gnode * 6 + v];
  // primitives (pressure from total minus kinetic energy) in FP64
  const Prim<double> w = cons2prim(q, uo);
  s_w[0 * N3 + t] = T(w.rho); s_w[1 * N3 + t] = T(w.v1); s_w[2 * N3 + t] = T(w.v2);
  s_w[3 * N3 + t] = T(w.v3); s_w[4 * N3 + t] = T(w.p);
  s_w[5 * N3 + t] = T(div_nr(w.rho_e, w.rho));
  s_w[6 * N3 + t] = T(w.phi);
  if constexpr (VOL == 0 || NC == 1) s_w[7 * N3 + t] = log_(T(w.rho));
  if constexpr (VOL == 0) s_w[8 * N3 + t] = log_(T(w.rho / w.p));
  if constexpr (MIXED) {
    s_phi[t] = w.phi;
    if (bg.on) {
      // hydrostatic background in FP64; FP32 carries only the perturbation
      const double pb = bg.p0 * exp(-w.phi * bg.inv_RT0);
      const double rb = pb * bg.inv_RT0;
      s_pb[t] = pb;
      s_w[9 * N3 + t] = T(w.p - pb);
      s_w[10 * N3 + t] = T(rb);
      s_w[11 * N3 + t] = T(w.rho - rb);
      s_w[12 * N3 + t] = log1pf(float((w.rho - rb) / rb));  // log(rho / rho_bg)
    }
  }
}

template <int VOL, int NC, int N3, typename T>
__device__ __forceinline__ FastPrim<T> load_fast_prim(const T* s_w, int node, const Background& bg) {
  FastPrim<T> l;
  l.rho = s_w[0 * N3 + node]; l.v1 = s_w[1 * N3 + node]; l.v2 = s_w[2 * N3 + node];
  l.v3 = s_w[3 * N3 + node]; l.p = s_w[4 * N3 + node]; l.e = s_w[5 * N3 + node];
  l.phi = s_w[6 * N3 + node];
  if constexpr (VOL == 0 || NC == 1) l.lrho = s_w[7 * N3 + node];
  if constexpr (VOL == 0) l.lrp = s_w[8 * N3 + node];
  if constexpr (sizeof(T) == 4) {
    if (bg.on) {
      l.pp = s_w[9 * N3 + node]; l.rb = s_w[10 * N3 + node];
      l.rp = s_w[11 * N3 + node]; l.lam = s_w[12 * N3 + node];
    }
  }
  return l;
}

// after the volume terms (acc, acc_nc) and the lifted face fluxes (s_sf) are
// complete: surface integral, -1/J, sources, then du or the RK update
template <int N, typename T, bool KAHAN>
__device__ __forceinline__ void stage_epilogue(
    const Acc<T, KAHAN>* acc, const Acc<T, KAHAN>* acc_nc, const T* s_sf, int ti, int tj, int tk,
    size_t gnode, const double* __restrict__ u, double* __restrict__ out,
    double* __restrict__ kreg, const double* __restrict__ invJ, int mode, double rk_a,
    double rk_dt, double rk_b, double coriolis, const Background& bg, double* un) {
  constexpr int NV = 6, N2 = Shape<N>::N2;
  double du[NV];
  du[0] = double(acc[0].value());
#pragma unroll
  for (int v = 1; v < 5; ++v)
    du[v] = double(acc[v].value()) + 0.5 * double(acc_nc[v - 1].value());
  du[5] = 0.0;
  constexpr int NF = 6 * N2;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2;
    const int od = d == 0 ? ti : (d == 1 ? tj : tk);
    const bool on_face = (f % 2 == 0) ? (od == 0) : (od == N - 1);
    if (!on_face) continue;
    const int a = d == 0 ? tj : ti;
    const int b = d == 2 ? tj : tk;
    const int fid = f * N2 + a + N * b;
#pragma unroll
    for (int v = 0; v < NV; ++v) du[v] += double(s_sf[v * NF + fid]);
  }
  const double fac = -invJ[gnode];
#pragma unroll
  for (int v = 0; v < NV; ++v) du[v] *= fac;
  double uo[NV];  // read again rather than kept live through the volume phase
#pragma unroll
  for (int v = 0; v < NV; ++v) uo[v] = u[gnode * NV + v];
  if (bg.on) {  // tendency of the hydrostatic background, computed in FP64
    du[1] += double(bg.B[gnode * 3 + 0]);
    du[2] += double(bg.B[gnode * 3 + 1]);
    du[3] += double(bg.B[gnode * 3 + 2]);
  }
  // calc_sources!: source_terms_coriolis (-2 Omega e_z x rho v), after the Jacobian
  du[1] += coriolis * uo[2];
  du[2] += -coriolis * uo[1];
  // phi has no tendency: its state entry is the same in both state buffers
  // (copied when the state is set), so an RK stage leaves it untouched, and
  // the RK register holds the other five variables only ([5, nodes]: whole
  // cache lines, 16 B per node less traffic than a phi row kept at zero)
  un[NV - 1] = uo[NV - 1];
  if (mode == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[gnode * NV + v] = du[v];
  } else {
#pragma unroll
    for (int v = 0; v < NV - 1; ++v) {
      const double kn = rk_a * kreg[gnode * (NV - 1) + v] + rk_dt * du[v];
      kreg[gnode * (NV - 1) + v] = kn;
      un[v] = uo[v] + rk_b * kn;
      out[gnode * NV + v] = un[v];
    }
  }
}

// per element: the sum over directions of the maximum scaled speed of the
// new states un (the resident loop's next step size), reduced over the
// element's slot of BE lanes with speed_max (exact, so any order gives the
// max-speed reduction's value); every thread of the workgroup takes part
template <int BE, int EPB = 1>
__device__ __forceinline__ void element_speed(const E3Params& q, bool active, const double* un,
                                              const double* __restrict__ cv64,
                                              const double* __restrict__ invJ, size_t gnode,
                                              int t, int e0, int n_valid, double (*red)[3],
                                              double* partial) {
  constexpr int WPS = BE / 64;  // waves per element slot
  double l[3] = {0.0, 0.0, 0.0};
  if (active) node_scaled_speeds(q, un, cv64 + gnode * 9, fabs(invJ[gnode]), l);
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l[d] = speed_max(l[d], __shfl_xor(l[d], o, 64));
  if ((t & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) red[t >> 6][d] = l[d];
  }
  __syncthreads();
  const int es = t / BE;
  if (t - es * BE == 0 && es < n_valid) {
    const int w0 = es * WPS;
    double m[3] = {red[w0][0], red[w0][1], red[w0][2]};
    for (int w = 1; w < WPS; ++w)
      for (int d = 0; d < 3; ++d) m[d] = speed_max(m[d], red[w0 + w][d]);
    partial[e0 + es] = m[0] + m[1] + m[2];
  }
}

// Pair-record form: per direction the unordered pairs of every line are
// evaluated once into LDS records, then gathered by both nodes.
template <int N, int VOL, int NC, typename T, bool KAHAN>
__global__ __launch_bounds__(Shape<N>::BLOCK) void e3_stage(
    E3Params q, DsplitTable dst, double inv_w, int n_elements, const double* __restrict__ u,
    double* __restrict__ out, double* __restrict__ kreg, const T* __restrict__ cv,
    const double* __restrict__ invJ, const int* __restrict__ nbr,
    const int* __restrict__ face_info, int surf, int snc, int mode, double rk_a, double rk_dt,
    ResidentLink res, double rk_b, double coriolis, Background bg) {
  using S = Shape<N>;
  using C = Chunking<N, T, NC>;
  constexpr int N2 = S::N2, N3 = S::N3, P = S::P, B = S::BLOCK;
  constexpr bool MIXED = sizeof(T) == 4;
  constexpr int NW = NodeFields<VOL, NC, T>::NW;
  __shared__ T s_w[NW * N3];
  __shared__ double s_phi[MIXED ? N3 : 1];  // FP64 phi for the jumps of the mixed path
  __shared__ double s_pb[MIXED ? N3 : 1];   // FP64 background pressure
  __shared__ T s_ja[3][N3];
  __shared__ T s_fl[C::FLEN];  // pair terms [9][NPC], later face fluxes [6][6 n^2]
  __shared__ T s_ds[N * N];
  __shared__ unsigned char s_pa[P], s_pm[P];  // node pair (a, m), a < m, of each pair slot

  // the FP64 form never splits off a background: make that a compile-time fact
  if constexpr (!MIXED) bg.on = 0;
  const int e = xcd_element(blockIdx.x, n_elements);
  if (e >= n_elements) return;  // whole workgroup: padding of the XCD mapping
  const int t = threadIdx.x;
  const bool active = t < N3;
  const int ti = t % N, tj = (t / N) % N, tk = t / N2;
  const size_t gnode = (size_t)e * N3 + t;
  if (t < N * N) s_ds[t] = T(dst.ds[t]);
  if (t < P) {
    int qq = t, a = 0;
    while (qq >= N - 1 - a) { qq -= N - 1 - a; ++a; }
    s_pa[t] = (unsigned char)a;
    s_pm[t] = (unsigned char)(a + 1 + qq);
  }
  if (active) node_fields_to_lds<N, VOL, NC, T>(q, u, gnode, t, s_w, s_phi, s_pb, bg);

  FacePrefetch<N> pf;
  face_prefetch<N>(pf, e, 1, t, u, nbr, face_info);

  Acc<T, KAHAN> acc[5], acc_nc[4];
  for (int d = 0; d < 3; ++d) {
    __syncthreads();
    if (active) {
      s_ja[0][t] = cv[gnode * 9 + 3 * d + 0];
      s_ja[1][t] = cv[gnode * 9 + 3 * d + 1];
      s_ja[2][t] = cv[gnode * 9 + 3 * d + 2];
    }
    const int stride = d == 0 ? 1 : (d == 1 ? N : N2);
    const int own_d = d == 0 ? ti : (d == 1 ? tj : tk);
    const int own_line = d == 0 ? tj + N * tk : (d == 1 ? ti + N * tk : ti + N * tj);
    for (int c = 0; c < C::CHUNKS; ++c) {
      const int line0 = c * C::LINES;
      const int n_lines = (N2 - line0) < C::LINES ? (N2 - line0) : C::LINES;
      __syncthreads();
      for (int pid = t; pid < n_lines * P; pid += B) {
        const int lrel = pid / P;
        const int line = line0 + lrel;
        const int slot = pid - lrel * P;
        const int a = s_pa[slot], m = s_pm[slot];
        const int base = d == 0 ? N * line : (d == 1 ? (line % N) + N2 * (line / N) : line);
        const int na = base + a * stride, nm = base + m * stride;
        const FastPrim<T> l = load_fast_prim<VOL, NC, N3, T>(s_w, na, bg);
        const FastPrim<T> r = load_fast_prim<VOL, NC, N3, T>(s_w, nm, bg);
        T dphi;
        if constexpr (MIXED) dphi = T(s_phi[nm] - s_phi[na]);
        else dphi = r.phi - l.phi;
        const T n0 = T(0.5) * (s_ja[0][na] + s_ja[0][nm]);
        const T n1 = T(0.5) * (s_ja[1][na] + s_ja[1][nm]);
        const T n2 = T(0.5) * (s_ja[2][na] + s_ja[2][nm]);
        T f[5], g[5];
        pair_flux<VOL, NC, T>(q, l, r, n0, n1, n2, dphi, bg, f, g);
#pragma unroll
        for (int v = 0; v < 5; ++v) s_fl[v * C::NPC + pid] = f[v];
        if constexpr (NC != 0) {
          s_fl[5 * C::NPC + pid] = g[4];  // sm
          s_fl[6 * C::NPC + pid] = g[3];  // energy entry
        }
      }
      __syncthreads();
      if (active && own_line >= line0 && own_line < line0 + n_lines) {
        T jo[3] = {T(0), T(0), T(0)};
        if constexpr (NC != 0) {
          jo[0] = s_ja[0][t];
          jo[1] = s_ja[1][t];
          jo[2] = s_ja[2][t];
        }
        // the partners in rotated order own_d + 1, ..., own_d + n - 1 (mod n):
        // every lane gathers in each step
#pragma unroll 1
        for (int j = 1; j < N; ++j) {
          const int m = own_d + j >= N ? own_d + j - N : own_d + j;
          const int a = m < own_d ? m : own_d, b = m < own_d ? own_d : m;
          const int pid = (own_line - line0) * P + a * (2 * N - a - 1) / 2 + (b - a - 1);
          const T ds = s_ds[own_d * N + m];
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v].add(ds * s_fl[v * C::NPC + pid]);
          if constexpr (NC != 0) {
            const T dsg = own_d < m ? ds : -ds;
            const int pn = t + (m - own_d) * stride;  // the partner node
            const T sm = s_fl[5 * C::NPC + pid];
            acc_nc[0].add(dsg * ((T(0.5) * (jo[0] + s_ja[0][pn])) * sm));
            acc_nc[1].add(dsg * ((T(0.5) * (jo[1] + s_ja[1][pn])) * sm));
            acc_nc[2].add(dsg * ((T(0.5) * (jo[2] + s_ja[2][pn])) * sm));
            acc_nc[3].add(dsg * s_fl[6 * C::NPC + pid]);
          }
        }
      }
    }
  }
  FaceNormals<N, T> fn;
  face_normals<N, T, 1, T, false>(fn, e, 1, t, cv, nullptr, 0, pf, nbr, face_info);
  __syncthreads();  // pair terms consumed; the buffer now takes the face fluxes
  face_fluxes_to_lds<N, T>(q, e, 1, t, inv_w, u, pf, fn, nbr, face_info, surf, snc, bg, s_w, 0,
                           s_pb, s_fl, 0);
  __syncthreads();
  double un[6];
  if (active)
    stage_epilogue<N, T, KAHAN>(acc, acc_nc, s_fl, ti, tj, tk, gnode, u, out, kreg, invJ, mode,
                                rk_a, res.dt ? res.dt[0] : rk_dt, rk_b, coriolis, bg, un);
  if (res.speed_partial) {
    __shared__ double s_red[B / 64][3];
    element_speed<B>(q, active, un, res.cv64, invJ, gnode, t, e, 1, s_red, res.speed_partial);
  }
}

// Direct form: every node evaluates the two-point flux with each of its n-1
// partners on each line itself (each unordered pair is evaluated by both of
// its nodes).  The fluxes are symmetric and the gravity terms antisymmetric
// under exchanging the two states bit for bit, so the result equals the
// pair-record form's exactly; in exchange for twice the flux arithmetic the
// volume phase needs no pair buffer and no barrier, which leaves LDS for more
// resident workgroups.
// G: storage type of the contravariant vectors in memory (T, or float for the
// FP64 form with FP32 geometry storage: widened to T where they are read)
template <int N, int VOL, int NC, typename T, bool KAHAN, int EPB = 1, typename G = T>
__global__ __launch_bounds__(EPB * Shape<N>::BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void e3_stage_direct(
    E3Params q, DsplitTable dst, double inv_w, int n_elements, const double* __restrict__ u,
    double* __restrict__ out, double* __restrict__ kreg, const G* __restrict__ cv,
    const double* __restrict__ invJ, const int* __restrict__ nbr,
    const int* __restrict__ face_info, int surf, int snc, int mode, double rk_a, double rk_dt,
    ResidentLink res, double rk_b, double coriolis, Background bg) {
  using S = Shape<N>;
  constexpr int N2 = S::N2, N3 = S::N3, BE = S::BLOCK;
  constexpr bool MIXED = sizeof(T) == 4;
  constexpr int NW = NodeFields<VOL, NC, T>::NW;
  constexpr int BUF = 9 * N3 > 36 * N2 ? 9 * N3 : 36 * N2;
  // flux_kennedy_gruber (with or without a gravity term) takes the sum form
  // below, and so does the FP64 flux_ranocha (alone or with the Waruszewski
  // term)
  constexpr bool KG_SUMS = VOL == 1;
  constexpr bool RANOCHA_SUMS = VOL == 0 && !MIXED && NC != 2;
  // EPB element slots of BE lanes each; slot es owns s_w_all[es], s_buf_all[es]
  __shared__ T s_w_all[EPB][NW * N3];
  __shared__ double s_phi_all[EPB][MIXED ? N3 : 1];
  __shared__ double s_pb_all[EPB][MIXED ? N3 : 1];
  __shared__ T s_buf_all[EPB][BUF];  // contravariant vectors [9][n^3], later face fluxes [6][6 n^2]
  __shared__ T s_ds[N * N];

  if constexpr (!MIXED) bg.on = 0;
  const int n_groups = (n_elements + EPB - 1) / EPB;
  const int grp = xcd_element(blockIdx.x, n_groups);
  if (grp >= n_groups) return;  // whole workgroup: padding of the XCD mapping
  const int e0 = grp * EPB;
  const int n_valid = EPB == 1 ? 1 : (n_elements - e0 < EPB ? n_elements - e0 : EPB);
  const int t = threadIdx.x;
  // element slot and lane within it (one slot: the lane itself)
  const int es = EPB == 1 ? 0 : t / BE, tl = EPB == 1 ? t : t - es * BE;
  const int e = e0 + es;
  const bool active = tl < N3 && es < n_valid;
  const int ti = tl % N, tj = (tl / N) % N, tk = tl / N2;
  const size_t gnode = (size_t)e * N3 + tl;
  T* s_w = s_w_all[es];
  double* s_phi = s_phi_all[es];
  double* s_pb = s_pb_all[es];
  T* s_buf = s_buf_all[es];
  if (t < N * N) s_ds[t] = T(dst.ds[t]);
  if (active) {
    node_fields_to_lds<N, VOL, NC, T>(q, u, gnode, tl, s_w, s_phi, s_pb, bg);
#pragma unroll
    for (int c = 0; c < 9; ++c) s_buf[c * N3 + tl] = T(cv[gnode * 9 + c]);
  }
  FacePrefetch<N, EPB> pf;
  face_prefetch<N, EPB>(pf, e0, n_valid, t, u, nbr, face_info);
  __syncthreads();

  Acc<T, KAHAN> acc[5], acc_nc[4];
  if (active) {
    const FastPrim<T> l = load_fast_prim<VOL, NC, N3, T>(s_w, tl, bg);
    double phi_own = 0.0;
    if constexpr (MIXED) phi_own = s_phi[tl];
    for (int d = 0; d < 3; ++d) {
      const int stride = d == 0 ? 1 : (d == 1 ? N : N2);
      const int own_d = d == 0 ? ti : (d == 1 ? tj : tk);
      const T* ja = s_buf + 3 * d * N3;
      const T jo0 = ja[tl], jo1 = ja[N3 + tl], jo2 = ja[2 * N3 + tl];
      // per lane: the step from which the partner index wraps, and the
      // Dsplit row of the own node
      const int lim = N - own_d;
      const T* ds_row = s_ds + own_d * N + own_d;
      // the partners in cyclic order own_d + 1, ..., own_d + n - 1 (mod n):
      // every lane has a partner in each of the n - 1 iterations (a loop over
      // the line's n nodes that skips the own one leaves a lane in n idle in
      // every iteration)
#pragma unroll 1
      for (int j = 1; j < N; ++j) {
        const bool wrap = j >= lim;
        const int jw = wrap ? j - N : j;  // partner index minus own index
        const int pn = tl + jw * stride;  // the partner node
        const FastPrim<T> r = load_fast_prim<VOL, NC, N3, T>(s_w, pn, bg);
        T dphi;
        if constexpr (MIXED) dphi = T(s_phi[pn] - phi_own);
        else dphi = r.phi - l.phi;
        const T ds = ds_row[jw];
        if constexpr (KG_SUMS) {
          // flux_kennedy_gruber and the gravity term from sums of the two
          // nodes' values: every average's factor 1/2 is a power of two, so
          // accumulating 8 f_rho, 16 f_m, 16 f_E, 2 g_m and 4 g_E and scaling
          // once at the end rounds exactly as the averaged form
          const T rs = l.rho + r.rho, ps = l.p + r.p, es = l.e + r.e;
          const T v1s = l.v1 + r.v1, v2s = l.v2 + r.v2, v3s = l.v3 + r.v3;
          const T n0s = jo0 + ja[pn], n1s = jo1 + ja[N3 + pn], n2s = jo2 + ja[2 * N3 + pn];
          const T vns = v1s * n0s + v2s * n1s + v3s * n2s;
          const T F = rs * vns;
          // the momentum rows carry p' = p - p_bg when the background is split off
          T pms = ps;
          if constexpr (MIXED) {
            if (bg.on) pms = l.pp + r.pp;
          }
          const T ps4 = T(4) * pms;
          acc[0].add(ds * F);
          acc[1].add(ds * (F * v1s + ps4 * n0s));
          acc[2].add(ds * (F * v2s + ps4 * n1s));
          acc[3].add(ds * (F * v3s + ps4 * n2s));
          acc[4].add(ds * (F * es + (T(2) * ps) * vns));
          if constexpr (NC != 0) {
            const T ra = NC == 1 ? ln_mean_logs_branch(l.rho, r.rho, l.lrho, r.lrho) : T(0.5) * rs;
            const T sj = ra * dphi;
            T sm = sj;  // momentum part: (rho_avg - rho_avg of the background) [phi]
            if constexpr (MIXED) {
              if (bg.on) {
                const T Ap = r.rp - l.rp, sp = l.rp + r.rp;
                const T Ab = r.rb - l.rb, sb = l.rb + r.rb;
                const T dmean = (NC == 1) ? delta_ln_mean(Ap + Ab, sp + sb, Ab, sb, Ap, sp,
                                                          -dphi * T(bg.inv_RT0), r.lam - l.lam)
                                          : T(0.5) * sp;
                sm = dmean * dphi;
              }
            }
            acc_nc[0].add(ds * (n0s * sm));
            acc_nc[1].add(ds * (n1s * sm));
            acc_nc[2].add(ds * (n2s * sm));
            acc_nc[3].add(ds * (sj * vns));
          }
        } else if constexpr (RANOCHA_SUMS) {
          // flux_ranocha and the Waruszewski term with every factor 1/2 of
          // the averages (normals, velocities, pressure) left out and applied
          // once at the end: 4 f_rho, 8 f_m, 8 f_E, 2 g_m and 4 g_E, each the
          // averaged form's value times a power of two, so the sums round as
          // the averaged form's do
          const T n0s = jo0 + ja[pn], n1s = jo1 + ja[N3 + pn], n2s = jo2 + ja[2 * N3 + pn];
          const T vnls = l.v1 * n0s + l.v2 * n1s + l.v3 * n2s;
          const T vnrs = r.v1 * n0s + r.v2 * n1s + r.v3 * n2s;
          const T rho_mean = ln_mean_logs(l.rho, r.rho, l.lrho, r.lrho);
          const T inv_rp = l.p * r.p * inv_ln_mean_logs(l.rho * r.p, r.rho * l.p, l.lrp - r.lrp);
          const T v1s = l.v1 + r.v1, v2s = l.v2 + r.v2, v3s = l.v3 + r.v3;
          const T ps2 = T(2) * (l.p + r.p);
          const T vvs = l.v1 * r.v1 + l.v2 * r.v2 + l.v3 * r.v3;
          const T F = rho_mean * (vnls + vnrs);
          acc[0].add(ds * F);
          acc[1].add(ds * (F * v1s + ps2 * n0s));
          acc[2].add(ds * (F * v2s + ps2 * n1s));
          acc[3].add(ds * (F * v3s + ps2 * n2s));
          acc[4].add(ds * (F * (vvs + inv_rp * T(2.0 * q.inv_gm1)) +
                           T(2) * (l.p * vnrs + r.p * vnls)));
          if constexpr (NC == 1) {
            const T sj = rho_mean * dphi;
            acc_nc[0].add(ds * (n0s * sj));
            acc_nc[1].add(ds * (n1s * sj));
            acc_nc[2].add(ds * (n2s * sj));
            acc_nc[3].add(ds * (n0s * sj * v1s + n1s * sj * v2s + n2s * sj * v3s));
          }
        } else {
          const T n0 = T(0.5) * (jo0 + ja[pn]);
          const T n1 = T(0.5) * (jo1 + ja[N3 + pn]);
          const T n2 = T(0.5) * (jo2 + ja[2 * N3 + pn]);
          T f[5], g[5];
          pair_flux<VOL, NC, T>(q, l, r, n0, n1, n2, dphi, bg, f, g);
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v].add(ds * f[v]);
          if constexpr (NC != 0) {
            acc_nc[0].add(ds * g[0]);
            acc_nc[1].add(ds * g[1]);
            acc_nc[2].add(ds * g[2]);
            acc_nc[3].add(ds * g[3]);
          }
        }
      }
    }
    if constexpr (KG_SUMS) {
      constexpr T sc[5] = {T(0.125), T(0.0625), T(0.0625), T(0.0625), T(0.0625)};
      constexpr T scn[4] = {T(0.5), T(0.5), T(0.5), T(0.25)};
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v].scale(sc[v]);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc_nc[v].scale(scn[v]);
    } else if constexpr (RANOCHA_SUMS) {
      constexpr T sc[5] = {T(0.25), T(0.125), T(0.125), T(0.125), T(0.125)};
      constexpr T scn[4] = {T(0.5), T(0.5), T(0.5), T(0.25)};
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v].scale(sc[v]);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc_nc[v].scale(scn[v]);
    }
  }
  FaceNormals<N, T, EPB> fn;
  face_normals<N, T, EPB, G>(fn, e0, n_valid, t, cv, s_buf_all[0], BUF, pf, nbr, face_info);
  __syncthreads();  // contravariant vectors consumed; the buffer takes the face fluxes
  face_fluxes_to_lds<N, T, EPB>(q, e0, n_valid, t, inv_w, u, pf, fn, nbr, face_info, surf, snc, bg,
                                s_w_all[0], NW * N3, s_pb_all[0], s_buf_all[0], BUF);
  __syncthreads();
  double un[6];
  if (active)
    stage_epilogue<N, T, KAHAN>(acc, acc_nc, s_buf, ti, tj, tk, gnode, u, out, kreg, invJ, mode,
                                rk_a, res.dt ? res.dt[0] : rk_dt, rk_b, coriolis, bg, un);
  if (res.speed_partial) {
    __shared__ double s_red[EPB * BE / 64][3];
    element_speed<BE, EPB>(q, active, un, res.cv64, invJ, gnode, t, e0, n_valid, s_red,
                           res.speed_partial);
  }
}

// which volume form a shape runs, from timings on MI355X: the direct form for
// FP64 at every polydeg and for FP32 at polydeg 3 (6 % faster) and 7 (EC:
// 1.8x faster than the pair form); the pair form for FP32 at polydeg 4 (2 %
// faster: its background-split flux is costly enough that evaluating every
// pair twice loses)
template <int N, int VOL, typename T>
constexpr bool use_direct_form() {
  return sizeof(T) == 8 || N != 5;
}

// two elements per workgroup is built where it can pay: polydeg 3 and 4 of
// the direct form, whose 6 n^2 face nodes leave a partly filled last face
// pass (polydeg 4: 150 face nodes on 128 lanes; two elements put 300 on 256)
template <int N, typename T>
constexpr bool has_two_element_groups() {
  return (N == 4 || N == 5) && use_direct_form<N, 1, T>();
}

template <int N, int VOL, int NC, typename T, bool KAHAN, int EPB, typename G>
void launch_direct(const E3Launch& L, const DsplitTable& dst, hipStream_t s) {
  const int n_groups = (L.n_elements + EPB - 1) / EPB;
  const int grid = 8 * ((n_groups + 7) / 8);
  hipLaunchKernelGGL((e3_stage_direct<N, VOL, NC, T, KAHAN, EPB, G>), dim3(grid),
                     dim3(EPB * Shape<N>::BLOCK), 0, s, L.q, dst, L.inv_w, L.n_elements, L.u,
                     L.out, L.k, reinterpret_cast<const G*>(L.cv_t), L.invJ, L.nbr, L.face_info,
                     L.surface_flux, L.surface_noncons, L.mode, L.rk_a, L.rk_dt, L.res, L.rk_b,
                     L.coriolis, L.bg);
}

// G != T (FP32 geometry storage of the FP64 form) exists in the direct form only
template <int N, int VOL, int NC, typename T, bool KAHAN, typename G = T>
void launch_stage_inst(const E3Launch& L, hipStream_t s) {
  DsplitTable dst;
  for (int q = 0; q < N * N; ++q) dst.ds[q] = L.dsplit[q];
  const int grid = 8 * ((L.n_elements + 7) / 8);
  static_assert(std::is_same<G, T>::value || use_direct_form<N, VOL, T>(),
                "geometry storage other than T needs the direct form");
  if constexpr (use_direct_form<N, VOL, T>()) {
    if constexpr (has_two_element_groups<N, T>()) {
      if (L.epb == 2) {
        launch_direct<N, VOL, NC, T, KAHAN, 2, G>(L, dst, s);
        return;
      }
    }
    launch_direct<N, VOL, NC, T, KAHAN, 1, G>(L, dst, s);
  } else {
    hipLaunchKernelGGL((e3_stage<N, VOL, NC, T, KAHAN>), dim3(grid), dim3(Shape<N>::BLOCK), 0, s,
                       L.q, dst, L.inv_w, L.n_elements, L.u, L.out, L.k,
                       reinterpret_cast<const T*>(L.cv_t), L.invJ, L.nbr, L.face_info,
                       L.surface_flux, L.surface_noncons, L.mode, L.rk_a, L.rk_dt, L.res, L.rk_b,
                       L.coriolis, L.bg);
  }
}

template <typename T, bool KAHAN, int N, int VOL, typename G = T>
bool dispatch_nc(const E3Launch& L, hipStream_t s) {
  switch (L.volume_noncons) {
    case 0: launch_stage_inst<N, VOL, 0, T, KAHAN, G>(L, s); return true;
    case 1: launch_stage_inst<N, VOL, 1, T, KAHAN, G>(L, s); return true;
    case 2: launch_stage_inst<N, VOL, 2, T, KAHAN, G>(L, s); return true;
  }
  return false;
}

template <typename T, bool KAHAN, int N, typename G = T>
bool dispatch_vol(const E3Launch& L, hipStream_t s) {
  switch (L.volume_flux) {
    case 0: return dispatch_nc<T, KAHAN, N, 0, G>(L, s);
    case 1: return dispatch_nc<T, KAHAN, N, 1, G>(L, s);
    case 2: return dispatch_nc<T, KAHAN, N, 2, G>(L, s);
  }
  return false;
}

template <typename T, bool KAHAN, typename G = T>
bool dispatch_stage(const E3Launch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return dispatch_vol<T, KAHAN, 2, G>(L, s);
    case 3: return dispatch_vol<T, KAHAN, 3, G>(L, s);
    case 4: return dispatch_vol<T, KAHAN, 4, G>(L, s);
    case 5: return dispatch_vol<T, KAHAN, 5, G>(L, s);
    case 6: return dispatch_vol<T, KAHAN, 6, G>(L, s);
    case 7: return dispatch_vol<T, KAHAN, 7, G>(L, s);
    case 8: return dispatch_vol<T, KAHAN, 8, G>(L, s);
  }
  return false;
}

}  // namespace atmo

#endif
