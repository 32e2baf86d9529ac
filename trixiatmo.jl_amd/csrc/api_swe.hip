// C ABI of libatmo_hip for the covariant shallow-water form: the 2D cubed
// sphere mesh, handle creation and the covariant paths of the generic entry
// points (rhs!, RK steps, max_dt, integrals), including ensembles of
// independent members advanced by the same launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "atmo_internal.hpp"
#include "handle.hpp"
#include "swe_kernels.hpp"

using namespace atmo;

struct atmo_mesh2d {
  Mesh2D m;
};

namespace {

inline int code2d(int code, int i, int n) {
  switch (code) {
    case kBegin: return 0;
    case kEnd: return n - 1;
    case kIFwd: return i;
    default: return n - 1 - i;
  }
}

// face 0 (-x) .. 3 (+y) from 2D node-index codes; -1 if malformed
int dir2d(const int32_t* idx) {
  const bool f0 = idx[0] == kBegin || idx[0] == kEnd;
  const bool f1 = idx[1] == kBegin || idx[1] == kEnd;
  const bool s0 = idx[0] == kIFwd || idx[0] == kIBwd;
  const bool s1 = idx[1] == kIFwd || idx[1] == kIBwd;
  if (f0 && s1) return idx[0] == kBegin ? 0 : 1;
  if (f1 && s0) return idx[1] == kBegin ? 2 : 3;
  return -1;
}

int launch(atmo_handle* h, const SweLaunch& L) {
  TimerPair tp;
  if (h->timing) timer_begin(h, 0, &tp);
  const bool ok = swe_launch_stage(L, h->stream);
  if (h->timing) timer_end(h, &tp);
  if (!ok) return fail(&h->err, ATMO_ERR_INVALID, "no covariant kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

}  // namespace

namespace atmo {

int swe_rhs_device(atmo_handle* h, double* du, const double* u) {
  SweLaunch L = h->swe_base;
  L.mode = 0;
  L.u = u;
  L.out = du;
  return launch(h, L);
}

namespace {

const double kA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                      -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double kB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                      1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                      2277821191437.0 / 14882151754819.0};

// the 5 stage launches of n_steps CarpenterKennedy2N54 steps, time steps read
// from d_dt_member
int enqueue_steps(atmo_handle* h, int n_steps, double* speed_partial = nullptr) {
  SweLaunch L = h->swe_base;
  L.mode = 1;
  L.k = h->d_k;
  L.dt_member = h->d_dt_member;
  for (int s = 0; s < n_steps; ++s)
    for (int st = 0; st < 5; ++st) {
      L.u = h->d_u;
      L.out = h->d_u2;
      L.rk_a = kA[st];
      L.rk_b = kB[st];
      L.speed_partial = st == 4 ? speed_partial : nullptr;
      const int rc = launch(h, L);
      if (rc != ATMO_OK) return rc;
      std::swap(h->d_u, h->d_u2);
    }
  return ATMO_OK;
}

// capture two steps (ten launches; the ping-pong buffers come back to their
// places) into a graph for replay
int ensure_graph(atmo_handle* h) {
  // the graph holds the state buffers' addresses: an odd number of steps
  // taken outside it swaps them, and it is captured again
  if (h->swe_graph && h->swe_graph_variant == h->swe_base.variant && h->swe_graph_u == h->d_u)
    return ATMO_OK;
  if (h->swe_graph) {
    HIP_TRY(h, hipGraphExecDestroy(h->swe_graph));
    h->swe_graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_steps(h, 2);
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (rc != ATMO_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(h, ec);
  const hipError_t ei = hipGraphInstantiate(&h->swe_graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  HIP_TRY(h, ei);
  h->swe_graph_variant = h->swe_base.variant;
  h->swe_graph_u = h->d_u;
  return ATMO_OK;
}

}  // namespace

// per-element speeds of the state a resident loop starts from; every later
// step's come from its own last stage
int swe_resident_prologue(atmo_handle* h) {
  SweLaunch L = h->swe_base;
  L.u = h->d_u;
  if (!swe_launch_member_max_speed(L, h->d_partial, h->d_member_first, h->n_members, h->d_result,
                                   h->stream))
    return fail(&h->err, ATMO_ERR_INVALID, "no covariant kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

// one step: every member's step size from its elements' speeds in d_partial,
// then the five stages, the last of which writes the new speeds there
int swe_enqueue_resident_step(atmo_handle* h) {
  launch_step_control_members_reduce(h->d_partial, h->d_member_first, h->n_members, h->n,
                                     h->d_clock, h->d_dt_member, h->stream);
  HIP_TRY(h, hipGetLastError());
  return enqueue_steps(h, 1, h->d_partial);
}

int swe_autotune(atmo_handle* h, double* ms_per_variant) {
  hipEvent_t a, b;
  HIP_TRY(h, hipEventCreate(&a));
  HIP_TRY(h, hipEventCreate(&b));
  // the shapes are timed on the RK stage, the launch the time loops make
  // (its traffic differs from the rhs's: the register is read and written):
  // the state is read and the scratch buffers d_u2 (the stage output) and d_du
  // (as the register) are written; the resident state is left untouched
  SweLaunch L = h->swe_base;
  L.mode = 1;
  L.u = h->d_u;
  L.out = h->d_u2;
  L.k = h->d_du;
  L.rk_a = 0.0;
  L.rk_dt = 0.0;
  L.rk_b = 0.0;
  L.dt_member = nullptr;
  L.speed_partial = nullptr;
  // two rounds over all shapes, 5 timed launches each; a shape's time is its
  // best round (the rounds interleave the shapes, so clock and cache drift
  // affect them alike)
  float best_of[kSweVariants];
  for (int v = 0; v < kSweVariants; ++v) best_of[v] = 1e30f;
  for (int round = 0; round < 2; ++round)
    for (int v = 0; v < kSweVariants; ++v) {
      L.variant = v;
      if (!swe_launch_stage(L, h->stream)) continue;
      HIP_TRY(h, hipEventRecord(a, h->stream));
      for (int r = 0; r < 5; ++r) swe_launch_stage(L, h->stream);
      HIP_TRY(h, hipEventRecord(b, h->stream));
      HIP_TRY(h, hipEventSynchronize(b));
      float ms = 0.0f;
      HIP_TRY(h, hipEventElapsedTime(&ms, a, b));
      ms /= 5.0f;
      if (ms < best_of[v]) best_of[v] = ms;
    }
  hipEventDestroy(a);
  hipEventDestroy(b);
  HIP_TRY(h, hipGetLastError());
  int best = 0;
  for (int v = 0; v < kSweVariants; ++v) {
    if (ms_per_variant) ms_per_variant[v] = best_of[v];
    if (best_of[v] < best_of[best]) best = v;
  }
  h->swe_base.variant = best;
  h->swe_tuned = true;
  h->du_valid = false;
  return ATMO_OK;
}

int swe_step(atmo_handle* h, const double* dt_members, double dt, int n_steps) {
  std::vector<double> dts;
  if (!dt_members) {
    dts.assign(h->n_members, dt);
    dt_members = dts.data();
  }
  HIP_TRY(h, hipMemcpyAsync(h->d_dt_member, dt_members, sizeof(double) * h->n_members,
                            hipMemcpyHostToDevice, h->stream));
  if (!h->swe_tuned) {
    const int rc = swe_autotune(h, nullptr);
    if (rc != ATMO_OK) return rc;
  }
  int done = 0;
  if (!h->timing && n_steps >= 2) {
    int rc = ensure_graph(h);
    if (rc != ATMO_OK) return rc;
    for (; done + 2 <= n_steps; done += 2) HIP_TRY(h, hipGraphLaunch(h->swe_graph, h->stream));
  }
  if (done < n_steps) {
    const int rc = enqueue_steps(h, n_steps - done);
    if (rc != ATMO_OK) return rc;
  }
  // the host copy of the time steps may go out of scope: finish the upload
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->du_valid = false;
  return ATMO_OK;
}

int swe_max_dt(atmo_handle* h, double cfl, double* dt_members) {
  SweLaunch L = h->swe_base;
  L.u = h->d_u;
  TimerPair tp;
  if (h->timing) timer_begin(h, 1, &tp);
  const bool ok = swe_launch_member_max_speed(L, h->d_partial, h->d_member_first, h->n_members,
                                              h->d_result, h->stream);
  if (h->timing) timer_end(h, &tp);
  if (!ok) return fail(&h->err, ATMO_ERR_INVALID, "no covariant kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  std::vector<double> mx(h->n_members);
  HIP_TRY(h, hipMemcpyAsync(mx.data(), h->d_result, sizeof(double) * h->n_members,
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->n_members; ++m) {
    if (!std::isfinite(mx[m])) return fail(&h->err, ATMO_ERR_NONFINITE, "non-finite wave speed");
    dt_members[m] = cfl * (2.0 / (h->n * mx[m]));  // max_dt (stepsize_dg2d.jl:70), then the cfl factor
  }
  return ATMO_OK;
}

int swe_integrals(atmo_handle* h, int fresh_rhs, double* out) {
  if (fresh_rhs) {
    const int rc = swe_rhs_device(h, h->d_du, h->d_u);
    if (rc != ATMO_OK) return rc;
    h->du_valid = true;
  }
  SweLaunch L = h->swe_base;
  L.u = h->d_u;
  // LGL weights live after the dt_member array
  const double* w = h->d_dt_member + h->n_members;
  if (!swe_launch_integrals(L, h->du_valid ? h->d_du : nullptr, w, h->d_partial, h->d_result,
                            h->stream))
    return fail(&h->err, ATMO_ERR_INVALID, "no covariant kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(out, h->d_result, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

}  // namespace atmo

extern "C" {

int atmo_mesh2d_cubed_sphere(int polydeg, int trees_per_face_dimension, double radius,
                             int32_t christoffel_collocation, atmo_mesh2d** out) {
  if (!out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (polydeg < 1 || polydeg > 15 || trees_per_face_dimension < 1 || !(radius > 0))
    return fail(nullptr, ATMO_ERR_INVALID, "bad cubed-sphere parameters");
  auto* mh = new atmo_mesh2d;
  mh->m.polydeg = polydeg;
  mh->m.n = polydeg + 1;
  Basis b(polydeg + 1);
  std::string err;
  if (!build_cubed_sphere_2d(mh->m, b, trees_per_face_dimension, radius,
                             christoffel_collocation ? 1 : 0, &err)) {
    delete mh;
    return fail(nullptr, ATMO_ERR_MESH, err);
  }
  *out = mh;
  return ATMO_OK;
}

int atmo_mesh2d_sizes(const atmo_mesh2d* mh, int32_t* ne, int32_t* ni) {
  if (!mh) return fail(nullptr, ATMO_ERR_INVALID, "null mesh");
  if (ne) *ne = mh->m.n_elements;
  if (ni) *ni = mh->m.n_interfaces;
  return ATMO_OK;
}

int atmo_mesh2d_arrays(const atmo_mesh2d* mh, const double** xyz, const double** aux,
                       const int32_t** ids, const int32_t** idx) {
  if (!mh) return fail(nullptr, ATMO_ERR_INVALID, "null mesh");
  if (xyz) *xyz = mh->m.node_coordinates.data();
  if (aux) *aux = mh->m.aux.data();
  if (ids) *ids = mh->m.interface_neighbor_ids.data();
  if (idx) *idx = mh->m.interface_node_indices.data();
  return ATMO_OK;
}

void atmo_mesh2d_destroy(atmo_mesh2d* mh) { delete mh; }

int atmo_covariant_aux_device(int polydeg, int32_t n_elements, const double* node_coordinates,
                              int32_t christoffel_collocation, const double* bottom_topography,
                              double* aux_node_vars) {
  if (polydeg < 1 || polydeg > 7 || n_elements < 1 || !node_coordinates || !aux_node_vars)
    return fail(nullptr, ATMO_ERR_INVALID, "bad arguments");
  const int n = polydeg + 1;
  Basis b(n);
  if (!swe_launch_aux(n, n_elements, b.x.data(), b.D.data(), node_coordinates, bottom_topography,
                      christoffel_collocation, nullptr, nullptr, aux_node_vars, nullptr))
    return fail(nullptr, ATMO_ERR_INVALID, "no kernel for this polydeg");
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
  return ATMO_OK;
}

// atmo_swe_create with n_interfaces < 0: the interface lists of each member
// (or of the one mesh) built on the device from its node coordinates, with
// the member's first element added to the ids, then the ordinary path
static int swe_create_with_device_connectivity(const atmo_swe_config* cfg,
                                               const atmo_geometry2d* g, atmo_handle** out) {
  const int n = cfg->polydeg + 1, E = cfg->n_elements;
  const size_t n2 = (size_t)n * n;
  const int M = g->member_first_element ? g->n_members : 1;
  if (M < 1) return fail(nullptr, ATMO_ERR_INVALID, "n_members must be positive");
  std::vector<int32_t> first(M + 1);
  for (int m = 0; m <= M; ++m)
    first[m] = g->member_first_element ? g->member_first_element[m] : (m == 0 ? 0 : E);
  if (first[0] != 0 || first[M] != E)
    return fail(nullptr, ATMO_ERR_INVALID, "member ranges must cover all elements");
  for (int m = 0; m < M; ++m)
    if (first[m + 1] <= first[m]) return fail(nullptr, ATMO_ERR_INVALID, "empty member range");
  double* d_xyz = nullptr;
  int32_t *d_ids = nullptr, *d_codes = nullptr;
  auto release = [&]() {
    if (d_xyz) (void)hipFree(d_xyz);
    if (d_ids) (void)hipFree(d_ids);
    if (d_codes) (void)hipFree(d_codes);
  };
  if (hipMalloc(&d_xyz, sizeof(double) * 3 * n2 * E) != hipSuccess ||
      hipMalloc(&d_ids, sizeof(int32_t) * 2 * 2 * (size_t)E) != hipSuccess ||
      hipMalloc(&d_codes, sizeof(int32_t) * 4 * 2 * (size_t)E) != hipSuccess ||
      hipMemcpy(d_xyz, g->node_coordinates, sizeof(double) * 3 * n2 * E,
                hipMemcpyHostToDevice) != hipSuccess) {
    release();
    return fail(nullptr, ATMO_ERR_DEVICE, "device allocation or upload failed");
  }
  std::vector<int32_t> ids, codes;
  for (int m = 0; m < M; ++m) {
    const int Em = first[m + 1] - first[m];
    int32_t ni = 0;
    const int rc = atmo_mesh2d_connectivity_device(cfg->polydeg, Em, d_xyz + 3 * n2 * first[m],
                                                   d_ids, d_codes, &ni);
    if (rc != ATMO_OK) {
      release();
      return rc;
    }
    const size_t i0 = ids.size() / 2;
    ids.resize(2 * (i0 + ni));
    codes.resize(4 * (i0 + ni));
    if (ni > 0 &&
        (hipMemcpy(ids.data() + 2 * i0, d_ids, sizeof(int32_t) * 2 * ni, hipMemcpyDeviceToHost) !=
             hipSuccess ||
         hipMemcpy(codes.data() + 4 * i0, d_codes, sizeof(int32_t) * 4 * ni,
                   hipMemcpyDeviceToHost) != hipSuccess)) {
      release();
      return fail(nullptr, ATMO_ERR_DEVICE, "device download failed");
    }
    for (size_t q = 2 * i0; q < ids.size(); ++q) ids[q] += first[m];
  }
  release();
  atmo_geometry2d gg = *g;
  gg.n_interfaces = (int32_t)(ids.size() / 2);
  gg.interface_neighbor_ids = ids.data();
  gg.interface_node_indices = codes.data();
  return atmo_swe_create(cfg, &gg, out);
}

int atmo_swe_create(const atmo_swe_config* cfg, const atmo_geometry2d* g, atmo_handle** out) {
  if (!cfg || !g || !out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (cfg->polydeg < 1 || cfg->polydeg > 7)
    return fail(nullptr, ATMO_ERR_INVALID, "polydeg must be in 1..7");
  if (cfg->n_elements < 1) return fail(nullptr, ATMO_ERR_INVALID, "n_elements must be positive");
  // the standard form (weak form, its flux_lax_friedrichs and sources) and the
  // split form's fluxes do not mix
  const bool standard = cfg->volume_flux == ATMO_SWE_VOL_WEAK_FORM;
  if (cfg->volume_flux < 0 || cfg->volume_flux > 2 || cfg->surface_flux < 0 ||
      cfg->surface_flux > 3 || cfg->surface_noncons < 0 || cfg->surface_noncons > 1 ||
      cfg->source_terms < 0 || cfg->source_terms > 2 ||
      standard != (cfg->surface_flux == ATMO_SWE_SURF_LAX_FRIEDRICHS_STANDARD) ||
      (cfg->source_terms != 0 && standard != (cfg->source_terms == 2)))
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported flux or source choice");
  if (cfg->precision != ATMO_PREC_FP64 && cfg->precision != ATMO_PREC_FP32_GEOMETRY)
    return fail(nullptr, ATMO_ERR_INVALID,
                "the covariant form runs in FP64, with FP64 or FP32 geometry storage");
  const bool geo32 = cfg->precision == ATMO_PREC_FP32_GEOMETRY;
  if (!(cfg->gravity > 0)) return fail(nullptr, ATMO_ERR_INVALID, "gravity must be positive");
  if (g->node_coordinates && g->n_interfaces < 0)
    return swe_create_with_device_connectivity(cfg, g, out);
  if (!g->node_coordinates || g->n_interfaces < 0 ||
      (g->n_interfaces > 0 && (!g->interface_neighbor_ids || !g->interface_node_indices)))
    return fail(nullptr, ATMO_ERR_INVALID, "missing geometry arrays");
  // no auxiliary variables given: they are computed on the device from the
  // node coordinates (init_auxiliary_node_variables!), and so are the stage
  // kernel's compact rows and face maps
  const bool aux_on_device = g->aux_node_vars == nullptr;
  const int n = cfg->polydeg + 1, n2 = n * n, E = cfg->n_elements;
  const int M = g->member_first_element ? g->n_members : 1;
  std::vector<int32_t> first(M + 1), member(E);
  if (g->member_first_element) {
    if (M < 1) return fail(nullptr, ATMO_ERR_INVALID, "n_members must be positive");
    for (int m = 0; m <= M; ++m) first[m] = g->member_first_element[m];
    if (first[0] != 0 || first[M] != E)
      return fail(nullptr, ATMO_ERR_INVALID, "member ranges must cover all elements");
    for (int m = 0; m < M; ++m) {
      if (first[m + 1] <= first[m]) return fail(nullptr, ATMO_ERR_INVALID, "empty member range");
      for (int e = first[m]; e < first[m + 1]; ++e) member[e] = m;
    }
  } else {
    first[0] = 0;
    first[1] = E;
  }
  std::vector<int32_t> nbr((size_t)E * 4 * n, INT32_MIN);
  for (int I = 0; I < g->n_interfaces; ++I) {
    const int ep = g->interface_neighbor_ids[2 * I], es = g->interface_neighbor_ids[2 * I + 1];
    const int32_t* ip = g->interface_node_indices + 4 * I;
    const int32_t* is = ip + 2;
    const int dp = dir2d(ip), ds = dir2d(is);
    if (ep < 0 || ep >= E || es < 0 || es >= E || dp < 0 || ds < 0 ||
        member[ep] != member[es])
      return fail(nullptr, ATMO_ERR_INVALID, "malformed interface " + std::to_string(I));
    for (int k = 0; k < n; ++k) {
      const int pi = code2d(ip[0], k, n), pj = code2d(ip[1], k, n);
      const int si = code2d(is[0], k, n), sj = code2d(is[1], k, n);
      const int kp = dp < 2 ? pj : pi, ks = ds < 2 ? sj : si;
      int32_t& a = nbr[((size_t)ep * 4 + dp) * n + kp];
      int32_t& b = nbr[((size_t)es * 4 + ds) * n + ks];
      if (a != INT32_MIN || b != INT32_MIN)
        return fail(nullptr, ATMO_ERR_INVALID, "face used twice (interface " + std::to_string(I) + ")");
      a = es * n2 + si + n * sj;
      b = ep * n2 + pi + n * pj;
    }
  }
  for (auto v : nbr)
    if (v == INT32_MIN) return fail(nullptr, ATMO_ERR_INVALID, "element face without interface");

  // compact geometry of the stage kernel (structure of arrays, only what the
  // stage reads) and the per-face-node basis maps B_own A_nbr, in which the
  // neighbour's momentum enters (contravariant2global then global2contravariant,
  // covariant_shallow_water.jl:140-153, as one 2 x 2 product)
  const size_t NN = (size_t)n2 * E;
  const size_t geo_len = kSweGeo * NN, fxf_len = (size_t)16 * n * E;
  std::vector<double> geo, fxf;
  if (!aux_on_device) {
  geo.assign(geo_len, 0.0);
  fxf.assign(fxf_len, 0.0);
  for (size_t gq = 0; gq < NN; ++gq) {
    const double* a = g->aux_node_vars + gq * 26;
    for (int c = 0; c < 3; ++c) geo[c * NN + gq] = a[13 + c];
    geo[3 * NN + gq] = a[19];
    for (int c = 0; c < 6; ++c) geo[(4 + c) * NN + gq] = a[20 + c];
    const double* x = g->node_coordinates + gq * 3;
    geo[10 * NN + gq] = x[2] / std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    geo[11 * NN + gq] = a[12];
    // the kernels recompute J and G^-1 from G: the given ones must agree
    double J, c11, c12, c22;
    metric_from_gcov(a[13], a[14], a[15], J, c11, c12, c22);
    const double cs = std::fabs(c11) + std::fabs(c22);
    if (!(std::fabs(J - a[12]) <= 1e-10 * std::fabs(a[12])) ||
        !(std::fabs(c11 - a[16]) <= 1e-10 * cs) || !(std::fabs(c12 - a[17]) <= 1e-10 * cs) ||
        !(std::fabs(c22 - a[18]) <= 1e-10 * cs))
      return fail(nullptr, ATMO_ERR_INVALID,
                  "aux_node_vars: J or the inverse metric is not sqrt(det G) or G^-1 at node " +
                      std::to_string(gq));
  }
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 4; ++f)
      for (int k = 0; k < n; ++k) {
        const int i = f == 0 ? 0 : (f == 1 ? n - 1 : k), j = f == 2 ? 0 : (f == 3 ? n - 1 : k);
        const double* bo = g->aux_node_vars + ((size_t)e * n2 + i + n * j) * 26 + 6;  // B (2 x 3)
        const double* an = g->aux_node_vars + (size_t)nbr[((size_t)e * 4 + f) * n + k] * 26;  // A (3 x 2)
        double* M = &fxf[(((size_t)e * 4 + f) * n + k) * 4];
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c)
            M[2 * r + c] = bo[r] * an[3 * c] + bo[r + 2] * an[1 + 3 * c] + bo[r + 4] * an[2 + 3 * c];
      }
  }

  // without bottom topography (h_s = 0 at every node) the stage skips its row
  int topography = 0;
  if (!aux_on_device) {
    for (size_t gq = 0; gq < NN && !topography; ++gq) topography = geo[3 * NN + gq] != 0.0;
  } else if (g->bottom_topography) {
    for (size_t gq = 0; gq < NN && !topography; ++gq) topography = g->bottom_topography[gq] != 0.0;
  }

  // FP32 storage: each value rounded to nearest once, here.  h_s stays FP64
  // (a row of its own): the lake at rest is balanced by h + h_s = const with
  // h in FP64, which a rounded h_s would break at the FP32 rounding of h_s
  std::vector<float> geo32v, fxf32v;
  std::vector<double> hs64;
  if (geo32 && !aux_on_device) {
    geo32v.assign(geo.begin(), geo.end());
    fxf32v.assign(fxf.begin(), fxf.end());
    if (topography) hs64.assign(geo.begin() + 3 * NN, geo.begin() + 4 * NN);
  }
  const bool need_hs64 = geo32 && topography;

  auto* h = new atmo_handle;
  h->eq = ATMO_EQ_COVARIANT_SWE_2D;
  h->swe = *cfg;
  h->n = n;
  h->n3 = n2;
  h->nv = 3;
  h->len = (int64_t)3 * n2 * E;
  h->n_members = M;
  std::memset(&h->cfg, 0, sizeof(h->cfg));
  h->cfg.n_elements = E;
  h->cfg.polydeg = cfg->polydeg;
  auto cleanup_fail = [&](int code, const char* msg) {
    fail(nullptr, code, msg);
    atmo_destroy(h);
    return code;
  };
  if (hipGetDevice(&h->device) != hipSuccess) return cleanup_fail(ATMO_ERR_DEVICE, "no HIP device");
  Basis b(n);
  const size_t sb = sizeof(double) * h->len;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&h->d_u, sb) != hipSuccess || hipMalloc(&h->d_u2, sb) != hipSuccess ||
      hipMalloc(&h->d_k, sb) != hipSuccess || hipMalloc(&h->d_du, sb) != hipSuccess ||
      hipMalloc(&h->d_uio, sb) != hipSuccess || hipMalloc(&h->d_duio, sb) != hipSuccess ||
      hipMalloc(&h->d_aux, sizeof(double) * 26 * n2 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_xyz, sizeof(double) * 3 * n2 * (size_t)E) != hipSuccess ||
      ((!geo32 || aux_on_device) && hipMalloc(&h->d_geo, sizeof(double) * geo_len) != hipSuccess) ||
      ((!geo32 || aux_on_device) && hipMalloc(&h->d_fxf, sizeof(double) * fxf_len) != hipSuccess) ||
      (geo32 && hipMalloc(&h->d_geo32, sizeof(float) * geo_len) != hipSuccess) ||
      (geo32 && hipMalloc(&h->d_fxf32, sizeof(float) * fxf_len) != hipSuccess) ||
      (need_hs64 && hipMalloc(&h->d_hs64, sizeof(double) * NN) != hipSuccess) ||
      hipMalloc(&h->d_nbr, sizeof(int32_t) * nbr.size()) != hipSuccess ||
      hipMalloc(&h->d_member, sizeof(int32_t) * E) != hipSuccess ||
      hipMalloc(&h->d_member_first, sizeof(int32_t) * (M + 1)) != hipSuccess ||
      hipMalloc(&h->d_dt_member, sizeof(double) * (M + n)) != hipSuccess ||
      hipMalloc(&h->d_partial, sizeof(double) * 3 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_result, sizeof(double) * (M > 16 ? M : 16)) != hipSuccess)
    return cleanup_fail(ATMO_ERR_DEVICE, "device allocation failed");
  if ((!aux_on_device && hipMemcpy(h->d_aux, g->aux_node_vars, sizeof(double) * 26 * n2 * (size_t)E,
                                   hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(h->d_xyz, g->node_coordinates, sizeof(double) * 3 * n2 * (size_t)E,
                hipMemcpyHostToDevice) != hipSuccess ||
      (!aux_on_device && !geo32 &&
       (hipMemcpy(h->d_geo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(h->d_fxf, fxf.data(), sizeof(double) * fxf.size(), hipMemcpyHostToDevice) !=
            hipSuccess)) ||
      (!aux_on_device && geo32 &&
       (hipMemcpy(h->d_geo32, geo32v.data(), sizeof(float) * geo32v.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_fxf32, fxf32v.data(), sizeof(float) * fxf32v.size(),
                  hipMemcpyHostToDevice) != hipSuccess)) ||
      (!aux_on_device && need_hs64 &&
       hipMemcpy(h->d_hs64, hs64.data(), sizeof(double) * hs64.size(), hipMemcpyHostToDevice) !=
           hipSuccess) ||
      hipMemcpy(h->d_nbr, nbr.data(), sizeof(int32_t) * nbr.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(h->d_member, member.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(h->d_member_first, first.data(), sizeof(int32_t) * (M + 1),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_dt_member + M, b.w.data(), sizeof(double) * n, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(h->d_u, 0, sb) != hipSuccess || hipMemset(h->d_k, 0, sb) != hipSuccess)
    return cleanup_fail(ATMO_ERR_DEVICE, "device upload failed");
  if (aux_on_device) {
    // the auxiliary variables, the compact rows and the face maps from the
    // node coordinates, in FP64; FP32 storage narrows them once, on the device
    // (after every upload above has landed: the handle's stream does not wait
    // for the null stream)
    if (hipDeviceSynchronize() != hipSuccess)
      return cleanup_fail(ATMO_ERR_DEVICE, "device upload failed");
    double* d_topo = nullptr;
    if (topography) {
      if (hipMalloc(&d_topo, sizeof(double) * NN) != hipSuccess)
        return cleanup_fail(ATMO_ERR_DEVICE, "device allocation failed");
      if (hipMemcpy(d_topo, g->bottom_topography, sizeof(double) * NN, hipMemcpyHostToDevice) !=
          hipSuccess) {
        (void)hipFree(d_topo);
        return cleanup_fail(ATMO_ERR_DEVICE, "device upload failed");
      }
    }
    bool ok = swe_launch_aux(n, E, b.x.data(), b.D.data(), h->d_xyz, d_topo,
                             g->christoffel_collocation, h->d_member, h->d_member_first,
                             h->d_aux, h->stream) &&
              swe_launch_face_maps(n, E, h->d_aux, h->d_nbr, h->d_fxf, h->stream);
    if (ok) {
      swe_launch_compact_rows(h->d_aux, h->d_xyz, (long long)NN, h->d_geo, h->stream);
      if (geo32) {
        launch_narrow_to_float(h->d_geo, h->d_geo32, geo_len, h->stream);
        launch_narrow_to_float(h->d_fxf, h->d_fxf32, fxf_len, h->stream);
        if (need_hs64)
          ok = hipMemcpyAsync(h->d_hs64, h->d_geo + 3 * NN, sizeof(double) * NN,
                              hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
      }
    }
    ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(h->stream) == hipSuccess;
    if (d_topo) (void)hipFree(d_topo);
    if (geo32) {  // the FP64 copies served only to narrow from
      (void)hipFree(h->d_geo);
      (void)hipFree(h->d_fxf);
      h->d_geo = nullptr;
      h->d_fxf = nullptr;
    }
    if (!ok) return cleanup_fail(ATMO_ERR_DEVICE, "auxiliary variables on the device failed");
  }
  SweLaunch& L = h->swe_base;
  L.n = n;
  L.n_elements = E;
  L.volume_flux = cfg->volume_flux;
  L.surface_flux = cfg->surface_flux;
  L.surface_noncons = cfg->surface_noncons;
  L.source_terms = cfg->source_terms;
  L.gravity = cfg->gravity;
  L.rotation_rate = cfg->rotation_rate;
  for (int q = 0; q < n * n; ++q) L.dsplit[q] = b.Dsplit[q];
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m)  // derivative_hat[i, m] = -D[m, i] w_m / w_i
      L.dhat[i * n + m] = -b.D[m * n + i] * b.w[m] / b.w[i];
  L.inv_w = 1.0 / b.w[0];
  L.aux = h->d_aux;
  L.xyz = h->d_xyz;
  L.geo = h->d_geo;
  L.fxf = h->d_fxf;
  L.geo32 = h->d_geo32;
  L.fxf32 = h->d_fxf32;
  L.hs64 = h->d_hs64;
  L.n_nodes = (long long)NN;
  L.nbr = h->d_nbr;
  L.member = h->d_member;
  L.dt_member = nullptr;
  L.topography = topography;
  L.variant = 0;
  *out = h;
  return ATMO_OK;
}

int atmo_autotune(atmo_handle* h, int32_t* variant, double* ms) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  if (h->eq == ATMO_EQ_EULER_ENERGY_GRAVITY_3D) {
    const int rc = e3_autotune(h, ms);
    if (rc != ATMO_OK) return rc;
    if (variant) *variant = h->e3_epb - 1;
    return ATMO_OK;
  }
  if (h->eq != ATMO_EQ_COVARIANT_SWE_2D) {
    if (variant) *variant = 0;
    return ATMO_OK;  // the rainy stage kernel has one shape
  }
  const int rc = swe_autotune(h, ms);
  if (rc != ATMO_OK) return rc;
  if (variant) *variant = h->swe_base.variant;
  return ATMO_OK;
}

int atmo_set_kernel_variant(atmo_handle* h, int32_t variant) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  if (h->eq == ATMO_EQ_EULER_ENERGY_GRAVITY_3D) {
    if (variant < 0 || variant > 1 ||
        (variant == 1 && !e3_has_two_element_groups(h->n, h->cfg.precision)))
      return fail(&h->err, ATMO_ERR_INVALID, "no such kernel variant");
    h->e3_epb = variant + 1;
    h->e3_tuned = true;
    return ATMO_OK;
  }
  if (h->eq != ATMO_EQ_COVARIANT_SWE_2D)
    return variant == 0 ? ATMO_OK : fail(&h->err, ATMO_ERR_INVALID, "one shape only");
  if (variant < 0 || variant >= kSweVariants)
    return fail(&h->err, ATMO_ERR_INVALID, "no such kernel variant");
  h->swe_base.variant = variant;
  h->swe_tuned = true;
  return ATMO_OK;
}

int atmo_max_dt_members(atmo_handle* h, double cfl, double* dt) {
  if (!h || !dt) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  if (h->eq != ATMO_EQ_COVARIANT_SWE_2D)
    return fail(&h->err, ATMO_ERR_INVALID, "ensembles are covariant shallow-water handles");
  return swe_max_dt(h, cfl, dt);
}

int atmo_step_members(atmo_handle* h, const double* dt, int32_t n_steps) {
  if (!h || !dt || n_steps < 0) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad argument");
  if (h->eq != ATMO_EQ_COVARIANT_SWE_2D)
    return fail(&h->err, ATMO_ERR_INVALID, "ensembles are covariant shallow-water handles");
  for (int m = 0; m < h->n_members; ++m)
    if (!(dt[m] > 0) || !std::isfinite(dt[m]))
      return fail(&h->err, ATMO_ERR_INVALID, "bad member time step");
  return swe_step(h, dt, 0.0, n_steps);
}

}  // extern "C"
