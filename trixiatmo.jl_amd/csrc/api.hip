// C ABI of libatmo_hip (declared in include/atmo_hip.h): handle management,
// host <-> device transfers, the RK driver and the mesh generators.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atmo_hip.h"
#include "atmo_internal.hpp"
#include "euler3d_kernels.hpp"
#include "handle.hpp"

namespace atmo {
void build_cartesian(Mesh3D& m, const Basis& b, const int trees[3], const double xmin[3],
                     const double xmax[3], const int periodic[3]);
void build_cubed_sphere_shell(Mesh3D& m, const Basis& b, int trees_per_face, int layers,
                              double inner_radius, double thickness);
}  // namespace atmo

using namespace atmo;

namespace atmo {
std::string& global_error() {
  thread_local std::string e;
  return e;
}
}  // namespace atmo

struct atmo_mesh3d {
  Mesh3D m;
  bool metric_done = false;  // metric terms are computed on the first request
};

namespace {

inline int code_index(int code, int a, int b, int n) {
  switch (code) {
    case kBegin: return 0;
    case kEnd: return n - 1;
    case kIFwd: return a;
    case kIBwd: return n - 1 - a;
    case kJFwd: return b;
    default: return n - 1 - b;
  }
}

// direction 0..5 from Trixi node-index codes; -1 if malformed
int indices2direction(const int32_t* idx) {
  int dir = -1, nfix = 0, nsurf = 0;
  for (int d = 0; d < 3; ++d) {
    if (idx[d] == kBegin || idx[d] == kEnd) { dir = 2 * d + (idx[d] == kEnd); ++nfix; }
    else if (idx[d] >= kIFwd && idx[d] <= kJBwd) ++nsurf;
    else return -1;
  }
  return (nfix == 1 && nsurf == 2) ? dir : -1;
}

// own-face canonical surface index (a, b) of volume node (i, j, k) on face dir
inline void face_ab(int dir, int i, int j, int k, int* a, int* b) {
  const int d = dir / 2;
  *a = d == 0 ? j : i;
  *b = d == 2 ? j : k;
}



}  // namespace

namespace atmo {
void timer_begin(atmo_handle* h, int cls, TimerPair* tp) {
  tp->cls = cls;
  hipEventCreate(&tp->a);
  hipEventCreate(&tp->b);
  hipEventRecord(tp->a, h->stream);
}

void timer_end(atmo_handle* h, TimerPair* tp) {
  hipEventRecord(tp->b, h->stream);
  h->timers.push_back(*tp);
}
}  // namespace atmo

namespace {

void collect_timers(atmo_handle* h) {
  for (auto& tp : h->timers) {
    hipEventSynchronize(tp.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, tp.a, tp.b);
    h->launches[tp.cls] += 1;
    h->total_ms[tp.cls] += ms;
    hipEventDestroy(tp.a);
    hipEventDestroy(tp.b);
  }
  h->timers.clear();
}

E3Launch base_launch(atmo_handle* h) {
  E3Launch L{};
  const double gamma = h->cfg.c_p / h->cfg.c_v;
  L.q.gamma = gamma;
  L.q.gm1 = gamma - 1.0;
  L.q.inv_gm1 = 1.0 / (gamma - 1.0);
  L.q.lmars_c = h->cfg.lmars_c;
  L.n = h->n;
  L.n_elements = h->cfg.n_elements;
  L.volume_flux = h->cfg.volume_flux;
  L.volume_noncons = h->cfg.volume_noncons;
  L.surface_flux = h->cfg.surface_flux;
  L.surface_noncons = h->cfg.surface_noncons;
  L.cv = h->d_cv;
  L.cv_t = h->cfg.precision == ATMO_PREC_FP64 ? (const void*)h->d_cv : (const void*)h->d_cv32;
  L.invJ = h->d_invJ;
  L.dsplit = h->dsplit.data();
  L.weights = h->weights.data();
  L.inv_w = 1.0 / h->weights[0];
  L.nbr = h->d_nbr;
  L.face_info = h->d_face_info;
  L.coriolis = h->cfg.source_terms == ATMO_SRC_CORIOLIS ? 2.0 * h->cfg.rotation_rate : 0.0;
  L.bg.on = h->bg_on ? 1 : 0;
  L.bg.p0 = h->cfg.background_p0;
  L.bg.inv_RT0 = h->bg_on ? 1.0 / ((h->cfg.c_p - h->cfg.c_v) * h->cfg.background_T0) : 0.0;
  L.bg.B = h->d_B;
  L.epb = h->e3_epb;
  return L;
}

bool launch_stage_only(atmo_handle* h, const E3Launch& L) {
  const int prec = h->cfg.precision;
  return prec == ATMO_PREC_FP64            ? e3_launch_stage_fp64(L, h->stream)
         : prec == ATMO_PREC_FP32_GEOMETRY ? e3_launch_stage_fp64_geo32(L, h->stream)
         : prec == ATMO_PREC_FP32_KAHAN    ? e3_launch_stage_fp32_kahan(L, h->stream)
                                           : e3_launch_stage_fp32(L, h->stream);
}

int launch_stage(atmo_handle* h, const E3Launch& L) {
  TimerPair tp;
  if (h->timing) timer_begin(h, 0, &tp);
  const bool ok = launch_stage_only(h, L);
  if (h->timing) timer_end(h, &tp);
  if (!ok) return fail(&h->err, ATMO_ERR_INVALID, "no kernel for this flux/polydeg combination");
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

// FP64 tendency of the hydrostatic background built from the geopotential of
// u (device pointer), stored as FP32 momentum rows for the mixed kernels
int compute_background(atmo_handle* h, const double* u) {
  if (!h->bg_on) return ATMO_OK;
  E3Launch L = base_launch(h);
  e3_launch_background_state(L, u, h->d_ubg, h->stream);
  L.bg.on = 0;
  L.mode = 0;
  L.cv_t = h->d_cv;
  L.u = h->d_ubg;
  L.out = h->d_dubg;
  if (!e3_launch_stage_fp64(L, h->stream))
    return fail(&h->err, ATMO_ERR_INVALID, "no kernel for this flux/polydeg combination");
  e3_launch_extract_momentum(h->d_dubg, h->d_B, (long)h->n3 * h->cfg.n_elements, h->stream);
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

int run_rhs(atmo_handle* h, double* du, const double* u) {
  E3Launch L = base_launch(h);
  L.mode = 0;
  L.u = u;
  L.out = du;
  return launch_stage(h, L);
}

// the five stages of one CarpenterKennedy2N54 step on the resident state
// (Carpenter & Kennedy (1994) five-stage fourth-order 2N coefficients)
int launch_rk_step(atmo_handle* h, E3Launch& L, double* speed_partial = nullptr) {
  static const double A[5] = {0.0, -567301805773.0 / 1357537059087.0,
                              -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0,
                              -1275806237668.0 / 842570457699.0};
  static const double B[5] = {1432997174477.0 / 9575080441755.0,
                              5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0,
                              3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};
  for (int st = 0; st < 5; ++st) {
    L.u = h->d_u;
    L.out = h->d_u2;
    L.rk_a = A[st];
    L.rk_b = B[st];
    L.res.speed_partial = st == 4 ? speed_partial : nullptr;
    const int rc = launch_stage(h, L);
    if (rc != ATMO_OK) return rc;
    std::swap(h->d_u, h->d_u2);
  }
  return ATMO_OK;
}

// two RK steps (ten stage launches, after which the ping-pong buffers are
// back in place) captured once per state buffer, dt read from d_step_dt
int ensure_e3_graph(atmo_handle* h) {
  if (h->e3_graph && h->e3_graph_u == h->d_u && h->e3_graph_epb == h->e3_epb) return ATMO_OK;
  if (h->e3_graph) {
    HIP_TRY(h, hipGraphExecDestroy(h->e3_graph));
    h->e3_graph = nullptr;
  }
  if (!h->d_step_dt) HIP_TRY(h, hipMalloc(&h->d_step_dt, sizeof(double)));
  E3Launch L = base_launch(h);
  L.mode = 1;
  L.k = h->d_k;
  L.rk_dt = 0.0;
  L.res.dt = h->d_step_dt;
  hipGraph_t g = nullptr;
  HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_rk_step(h, L);
  if (rc == ATMO_OK) rc = launch_rk_step(h, L);
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (rc != ATMO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(h, ec);
  const hipError_t ei = hipGraphInstantiate(&h->e3_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_TRY(h, ei);
  h->e3_graph_u = h->d_u;
  h->e3_graph_epb = h->e3_epb;
  return ATMO_OK;
}

}  // namespace

namespace atmo {

// Autotuner of the 3D stage (no counterpart in the reference): where the
// stage has a two-element workgroup shape, times one and two elements per
// workgroup on the RK stage (the launch the time loops make) and keeps the
// faster.  The resident state is read; the scratch buffers d_u2 (the stage
// output) and d_du (as the register) are written.  Every shape gives the same
// bits (each element's arithmetic does not depend on its workgroup).
int e3_autotune(atmo_handle* h, double* ms_per_variant) {
  if (ms_per_variant) ms_per_variant[0] = ms_per_variant[1] = 0.0;
  if (!e3_has_two_element_groups(h->n, h->cfg.precision)) {
    h->e3_epb = 1;
    h->e3_tuned = true;
    return ATMO_OK;
  }
  hipEvent_t a, b;
  HIP_TRY(h, hipEventCreate(&a));
  HIP_TRY(h, hipEventCreate(&b));
  E3Launch L = base_launch(h);
  L.mode = 1;
  L.u = h->d_u;
  L.out = h->d_u2;
  L.k = h->d_du;
  L.rk_a = L.rk_dt = L.rk_b = 0.0;
  float best_of[2] = {1e30f, 1e30f};
  for (int round = 0; round < 3; ++round)
    for (int v = 0; v < 2; ++v) {
      L.epb = v + 1;
      if (!launch_stage_only(h, L)) continue;
      HIP_TRY(h, hipEventRecord(a, h->stream));
      for (int r = 0; r < 5; ++r) launch_stage_only(h, L);
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
  h->e3_epb = best_of[1] < best_of[0] ? 2 : 1;
  if (ms_per_variant) {
    ms_per_variant[0] = best_of[0];
    ms_per_variant[1] = best_of[1];
  }
  h->e3_tuned = true;
  h->du_valid = false;
  return ATMO_OK;
}

// per-element speeds of the state a resident loop starts from; every later
// step's come from its own last stage
int e3_resident_prologue(atmo_handle* h) {
  E3Launch L = base_launch(h);
  L.u = h->d_u;
  e3_launch_reduction(L, 0, nullptr, h->d_partial, h->d_result, h->stream);
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

// one step: the step size from the per-element speeds in d_partial, then the
// five stages, the last of which writes the new state's speeds there
int e3_enqueue_resident_step(atmo_handle* h) {
  launch_step_control_reduce(h->d_partial, h->cfg.n_elements, h->n, h->d_clock, h->d_clock + 1,
                             h->stream);
  HIP_TRY(h, hipGetLastError());
  E3Launch L = base_launch(h);
  L.mode = 1;
  L.k = h->d_k;
  L.rk_dt = 0.0;
  L.res.dt = h->d_clock + 1;
  L.res.cv64 = h->d_cv;
  return launch_rk_step(h, L, h->d_partial);
}

}  // namespace atmo

extern "C" {

int atmo_lgl_basis(int n, double* nodes, double* weights, double* D, double* Dsplit) {
  if (n < 2 || n > 16) return fail(nullptr, ATMO_ERR_INVALID, "n must be in 2..16");
  Basis b(n);
  if (nodes) std::memcpy(nodes, b.x.data(), sizeof(double) * n);
  if (weights) std::memcpy(weights, b.w.data(), sizeof(double) * n);
  if (D) std::memcpy(D, b.D.data(), sizeof(double) * n * n);
  if (Dsplit) std::memcpy(Dsplit, b.Dsplit.data(), sizeof(double) * n * n);
  return ATMO_OK;
}

int atmo_mesh3d_cartesian(int polydeg, const int32_t trees[3], const double xmin[3],
                          const double xmax[3], const int32_t periodic[3], atmo_mesh3d** out) {
  if (!out || !trees || !xmin || !xmax || !periodic)
    return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (polydeg < 1 || polydeg > 15) return fail(nullptr, ATMO_ERR_INVALID, "polydeg out of range");
  for (int c = 0; c < 3; ++c)
    if (trees[c] < 1 || !(xmax[c] > xmin[c]))
      return fail(nullptr, ATMO_ERR_INVALID, "bad brick dimensions");
  auto* mh = new atmo_mesh3d;
  mh->m.polydeg = polydeg;
  mh->m.n = polydeg + 1;
  Basis b(polydeg + 1);
  const int t[3] = {trees[0], trees[1], trees[2]};
  const int p[3] = {periodic[0], periodic[1], periodic[2]};
  build_cartesian(mh->m, b, t, xmin, xmax, p);
  std::string err;
  if (!build_connectivity(mh->m, &err)) {
    delete mh;
    return fail(nullptr, ATMO_ERR_MESH, err);
  }
  *out = mh;
  return ATMO_OK;
}

int atmo_mesh3d_cubed_sphere_shell(int polydeg, int trees_per_face_dimension, int layers,
                                   double inner_radius, double thickness, atmo_mesh3d** out) {
  if (!out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (polydeg < 1 || polydeg > 15 || trees_per_face_dimension < 1 || layers < 1 ||
      !(inner_radius > 0) || !(thickness > 0))
    return fail(nullptr, ATMO_ERR_INVALID, "bad shell parameters");
  auto* mh = new atmo_mesh3d;
  mh->m.polydeg = polydeg;
  mh->m.n = polydeg + 1;
  Basis b(polydeg + 1);
  build_cubed_sphere_shell(mh->m, b, trees_per_face_dimension, layers, inner_radius, thickness);
  std::string err;
  if (!build_connectivity(mh->m, &err)) {
    delete mh;
    return fail(nullptr, ATMO_ERR_MESH, err);
  }
  *out = mh;
  return ATMO_OK;
}

int atmo_mesh3d_sizes(const atmo_mesh3d* mh, int32_t* ne, int32_t* ni, int32_t* nb) {
  if (!mh) return fail(nullptr, ATMO_ERR_INVALID, "null mesh");
  if (ne) *ne = mh->m.n_elements;
  if (ni) *ni = mh->m.n_interfaces;
  if (nb) *nb = mh->m.n_boundaries;
  return ATMO_OK;
}

int atmo_mesh3d_arrays(const atmo_mesh3d* mh, const double** xyz, const double** cv,
                       const double** invJ, const int32_t** iids, const int32_t** iidx,
                       const int32_t** be, const int32_t** bidx, const int32_t** bk) {
  if (!mh) return fail(nullptr, ATMO_ERR_INVALID, "null mesh");
  if ((cv || invJ) && !mh->metric_done) {
    auto* mm = const_cast<atmo_mesh3d*>(mh);  // a cache filled on first use
    compute_metric_terms(mm->m, Basis(mm->m.n));
    mm->metric_done = true;
  }
  if (xyz) *xyz = mh->m.node_coordinates.data();
  if (cv) *cv = mh->m.contravariant.data();
  if (invJ) *invJ = mh->m.inverse_jacobian.data();
  if (iids) *iids = mh->m.interface_neighbor_ids.data();
  if (iidx) *iidx = mh->m.interface_node_indices.data();
  if (be) *be = mh->m.boundary_element.data();
  if (bidx) *bidx = mh->m.boundary_node_indices.data();
  if (bk) *bk = mh->m.boundary_kind.data();
  return ATMO_OK;
}

void atmo_mesh3d_destroy(atmo_mesh3d* mh) { delete mh; }

int atmo_metric_terms3d(int polydeg, int32_t ne, const double* xyz, double* cv, double* invJ) {
  if (polydeg < 1 || polydeg > 15 || ne < 0 || (ne > 0 && (!xyz || !cv || !invJ)))
    return fail(nullptr, ATMO_ERR_INVALID, "bad metric-term arguments");
  const int n = polydeg + 1;
  Basis b(n);
  std::vector<double> jac((size_t)9 * n * n * n * ne);
  compute_metric_terms(n, ne, xyz, b.D.data(), jac.data(), cv, invJ);
  return ATMO_OK;
}

int atmo_metric_terms3d_device(int polydeg, int32_t ne, const double* xyz, double* cv,
                               double* invJ) {
  if (polydeg < 1 || polydeg > 7 || ne < 0 || (ne > 0 && (!xyz || !cv || !invJ)))
    return fail(nullptr, ATMO_ERR_INVALID, "bad metric-term arguments (polydeg 1..7)");
  if (ne == 0) return ATMO_OK;
  Basis b(polydeg + 1);
  if (!e3_launch_metric_terms(polydeg + 1, ne, b.D.data(), xyz, cv, invJ, nullptr))
    return fail(nullptr, ATMO_ERR_INVALID, "polydeg out of range");
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return fail(nullptr, ATMO_ERR_DEVICE, hipGetErrorString(e));
  return ATMO_OK;
}

const char* atmo_last_error(const atmo_handle* h) {
  return h ? h->err.c_str() : global_error().c_str();
}

// device buffers of a set-up step, released on every return path
struct DeviceBuffers {
  std::vector<void*> ptrs;
  ~DeviceBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  void* get(size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes + 16) != hipSuccess) return nullptr;
    ptrs.push_back(q);
    return q;
  }
};

// atmo_create with n_interfaces < 0: the interface and boundary lists are
// built on the device from the node coordinates, copied to the host (the
// create path validates them and builds the per-element link table there)
// and handed to the ordinary path with each boundary's condition by kind
static int create_with_device_connectivity(const atmo_config* cfg, const atmo_geometry3d* g,
                                           atmo_handle** out) {
  if (!g->node_coordinates) return fail(nullptr, ATMO_ERR_INVALID, "missing node coordinates");
  const int n = cfg->polydeg + 1, E = cfg->n_elements;
  const size_t n3 = (size_t)n * n * n;
  DeviceBuffers db;
  auto* d_xyz = static_cast<double*>(db.get(sizeof(double) * 3 * n3 * E));
  auto* d_ids = static_cast<int32_t*>(db.get(sizeof(int32_t) * 2 * 3 * (size_t)E));
  auto* d_codes = static_cast<int32_t*>(db.get(sizeof(int32_t) * 6 * 3 * (size_t)E));
  auto* d_be = static_cast<int32_t*>(db.get(sizeof(int32_t) * 6 * (size_t)E));
  auto* d_bcodes = static_cast<int32_t*>(db.get(sizeof(int32_t) * 3 * 6 * (size_t)E));
  auto* d_kind = static_cast<int32_t*>(db.get(sizeof(int32_t) * 6 * (size_t)E));
  if (!d_xyz || !d_ids || !d_codes || !d_be || !d_bcodes || !d_kind)
    return fail(nullptr, ATMO_ERR_DEVICE, "device allocation failed");
  if (hipMemcpy(d_xyz, g->node_coordinates, sizeof(double) * 3 * n3 * E, hipMemcpyHostToDevice) !=
      hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "device upload failed");
  int32_t ni = 0, nb = 0;
  const int rc = atmo_mesh3d_connectivity_device(cfg->polydeg, E, d_xyz, g->period, d_ids,
                                                 d_codes, d_be, d_bcodes, d_kind, &ni, &nb);
  if (rc != ATMO_OK) return rc;
  std::vector<int32_t> ids(2 * (size_t)ni), codes(6 * (size_t)ni), be(nb), bcodes(3 * (size_t)nb),
      kind(nb), bc(nb);
  const bool ok =
      (ni == 0 || (hipMemcpy(ids.data(), d_ids, sizeof(int32_t) * ids.size(),
                             hipMemcpyDeviceToHost) == hipSuccess &&
                   hipMemcpy(codes.data(), d_codes, sizeof(int32_t) * codes.size(),
                             hipMemcpyDeviceToHost) == hipSuccess)) &&
      (nb == 0 || (hipMemcpy(be.data(), d_be, sizeof(int32_t) * nb, hipMemcpyDeviceToHost) ==
                       hipSuccess &&
                   hipMemcpy(bcodes.data(), d_bcodes, sizeof(int32_t) * bcodes.size(),
                             hipMemcpyDeviceToHost) == hipSuccess &&
                   hipMemcpy(kind.data(), d_kind, sizeof(int32_t) * nb, hipMemcpyDeviceToHost) ==
                       hipSuccess));
  if (!ok) return fail(nullptr, ATMO_ERR_DEVICE, "device download failed");
  for (int b = 0; b < nb; ++b) bc[b] = g->boundary_condition_by_kind[kind[b]];
  atmo_geometry3d gg = *g;
  gg.n_interfaces = ni;
  gg.interface_neighbor_ids = ids.data();
  gg.interface_node_indices = codes.data();
  gg.n_boundaries = nb;
  gg.boundary_element = be.data();
  gg.boundary_node_indices = bcodes.data();
  gg.boundary_condition = bc.data();
  return atmo_create(cfg, &gg, out);
}

int atmo_create(const atmo_config* cfg, const atmo_geometry3d* g, atmo_handle** out) {
  if (!cfg || !g || !out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (cfg->equations != ATMO_EQ_EULER_ENERGY_GRAVITY_3D)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported equations");
  if (cfg->polydeg < 1 || cfg->polydeg > 7)
    return fail(nullptr, ATMO_ERR_INVALID, "polydeg must be in 1..7");
  if (cfg->n_elements < 1) return fail(nullptr, ATMO_ERR_INVALID, "n_elements must be positive");
  if (cfg->volume_flux < 0 || cfg->volume_flux > 2)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported volume flux");
  if (cfg->volume_noncons < 0 || cfg->volume_noncons > 2 || cfg->surface_noncons < 0 ||
      cfg->surface_noncons > 2)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported nonconservative flux");
  if (cfg->surface_flux < 0 || cfg->surface_flux > 5)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported surface flux");
  if (cfg->precision < ATMO_PREC_FP64 || cfg->precision > ATMO_PREC_FP32_GEOMETRY)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported precision");
  if (!(cfg->c_p > cfg->c_v) || !(cfg->c_v > 0))
    return fail(nullptr, ATMO_ERR_INVALID, "need c_p > c_v > 0");
  if (cfg->surface_flux == ATMO_SURF_LMARS && !(cfg->lmars_c > 0))
    return fail(nullptr, ATMO_ERR_INVALID, "FluxLMARS needs a positive speed of sound");
  if (cfg->source_terms != ATMO_SRC_NONE && cfg->source_terms != ATMO_SRC_CORIOLIS)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported source terms");
  if (g->n_interfaces < 0) return create_with_device_connectivity(cfg, g, out);
  const bool metric_on_device = !g->contravariant_vectors && !g->inverse_jacobian;
  if ((!metric_on_device && (!g->contravariant_vectors || !g->inverse_jacobian)) ||
      (metric_on_device && !g->node_coordinates))
    return fail(nullptr, ATMO_ERR_INVALID, "missing geometry arrays");
  if (g->n_interfaces < 0 || g->n_boundaries < 0 ||
      (g->n_interfaces > 0 && (!g->interface_neighbor_ids || !g->interface_node_indices)) ||
      (g->n_boundaries > 0 &&
       (!g->boundary_element || !g->boundary_node_indices || !g->boundary_condition)))
    return fail(nullptr, ATMO_ERR_INVALID, "missing connectivity arrays");

  const int n = cfg->polydeg + 1, n2 = n * n, n3 = n2 * n, E = cfg->n_elements;
  // face-node gather map: neighbour global node, or -1 for a slip wall
  std::vector<int32_t> nbr((size_t)E * 6 * n2, INT32_MIN);
  std::vector<int32_t> face_info((size_t)E * 6, 0);
  auto set_link = [&](int e, int dir, int a, int b, int32_t v) -> bool {
    int32_t& slot = nbr[((size_t)e * 6 + dir) * n2 + a + n * b];
    if (slot != INT32_MIN) return false;
    slot = v;
    return true;
  };
  for (int I = 0; I < g->n_interfaces; ++I) {
    const int ep = g->interface_neighbor_ids[2 * I], es = g->interface_neighbor_ids[2 * I + 1];
    const int32_t* ip = g->interface_node_indices + 6 * I;
    const int32_t* is = ip + 3;
    const int dp = indices2direction(ip), ds = indices2direction(is);
    if (ep < 0 || ep >= E || es < 0 || es >= E || dp < 0 || ds < 0)
      return fail(nullptr, ATMO_ERR_INVALID, "malformed interface " + std::to_string(I));
    face_info[(size_t)es * 6 + ds] = 8 | dp;
    for (int bb = 0; bb < n; ++bb)
      for (int aa = 0; aa < n; ++aa) {
        const int pi = code_index(ip[0], aa, bb, n), pj = code_index(ip[1], aa, bb, n),
                  pk = code_index(ip[2], aa, bb, n);
        const int si = code_index(is[0], aa, bb, n), sj = code_index(is[1], aa, bb, n),
                  sk = code_index(is[2], aa, bb, n);
        int fa, fb;
        face_ab(dp, pi, pj, pk, &fa, &fb);
        const int32_t sglob = es * n3 + si + n * (sj + n * sk);
        const int32_t pglob = ep * n3 + pi + n * (pj + n * pk);
        if (!set_link(ep, dp, fa, fb, sglob))
          return fail(nullptr, ATMO_ERR_INVALID, "face used twice (interface " + std::to_string(I) + ")");
        face_ab(ds, si, sj, sk, &fa, &fb);
        if (!set_link(es, ds, fa, fb, pglob))
          return fail(nullptr, ATMO_ERR_INVALID, "face used twice (interface " + std::to_string(I) + ")");
      }
  }
  for (int B = 0; B < g->n_boundaries; ++B) {
    const int e = g->boundary_element[B];
    const int32_t* ib = g->boundary_node_indices + 3 * B;
    const int dir = indices2direction(ib);
    if (e < 0 || e >= E || dir < 0 || g->boundary_condition[B] != ATMO_BC_SLIP_WALL)
      return fail(nullptr, ATMO_ERR_INVALID, "malformed boundary " + std::to_string(B));
    for (int bb = 0; bb < n; ++bb)
      for (int aa = 0; aa < n; ++aa) {
        int fa, fb;
        face_ab(dir, code_index(ib[0], aa, bb, n), code_index(ib[1], aa, bb, n),
                code_index(ib[2], aa, bb, n), &fa, &fb);
        if (!set_link(e, dir, fa, fb, -1))
          return fail(nullptr, ATMO_ERR_INVALID, "face used twice (boundary " + std::to_string(B) + ")");
      }
  }
  for (auto v : nbr)
    if (v == INT32_MIN) return fail(nullptr, ATMO_ERR_INVALID, "element face without interface or boundary");

  auto* h = new atmo_handle;
  h->cfg = *cfg;
  h->n = n;
  h->n3 = n3;
  h->len = (int64_t)6 * n3 * E;
  // the hydrostatic background split serves the FP32 flux forms only
  h->bg_on = (cfg->precision == ATMO_PREC_FP32_KAHAN || cfg->precision == ATMO_PREC_FP32) &&
             cfg->background_p0 > 0 && cfg->background_T0 > 0;
  auto cleanup_fail = [&](int code) {
    atmo_destroy(h);
    return code;
  };
  if (hipGetDevice(&h->device) != hipSuccess) {
    fail(nullptr, ATMO_ERR_DEVICE, "no HIP device");
    return cleanup_fail(ATMO_ERR_DEVICE);
  }
  {
    Basis b(n);
    h->dsplit = b.Dsplit;
    h->weights = b.w;
  }
  const size_t sb = sizeof(double) * h->len;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&h->d_u, sb) != hipSuccess || hipMalloc(&h->d_u2, sb) != hipSuccess ||
      hipMalloc(&h->d_k, sb) != hipSuccess || hipMalloc(&h->d_du, sb) != hipSuccess ||
      hipMalloc(&h->d_uio, sb) != hipSuccess || hipMalloc(&h->d_duio, sb) != hipSuccess ||
      hipMalloc(&h->d_cv, sizeof(double) * 9 * n3 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_invJ, sizeof(double) * n3 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_nbr, sizeof(int32_t) * nbr.size()) != hipSuccess ||
      hipMalloc(&h->d_face_info, sizeof(int32_t) * face_info.size()) != hipSuccess ||
      hipMalloc(&h->d_partial, sizeof(double) * 9 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_result, sizeof(double) * 16) != hipSuccess ||
      (cfg->precision != ATMO_PREC_FP64 &&
       hipMalloc(&h->d_cv32, sizeof(float) * 9 * n3 * (size_t)E) != hipSuccess) ||
      (h->bg_on && (hipMalloc(&h->d_ubg, sb) != hipSuccess || hipMalloc(&h->d_dubg, sb) != hipSuccess ||
                    hipMalloc(&h->d_B, sizeof(float) * 3 * n3 * (size_t)E) != hipSuccess))) {
    fail(nullptr, ATMO_ERR_DEVICE, "device allocation failed");
    return cleanup_fail(ATMO_ERR_DEVICE);
  }
  if (metric_on_device) {
    // upload the node coordinates only and compute the metric terms here
    double* d_xyz = nullptr;
    Basis b(n);
    const bool ok = hipMalloc(&d_xyz, sizeof(double) * 3 * n3 * (size_t)E) == hipSuccess &&
                    hipMemcpy(d_xyz, g->node_coordinates, sizeof(double) * 3 * n3 * (size_t)E,
                              hipMemcpyHostToDevice) == hipSuccess &&
                    e3_launch_metric_terms(n, E, b.D.data(), d_xyz, h->d_cv, h->d_invJ, h->stream) &&
                    hipGetLastError() == hipSuccess;
    if (ok && h->d_cv32) launch_narrow_to_float(h->d_cv, h->d_cv32, (size_t)9 * n3 * E, h->stream);
    const bool done = ok && hipStreamSynchronize(h->stream) == hipSuccess;
    if (d_xyz) (void)hipFree(d_xyz);
    if (!done) {
      fail(nullptr, ATMO_ERR_DEVICE, "metric terms on the device failed");
      return cleanup_fail(ATMO_ERR_DEVICE);
    }
  } else if (hipMemcpy(h->d_cv, g->contravariant_vectors, sizeof(double) * 9 * n3 * (size_t)E,
                       hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(h->d_invJ, g->inverse_jacobian, sizeof(double) * n3 * (size_t)E,
                       hipMemcpyHostToDevice) != hipSuccess) {
    fail(nullptr, ATMO_ERR_DEVICE, "device upload failed");
    return cleanup_fail(ATMO_ERR_DEVICE);
  }
  if (hipMemcpy(h->d_nbr, nbr.data(), sizeof(int32_t) * nbr.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(h->d_face_info, face_info.data(), sizeof(int32_t) * face_info.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->d_u, 0, sb) != hipSuccess || hipMemset(h->d_u2, 0, sb) != hipSuccess ||
      hipMemset(h->d_k, 0, sb) != hipSuccess) {
    fail(nullptr, ATMO_ERR_DEVICE, "device upload failed");
    return cleanup_fail(ATMO_ERR_DEVICE);
  }
  if (h->d_cv32 && !metric_on_device) {
    std::vector<float> cv32((size_t)9 * n3 * E);
    for (size_t q = 0; q < cv32.size(); ++q) cv32[q] = (float)g->contravariant_vectors[q];
    if (hipMemcpy(h->d_cv32, cv32.data(), sizeof(float) * cv32.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
      fail(nullptr, ATMO_ERR_DEVICE, "device upload failed");
      return cleanup_fail(ATMO_ERR_DEVICE);
    }
  }
  *out = h;
  return ATMO_OK;
}

void atmo_destroy(atmo_handle* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto& tp : h->timers) {
    hipEventDestroy(tp.a);
    hipEventDestroy(tp.b);
  }
  double* bufs[] = {h->d_u, h->d_u2, h->d_k, h->d_du, h->d_uio, h->d_duio,
                    h->d_cv, h->d_invJ, h->d_partial, h->d_result};
  for (double* p : bufs)
    if (p) hipFree(p);
  if (h->d_nbr) hipFree(h->d_nbr);
  if (h->d_cv32) hipFree(h->d_cv32);
  if (h->d_ubg) hipFree(h->d_ubg);
  if (h->d_dubg) hipFree(h->d_dubg);
  if (h->d_B) hipFree(h->d_B);
  double* swe_bufs[] = {h->d_aux, h->d_xyz, h->d_dt_member, h->d_geo, h->d_fxf};
  for (double* p : swe_bufs)
    if (p) hipFree(p);
  if (h->d_geo32) hipFree(h->d_geo32);
  if (h->d_fxf32) hipFree(h->d_fxf32);
  if (h->d_hs64) hipFree(h->d_hs64);
  if (h->d_clock) hipFree(h->d_clock);
  if (h->d_step_dt) hipFree(h->d_step_dt);
  if (h->e3_graph) hipGraphExecDestroy(h->e3_graph);
  if (h->d_member) hipFree(h->d_member);
  if (h->d_member_first) hipFree(h->d_member_first);
  if (h->d_face_info) hipFree(h->d_face_info);
  if (h->swe_graph) hipGraphExecDestroy(h->swe_graph);
  rainy_free(h);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int64_t atmo_state_length(const atmo_handle* h) { return h ? h->len : -1; }

int atmo_rhs(atmo_handle* h, double* du, const double* u, double t) {
  if (!h || !du || !u) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  const size_t sb = sizeof(double) * h->len;
  HIP_TRY(h, hipMemcpyAsync(h->d_uio, u, sb, hipMemcpyHostToDevice, h->stream));
  int rc = ATMO_OK;
  if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) {
    rc = swe_rhs_device(h, h->d_duio, h->d_uio);
  } else if (h->eq == ATMO_EQ_RAINY_EULER_2D) {
    rc = rainy_rhs_device(h, h->d_duio, h->d_uio, t);
  } else {
    rc = compute_background(h, h->d_uio);
    if (rc != ATMO_OK) return rc;
    rc = run_rhs(h, h->d_duio, h->d_uio);
  }
  if (rc != ATMO_OK) return rc;
  HIP_TRY(h, hipMemcpyAsync(du, h->d_duio, sb, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

int atmo_rhs_device(atmo_handle* h, double* du, const double* u, double t) {
  if (!h || !du || !u) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) return swe_rhs_device(h, du, u);
  if (h->eq == ATMO_EQ_RAINY_EULER_2D) return rainy_rhs_device(h, du, u, t);
  int rc = compute_background(h, u);
  if (rc != ATMO_OK) return rc;
  return run_rhs(h, du, u);
}

int atmo_set_state(atmo_handle* h, const double* u) {
  if (!h || !u) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  HIP_TRY(h, hipMemcpyAsync(h->d_u, u, sizeof(double) * h->len, hipMemcpyHostToDevice, h->stream));
  // the RK stages of the 3D form never write phi: both state buffers hold it
  if (h->eq == ATMO_EQ_EULER_ENERGY_GRAVITY_3D)
    HIP_TRY(h, hipMemcpyAsync(h->d_u2, h->d_u, sizeof(double) * h->len, hipMemcpyDeviceToDevice,
                              h->stream));
  HIP_TRY(h, hipMemsetAsync(h->d_k, 0, sizeof(double) * h->len, h->stream));
  int rc = compute_background(h, h->d_u);
  if (rc != ATMO_OK) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->du_valid = false;
  return ATMO_OK;
}

int atmo_get_state(atmo_handle* h, double* u) {
  if (!h || !u) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  HIP_TRY(h, hipMemcpyAsync(u, h->d_u, sizeof(double) * h->len, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

int atmo_step(atmo_handle* h, double dt, int32_t n_steps) {
  if (!h || n_steps < 0 || !(dt > 0) || !std::isfinite(dt))
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad step arguments");
  if (h->eq == ATMO_EQ_RAINY_EULER_2D) return rainy_step(h, dt, n_steps);
  if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) {
    const int rc = swe_step(h, nullptr, dt, n_steps);
    if (rc != ATMO_OK) return rc;
    for (int s = 0; s < n_steps; ++s) h->time += dt;
    return ATMO_OK;
  }
  if (!h->e3_tuned) {
    const int rc = e3_autotune(h, nullptr);
    if (rc != ATMO_OK) return rc;
  }
  E3Launch L = base_launch(h);
  L.mode = 1;
  L.k = h->d_k;
  L.rk_dt = dt;
  int done = 0;
  if (!h->timing && n_steps >= 2) {
    // pairs of steps from a captured graph whose stages read dt from memory
    const int rc = ensure_e3_graph(h);
    if (rc != ATMO_OK) return rc;
    launch_set_scalar(h->d_step_dt, dt, h->stream);
    HIP_TRY(h, hipGetLastError());
    for (; done + 2 <= n_steps; done += 2) {
      HIP_TRY(h, hipGraphLaunch(h->e3_graph, h->stream));
      h->time += dt;
      h->time += dt;
    }
  }
  for (; done < n_steps; ++done) {
    const int rc = launch_rk_step(h, L);
    if (rc != ATMO_OK) return rc;
    h->time += dt;
  }
  h->du_valid = false;
  return ATMO_OK;
}

int atmo_max_dt(atmo_handle* h, double cfl, double* dt) {
  if (!h || !dt) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  if (h->eq == ATMO_EQ_RAINY_EULER_2D) return rainy_max_dt(h, cfl, dt);
  if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) {  // the smallest member time step
    std::vector<double> dtm(h->n_members);
    const int rc = swe_max_dt(h, cfl, dtm.data());
    if (rc != ATMO_OK) return rc;
    double m = dtm[0];
    for (double v : dtm) m = std::fmin(m, v);
    *dt = m;
    return ATMO_OK;
  }
  E3Launch L = base_launch(h);
  L.u = h->d_u;
  TimerPair tp;
  if (h->timing) timer_begin(h, 1, &tp);
  e3_launch_reduction(L, 0, nullptr, h->d_partial, h->d_result, h->stream);
  if (h->timing) timer_end(h, &tp);
  HIP_TRY(h, hipGetLastError());
  double mx = 0;
  HIP_TRY(h, hipMemcpyAsync(&mx, h->d_result, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (!std::isfinite(mx)) return fail(&h->err, ATMO_ERR_NONFINITE, "non-finite wave speed in state");
  if (mx < 5e-324) mx = 5e-324;
  *dt = cfl * (2.0 / (h->n * mx));  // max_dt (stepsize_dg2d.jl:43), then StepsizeCallback's cfl factor
  return ATMO_OK;
}

int atmo_integrals(atmo_handle* h, int32_t fresh_rhs, double* out) {
  if (!h || !out) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "null argument");
  if (h->eq == ATMO_EQ_COVARIANT_SWE_2D) return swe_integrals(h, fresh_rhs, out);
  if (h->eq == ATMO_EQ_RAINY_EULER_2D) return rainy_integrals(h, out);
  if (fresh_rhs) {
    int rc = run_rhs(h, h->d_du, h->d_u);
    if (rc != ATMO_OK) return rc;
    h->du_valid = true;
  }
  E3Launch L = base_launch(h);
  L.u = h->d_u;
  TimerPair tp;
  if (h->timing) timer_begin(h, 1, &tp);
  e3_launch_reduction(L, 1, h->du_valid ? h->d_du : nullptr, h->d_partial, h->d_result, h->stream);
  if (h->timing) timer_end(h, &tp);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(out, h->d_result, sizeof(double) * 9, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

int atmo_synchronize(atmo_handle* h) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

int atmo_enable_timing(atmo_handle* h, int32_t enable) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  collect_timers(h);
  h->timing = enable != 0;
  h->launches[0] = h->launches[1] = 0;
  h->total_ms[0] = h->total_ms[1] = 0;
  return ATMO_OK;
}

int atmo_kernel_times(atmo_handle* h, int32_t cls, int64_t* launches, double* total_ms) {
  if (!h || cls < 0 || cls > 1) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad argument");
  collect_timers(h);
  if (launches) *launches = h->launches[cls];
  if (total_ms) *total_ms = h->total_ms[cls];
  return ATMO_OK;
}

}  // extern "C"
