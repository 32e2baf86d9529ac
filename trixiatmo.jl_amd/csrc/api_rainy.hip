// C ABI of libatmo_hip for the rainy Euler equations on a uniform Cartesian
// StructuredMesh (handle creation, the rainy paths of the generic entry
// points, the saturation Newton solve) and for the MLP that seeds that solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "atmo_internal.hpp"
#include "handle.hpp"
#include "rainy_kernels.hpp"

using namespace atmo;

struct atmo_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *d_params = nullptr, *d_norm = nullptr, *d_m = nullptr, *d_v = nullptr;
  float *d_partial = nullptr, *d_grad = nullptr;
  int64_t partial_blocks = 0;
  int step = 0;
  std::string err;
};

namespace {

int launch_stage(atmo_handle* h, const RainyLaunch& L) {
  TimerPair tp;
  if (h->timing) timer_begin(h, 0, &tp);
  const bool ok = rainy_launch_stage(L, h->stream);
  if (h->timing) timer_end(h, &tp);
  if (!ok) return fail(&h->err, ATMO_ERR_INVALID, "no rainy kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

void mlp_free(atmo_mlp* m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  float* bufs[] = {m->d_params, m->d_norm, m->d_m, m->d_v, m->d_partial, m->d_grad};
  for (float* p : bufs)
    if (p) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

#define MLP_TRY(m, expr)                                                            \
  do {                                                                              \
    hipError_t e__ = (expr);                                                        \
    if (e__ != hipSuccess)                                                          \
      return atmo::fail(&(m)->err, ATMO_ERR_DEVICE,                                 \
                        std::string(#expr) + ": " + hipGetErrorString(e__));        \
  } while (0)

int mlp_ensure_partials(atmo_mlp* m, int64_t blocks) {
  if (blocks <= m->partial_blocks) return ATMO_OK;
  if (m->d_partial) MLP_TRY(m, hipFree(m->d_partial));
  m->d_partial = nullptr;
  MLP_TRY(m, hipMalloc(&m->d_partial, sizeof(float) * (kMlpParams + 1) * (size_t)blocks));
  m->partial_blocks = blocks;
  return ATMO_OK;
}

}  // namespace

namespace atmo {

int rainy_rhs_device(atmo_handle* h, double* du, const double* u, double t) {
  RainyLaunch L = h->rainy_base;
  L.mode = 0;
  L.u = u;
  L.out = du;
  L.t = t;
  return launch_stage(h, L);
}

namespace {

// Carpenter & Kennedy (1994) five-stage fourth-order 2N coefficients and the
// stage times c
const double kA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                      -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double kB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                      1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                      2277821191437.0 / 14882151754819.0};
const double kC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                      2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

}  // namespace

// per-element speeds of the state a resident loop starts from; every later
// step's come from its own last stage
int rainy_resident_prologue(atmo_handle* h) {
  RainyLaunch L = h->rainy_base;
  L.u = h->d_u;
  if (!rainy_launch_max_speed(L, h->d_partial, h->d_result, h->stream))
    return fail(&h->err, ATMO_ERR_INVALID, "no rainy kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  return ATMO_OK;
}

// one step: the step size from the per-element speeds in d_partial, then the
// five stages (each with its stage limiter), the last of which writes the
// limited new state's speeds there
int rainy_enqueue_resident_step(atmo_handle* h) {
  launch_step_control_reduce(h->d_partial, h->rainy_base.n_elements, h->n, h->d_clock,
                             h->d_clock + 1, h->stream);
  HIP_TRY(h, hipGetLastError());
  RainyLaunch L = h->rainy_base;
  L.mode = 1;
  L.k = h->d_k;
  L.clock = h->d_clock;
  for (int st = 0; st < 5; ++st) {
    L.u = h->d_u;
    L.out = h->d_u2;
    L.rk_a = kA[st];
    L.rk_b = kB[st];
    L.c_stage = kC[st];
    L.speed_partial = st == 4 ? h->d_partial : nullptr;
    const int rc = launch_stage(h, L);
    if (rc != ATMO_OK) return rc;
    std::swap(h->d_u, h->d_u2);
  }
  return ATMO_OK;
}

int rainy_step(atmo_handle* h, double dt, int n_steps) {
  const double* A = kA;
  const double* Bc = kB;
  const double* c = kC;
  RainyLaunch L = h->rainy_base;
  L.mode = 1;
  L.k = h->d_k;
  L.rk_dt = dt;
  for (int s = 0; s < n_steps; ++s) {
    for (int st = 0; st < 5; ++st) {
      L.u = h->d_u;
      L.out = h->d_u2;
      L.rk_a = A[st];
      L.rk_b = Bc[st];
      L.t = h->time + c[st] * dt;
      const int rc = launch_stage(h, L);
      if (rc != ATMO_OK) return rc;
      std::swap(h->d_u, h->d_u2);
    }
    h->time += dt;
  }
  h->du_valid = false;
  return ATMO_OK;
}

int rainy_max_dt(atmo_handle* h, double cfl, double* dt) {
  RainyLaunch L = h->rainy_base;
  L.u = h->d_u;
  TimerPair tp;
  if (h->timing) timer_begin(h, 1, &tp);
  const bool ok = rainy_launch_max_speed(L, h->d_partial, h->d_result, h->stream);
  if (h->timing) timer_end(h, &tp);
  if (!ok) return fail(&h->err, ATMO_ERR_INVALID, "no rainy kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  double mx = 0;
  HIP_TRY(h, hipMemcpyAsync(&mx, h->d_result, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (!std::isfinite(mx)) return fail(&h->err, ATMO_ERR_NONFINITE, "non-finite wave speed");
  *dt = cfl * (2.0 / (h->n * mx));  // max_dt (stepsize_dg2d.jl:43), then StepsizeCallback's cfl factor
  return ATMO_OK;
}

int rainy_integrals(atmo_handle* h, double* out) {
  RainyLaunch L = h->rainy_base;
  L.u = h->d_u;
  // d_dt_member holds the LGL weights on rainy handles
  if (!rainy_launch_integrals(L, h->d_dt_member, h->d_partial, h->d_result, h->stream))
    return fail(&h->err, ATMO_ERR_INVALID, "no rainy kernel for this polydeg");
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(out, h->d_result, sizeof(double) * 9, hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

void rainy_free(atmo_handle* h) {
  if (h->d_stats) hipFree(h->d_stats);
  h->d_stats = nullptr;
}

}  // namespace atmo

extern "C" {

int atmo_rainy_create(const atmo_rainy_config* cfg, atmo_handle** out) {
  if (!cfg || !out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  if (cfg->polydeg < 1 || cfg->polydeg > 7)
    return fail(nullptr, ATMO_ERR_INVALID, "polydeg must be in 1..7");
  if (cfg->cells_x < 1 || cfg->cells_y < 1)
    return fail(nullptr, ATMO_ERR_INVALID, "cells_per_dimension must be positive");
  if (!(cfg->x_max > cfg->x_min) || !(cfg->y_max > cfg->y_min))
    return fail(nullptr, ATMO_ERR_INVALID, "empty domain");
  if (cfg->volume_integral < 0 || cfg->volume_integral > 1 || cfg->surface_flux < 0 ||
      cfg->surface_flux > 1 || cfg->source_terms < 0 || cfg->source_terms > 3)
    return fail(nullptr, ATMO_ERR_INVALID, "unsupported flux or source choice");
  if (cfg->newton_max_iterations < 0 || cfg->newton_max_iterations > 1000 ||
      (cfg->newton_max_iterations > 0 && !(cfg->newton_tolerance > 0)))
    return fail(nullptr, ATMO_ERR_INVALID, "bad Newton settings");
  const int64_t E64 = (int64_t)cfg->cells_x * cfg->cells_y;
  if (E64 > (int64_t)1 << 28) return fail(nullptr, ATMO_ERR_INVALID, "mesh too large");
  const int n = cfg->polydeg + 1, n2 = n * n, E = (int)E64;

  auto* h = new atmo_handle;
  h->eq = ATMO_EQ_RAINY_EULER_2D;
  h->rainy = *cfg;
  h->n = n;
  h->n3 = n2;
  h->nv = 9;
  h->len = (int64_t)9 * n2 * E;
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
  // one statistics slot per wave: the stage launches at most one wave per
  // element, the Newton kernel four per 256 nodes
  {
    const size_t nodes = (size_t)n * n * E;
    h->n_stat_slots = (int)std::max<size_t>((size_t)E, 4 * ((nodes + 255) / 256));
  }
  const size_t stat_bytes = sizeof(unsigned long long) * (4 * (size_t)h->n_stat_slots + 4);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&h->d_u, sb) != hipSuccess || hipMalloc(&h->d_u2, sb) != hipSuccess ||
      hipMalloc(&h->d_k, sb) != hipSuccess || hipMalloc(&h->d_du, sb) != hipSuccess ||
      hipMalloc(&h->d_uio, sb) != hipSuccess || hipMalloc(&h->d_duio, sb) != hipSuccess ||
      hipMalloc(&h->d_dt_member, sizeof(double) * 16) != hipSuccess ||
      hipMalloc(&h->d_partial, sizeof(double) * 9 * (size_t)E) != hipSuccess ||
      hipMalloc(&h->d_result, sizeof(double) * 16) != hipSuccess ||
      hipMalloc(&h->d_stats, stat_bytes) != hipSuccess)
    return cleanup_fail(ATMO_ERR_DEVICE, "device allocation failed");
  if (hipMemcpy(h->d_dt_member, b.w.data(), sizeof(double) * n, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(h->d_u, 0, sb) != hipSuccess || hipMemset(h->d_k, 0, sb) != hipSuccess ||
      hipMemset(h->d_stats, 0, stat_bytes) != hipSuccess)
    return cleanup_fail(ATMO_ERR_DEVICE, "device upload failed");
  RainyLaunch& L = h->rainy_base;
  std::memset(&L, 0, sizeof(L));
  L.n = n;
  L.nx = cfg->cells_x;
  L.ny = cfg->cells_y;
  L.n_elements = E;
  L.periodic_y = cfg->periodic_y ? 1 : 0;
  L.volume_integral = cfg->volume_integral;
  L.surface_flux = cfg->surface_flux;
  L.source = cfg->source_terms;
  L.x0 = cfg->x_min;
  L.y0 = cfg->y_min;
  L.dx = (cfg->x_max - cfg->x_min) / cfg->cells_x;
  L.dy = (cfg->y_max - cfg->y_min) / cfg->cells_y;
  // uniform Cartesian cells: Ja^1 = (dy/2, 0), Ja^2 = (0, dx/2), J = dx dy / 4
  L.Ja11 = 0.5 * L.dy;
  L.Ja22 = 0.5 * L.dx;
  L.inv_J = 1.0 / (0.25 * L.dx * L.dy);
  for (int i = 0; i < n; ++i) L.xi[i] = b.x[i];
  L.inv_w0 = 1.0 / b.w[0];
  L.inv_wn = 1.0 / b.w[n - 1];
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m) {
      L.dsplit[i * n + m] = b.Dsplit[i * n + m];
      // Trixi's derivative_dhat: dhat[i, m] = -D[m, i] w_m / w_i
      L.dhat[i * n + m] = -b.D[m * n + i] * b.w[m] / b.w[i];
    }
  L.newton_tol = cfg->newton_tolerance;
  L.newton_max_iter = cfg->newton_max_iterations;
  L.stats = nullptr;
  *out = h;
  return ATMO_OK;
}

int atmo_set_time(atmo_handle* h, double t) {
  if (!h || !std::isfinite(t)) return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad time");
  h->time = t;
  return ATMO_OK;
}

int atmo_saturation_newton(atmo_handle* h, int32_t seeded, int32_t* iterations,
                           atmo_newton_stats* stats) {
  if (!h) return fail(nullptr, ATMO_ERR_INVALID, "null handle");
  if (h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(&h->err, ATMO_ERR_INVALID, "the saturation solve belongs to rainy handles");
  if (seeded < 0 || seeded > 2) return fail(&h->err, ATMO_ERR_INVALID, "seeded must be 0, 1 or 2");
  if (seeded == 1 && !h->mlp) return fail(&h->err, ATMO_ERR_INVALID, "no MLP attached");
  if (h->rainy.newton_max_iterations < 1)
    return fail(&h->err, ATMO_ERR_INVALID, "the handle has no Newton settings");
  RainyLaunch L = h->rainy_base;
  // counts accumulated by earlier stage limiters are discarded: the slots
  // then hold this solve's counts only
  unsigned long long* out = h->d_stats + 4 * (size_t)h->n_stat_slots;
  rainy_launch_stats_reduce(h->d_stats, h->n_stat_slots, out, h->stream);
  L.stats = h->d_stats;
  L.closed_form_seed = seeded == 2;
  MlpView mv{};
  if (seeded == 1) mv = MlpView{h->mlp->d_params, h->mlp->d_norm};
  TimerPair tp;
  if (h->timing) timer_begin(h, 1, &tp);
  rainy_launch_newton(L, h->d_u, seeded == 1 ? &mv : nullptr, iterations, h->stream);
  if (h->timing) timer_end(h, &tp);
  HIP_TRY(h, hipGetLastError());
  rainy_launch_stats_reduce(h->d_stats, h->n_stat_slots, out, h->stream);
  HIP_TRY(h, hipGetLastError());
  unsigned long long st[4];
  HIP_TRY(h, hipMemcpyAsync(st, out, sizeof(st), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (stats) {
    stats->nodes = h->len / 9;
    stats->updates = (int64_t)st[0];
    stats->failures = (int64_t)st[1];
    stats->max_updates = (int64_t)st[2];
    stats->fallbacks = (int64_t)st[3];
  }
  h->du_valid = false;
  return ATMO_OK;
}

int atmo_enable_newton_stats(atmo_handle* h, int32_t enable) {
  if (!h || h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "not a rainy handle");
  // clear the slots (the reduction clears what it reads)
  rainy_launch_stats_reduce(h->d_stats, h->n_stat_slots, h->d_stats + 4 * (size_t)h->n_stat_slots,
                            h->stream);
  HIP_TRY(h, hipGetLastError());
  h->rainy_base.stats = enable ? h->d_stats : nullptr;
  return ATMO_OK;
}

int atmo_newton_stats_read(atmo_handle* h, atmo_newton_stats* stats) {
  if (!h || !stats || h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "not a rainy handle");
  unsigned long long* out = h->d_stats + 4 * (size_t)h->n_stat_slots;
  rainy_launch_stats_reduce(h->d_stats, h->n_stat_slots, out, h->stream);
  HIP_TRY(h, hipGetLastError());
  unsigned long long st[4];
  HIP_TRY(h, hipMemcpyAsync(st, out, sizeof(st), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  stats->nodes = h->len / 9;
  stats->updates = (int64_t)st[0];
  stats->failures = (int64_t)st[1];
  stats->max_updates = (int64_t)st[2];
  stats->fallbacks = (int64_t)st[3];
  return ATMO_OK;
}

int atmo_rainy_features(atmo_handle* h, float* x, float* y) {
  if (!h || !x || !y || h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "bad arguments");
  rainy_launch_features(h->rainy_base, h->d_u, x, y, h->stream);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return ATMO_OK;
}

int atmo_rainy_attach_mlp(atmo_handle* h, atmo_mlp* m) {
  if (!h || h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "not a rainy handle");
  h->mlp = m;
  if (!m) h->rainy_base.mlp = MlpView{nullptr, nullptr};
  else if (h->rainy_base.mlp.params) h->rainy_base.mlp = MlpView{m->d_params, m->d_norm};
  return ATMO_OK;
}

int atmo_rainy_seed_stage_limiter(atmo_handle* h, int32_t enable) {
  if (!h || h->eq != ATMO_EQ_RAINY_EULER_2D)
    return fail(h ? &h->err : nullptr, ATMO_ERR_INVALID, "not a rainy handle");
  if (enable < 0 || enable > 2) return fail(&h->err, ATMO_ERR_INVALID, "enable must be 0, 1 or 2");
  if (enable == 1 && !h->mlp) return fail(&h->err, ATMO_ERR_INVALID, "no MLP attached");
  h->rainy_base.mlp = enable == 1 ? MlpView{h->mlp->d_params, h->mlp->d_norm} : MlpView{nullptr, nullptr};
  h->rainy_base.closed_form_seed = enable == 2;
  return ATMO_OK;
}

int atmo_mlp_create(uint64_t seed, atmo_mlp** out) {
  if (!out) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  auto* m = new atmo_mlp;
  auto bad = [&](const char* msg) {
    fail(nullptr, ATMO_ERR_DEVICE, msg);
    mlp_free(m);
    return ATMO_ERR_DEVICE;
  };
  if (hipGetDevice(&m->device) != hipSuccess) return bad("no HIP device");
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->d_params, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&m->d_norm, sizeof(float) * 10) != hipSuccess ||
      hipMalloc(&m->d_m, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&m->d_v, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&m->d_grad, sizeof(float) * (kMlpParams + 1)) != hipSuccess)
    return bad("device allocation failed");
  // Glorot-uniform weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> p(kMlpParams, 0.0f);
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> U(-a, a);
    for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = (float)U(rng);
  };
  fill(0, kMlpHidden, kMlpIn);
  fill(kMlpHidden * kMlpIn + kMlpHidden, kMlpHidden, kMlpHidden);
  fill(kMlpHidden * kMlpIn + 2 * kMlpHidden + kMlpHidden * kMlpHidden, kMlpOut, kMlpHidden);
  const float norm[10] = {0, 0, 0, 1, 1, 1, 0, 0, 1, 1};
  if (hipMemcpy(m->d_params, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(m->d_norm, norm, sizeof(norm), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMemset(m->d_v, 0, sizeof(float) * kMlpParams) != hipSuccess)
    return bad("device upload failed");
  *out = m;
  return ATMO_OK;
}

void atmo_mlp_destroy(atmo_mlp* m) { mlp_free(m); }

int atmo_mlp_set_normalization(atmo_mlp* m, const float* norm) {
  if (!m || !norm) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  for (int i = 3; i < 6; ++i)
    if (!(norm[i] > 0)) return fail(&m->err, ATMO_ERR_INVALID, "input scales must be positive");
  if (!(norm[8] > 0) || !(norm[9] > 0))
    return fail(&m->err, ATMO_ERR_INVALID, "output scales must be positive");
  MLP_TRY(m, hipMemcpy(m->d_norm, norm, sizeof(float) * 10, hipMemcpyHostToDevice));
  return ATMO_OK;
}

int atmo_mlp_set_params(atmo_mlp* m, const float* params) {
  if (!m || !params) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  MLP_TRY(m, hipMemcpy(m->d_params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  MLP_TRY(m, hipMemset(m->d_m, 0, sizeof(float) * kMlpParams));
  MLP_TRY(m, hipMemset(m->d_v, 0, sizeof(float) * kMlpParams));
  m->step = 0;
  return ATMO_OK;
}

int atmo_mlp_get_params(atmo_mlp* m, float* params) {
  if (!m || !params) return fail(nullptr, ATMO_ERR_INVALID, "null argument");
  MLP_TRY(m, hipStreamSynchronize(m->stream));
  MLP_TRY(m, hipMemcpy(params, m->d_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return ATMO_OK;
}

int atmo_mlp_loss_grad(atmo_mlp* m, const float* x, const float* y, int64_t n, float* loss,
                       float* grad) {
  if (!m || !x || !y || n < 1) return fail(m ? &m->err : nullptr, ATMO_ERR_INVALID, "bad arguments");
  const int64_t blocks = (n + 63) / 64;
  if (blocks > (int64_t)1 << 24) return fail(&m->err, ATMO_ERR_INVALID, "too many samples");
  int rc = mlp_ensure_partials(m, blocks);
  if (rc != ATMO_OK) return rc;
  const int nb = mlp_launch_loss_grad(m->d_params, m->d_norm, x, y, nullptr, 0, n, m->d_partial,
                                      m->stream);
  mlp_launch_reduce(m->d_partial, nb, m->d_grad, m->stream);
  MLP_TRY(m, hipGetLastError());
  std::vector<float> g(kMlpParams + 1);
  MLP_TRY(m, hipMemcpyAsync(g.data(), m->d_grad, sizeof(float) * (kMlpParams + 1),
                            hipMemcpyDeviceToHost, m->stream));
  MLP_TRY(m, hipStreamSynchronize(m->stream));
  if (grad) std::memcpy(grad, g.data(), sizeof(float) * kMlpParams);
  if (loss) *loss = g[kMlpParams];
  return ATMO_OK;
}

int atmo_mlp_train(atmo_mlp* m, const float* x, const float* y, const int32_t* perm, int64_t n,
                   int32_t batch, int32_t n_batches, float lr, float* last_loss) {
  if (!m || !x || !y || n < 1 || batch < 1 || n_batches < 0 || !(lr > 0))
    return fail(m ? &m->err : nullptr, ATMO_ERR_INVALID, "bad arguments");
  if ((int64_t)batch * n_batches > n)
    return fail(&m->err, ATMO_ERR_INVALID, "batches exceed the samples");
  int rc = mlp_ensure_partials(m, (batch + 63) / 64);
  if (rc != ATMO_OK) return rc;
  for (int b = 0; b < n_batches; ++b) {
    const int nb = mlp_launch_loss_grad(m->d_params, m->d_norm, x, y, perm, (int64_t)b * batch,
                                        batch, m->d_partial, m->stream);
    mlp_launch_reduce(m->d_partial, nb, m->d_grad, m->stream);
    mlp_launch_adam(m->d_params, m->d_m, m->d_v, m->d_grad, lr, 0.9f, 0.999f, 1e-8f, ++m->step,
                    m->stream);
  }
  MLP_TRY(m, hipGetLastError());
  float g = 0.0f;
  if (n_batches > 0)
    MLP_TRY(m, hipMemcpyAsync(&g, m->d_grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost,
                              m->stream));
  MLP_TRY(m, hipStreamSynchronize(m->stream));
  if (last_loss) *last_loss = g;
  return ATMO_OK;
}

int atmo_mlp_forward(atmo_mlp* m, const float* x, float* y, int64_t n) {
  if (!m || !x || !y || n < 1) return fail(m ? &m->err : nullptr, ATMO_ERR_INVALID, "bad arguments");
  mlp_launch_forward(m->d_params, m->d_norm, x, y, n, m->stream);
  MLP_TRY(m, hipGetLastError());
  MLP_TRY(m, hipStreamSynchronize(m->stream));
  return ATMO_OK;
}

}  // extern "C"
