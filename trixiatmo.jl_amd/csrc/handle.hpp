// Handle of libatmo_hip (internal): device buffers, stream and timers.
#ifndef ATMO_HANDLE_HPP
#define ATMO_HANDLE_HPP

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/atmo_hip.h"
#include "rainy_kernels.hpp"
#include "swe_kernels.hpp"

namespace atmo {

struct TimerPair {
  hipEvent_t a, b;
  int cls;
};

std::string& global_error();

inline int fail(std::string* slot, int code, const std::string& msg) {
  if (slot) *slot = msg;
  global_error() = msg;
  return code;
}

}  // namespace atmo

struct atmo_handle {
  int eq = ATMO_EQ_EULER_ENERGY_GRAVITY_3D;
  atmo_config cfg;
  int n = 0, n3 = 0, nv = 6;
  int64_t len = 0;  // doubles per state
  int device = 0;
  hipStream_t stream = nullptr;
  double *d_u = nullptr, *d_u2 = nullptr, *d_k = nullptr, *d_du = nullptr;
  double *d_uio = nullptr, *d_duio = nullptr;  // staging for atmo_rhs with host arrays
  double *d_cv = nullptr, *d_invJ = nullptr;
  float* d_cv32 = nullptr;  // FP32 geometry of the mixed-precision path
  bool bg_on = false;       // hydrostatic background split (mixed path)
  double *d_ubg = nullptr, *d_dubg = nullptr;
  float* d_B = nullptr;
  std::vector<double> dsplit, weights;
  int* d_nbr = nullptr;
  int* d_face_info = nullptr;
  double *d_partial = nullptr, *d_result = nullptr;
  // covariant shallow water
  atmo_swe_config swe{};
  atmo::SweLaunch swe_base{};
  double *d_aux = nullptr, *d_xyz = nullptr, *d_dt_member = nullptr;
  double *d_geo = nullptr, *d_fxf = nullptr;  // stage-kernel geometry (compact, face transforms)
  float *d_geo32 = nullptr, *d_fxf32 = nullptr;  // the same in FP32 (ATMO_PREC_FP32_GEOMETRY)
  double* d_hs64 = nullptr;  // FP32 geometry with bottom topography: h_s kept in FP64
  int *d_member = nullptr, *d_member_first = nullptr;
  int n_members = 1;
  bool swe_tuned = false;                  // workgroup shape chosen by the autotuner
  hipGraphExec_t swe_graph = nullptr;      // two RK steps captured for replay
  int swe_graph_variant = -1;
  double* swe_graph_u = nullptr;           // state buffer the graph was captured with
  // rainy Euler
  atmo_rainy_config rainy{};
  atmo::RainyLaunch rainy_base{};
  double time = 0.0;                       // time of the resident state
  double* d_clock = nullptr;               // device clock of the resident time loop
  // 3D atmo_step: two RK steps captured for replay, their step size read from d_step_dt
  hipGraphExec_t e3_graph = nullptr;
  double* e3_graph_u = nullptr;            // state buffer the graph was captured with
  int e3_graph_epb = 0;                    // and the stage shape
  int e3_epb = 1;                          // elements per workgroup of the 3D stage
  bool e3_tuned = false;
  double* d_step_dt = nullptr;
  // Newton statistics: per-wave slots [n_stat_slots][4], then the reduced [4]
  unsigned long long* d_stats = nullptr;
  int n_stat_slots = 0;
  atmo_mlp* mlp = nullptr;                 // seed of the saturation solve
  bool du_valid = false;
  bool timing = false;
  std::vector<atmo::TimerPair> timers;
  int64_t launches[2] = {0, 0};
  double total_ms[2] = {0, 0};
  std::string err;
};

#define HIP_TRY(h, expr)                                                         \
  do {                                                                           \
    hipError_t e__ = (expr);                                                     \
    if (e__ != hipSuccess)                                                       \
      return atmo::fail((h) ? &(h)->err : nullptr, ATMO_ERR_DEVICE,              \
                        std::string(#expr) + ": " + hipGetErrorString(e__));     \
  } while (0)

namespace atmo {
void timer_begin(atmo_handle* h, int cls, TimerPair* tp);
void timer_end(atmo_handle* h, TimerPair* tp);
// covariant shallow water paths of the generic entry points
int swe_rhs_device(atmo_handle* h, double* du, const double* u);
int swe_step(atmo_handle* h, const double* dt_members, double dt, int n_steps);
int swe_max_dt(atmo_handle* h, double cfl, double* dt_members);
int swe_integrals(atmo_handle* h, int fresh_rhs, double* out);
int swe_autotune(atmo_handle* h, double* ms_per_variant);
// 3D stage: one or two elements per workgroup (ms_per_variant[2] may be NULL)
int e3_autotune(atmo_handle* h, double* ms_per_variant);
// whether the 3D stage of this polydeg / precision has the two-element shape
bool e3_has_two_element_groups(int n, int precision);
// rainy Euler paths of the generic entry points
int rainy_rhs_device(atmo_handle* h, double* du, const double* u, double t);
int rainy_step(atmo_handle* h, double dt, int n_steps);
int rainy_max_dt(atmo_handle* h, double cfl, double* dt);
int rainy_integrals(atmo_handle* h, double* out);
void rainy_free(atmo_handle* h);
// the resident time loop: one step (maximum wave speed, step-size control,
// five stages reading the step size from d_clock) enqueued on the stream
int e3_enqueue_resident_step(atmo_handle* h);
int swe_enqueue_resident_step(atmo_handle* h);
int rainy_enqueue_resident_step(atmo_handle* h);
// *ptr = value, stream-ordered (a value for kernels that read it from memory)
void launch_set_scalar(double* ptr, double value, hipStream_t s);
// dt[m] and the clock of n_members members from their maximum wave speeds
void launch_step_control(const double* max_speed, int n_members, int n, double* clock, double* dt,
                         hipStream_t s);
// the same for one member from per-element partial maxima
void launch_step_control_reduce(const double* partial, int n_items, int n, double* clock,
                                double* dt, hipStream_t s);
// set-up of a resident loop before its steps are enqueued (the first step's
// speeds where the stages provide the later ones)
int e3_resident_prologue(atmo_handle* h);
int swe_resident_prologue(atmo_handle* h);
int rainy_resident_prologue(atmo_handle* h);
// the same per ensemble member over its element range
void launch_step_control_members_reduce(const double* partial, const int* member_first,
                                        int n_members, int n, double* clock, double* dt,
                                        hipStream_t s);
}  // namespace atmo

#endif
