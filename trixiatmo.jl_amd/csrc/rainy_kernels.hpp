// Launch interface of the rainy-Euler kernels (internal).
#ifndef ATMO_RAINY_KERNELS_HPP
#define ATMO_RAINY_KERNELS_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

namespace atmo {

// small MLP 3 -> 32 -> 32 -> 2 (tanh) that seeds the saturation solve; the
// parameters are packed W1[32][3], b1[32], W2[32][32], b2[32], W3[2][32], b3[2]
constexpr int kMlpIn = 3, kMlpHidden = 32, kMlpOut = 2;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpOut * kMlpHidden + kMlpOut;  // 1250

// tanh of the network's hidden layers: 1 - 2 / (1 + e^(2x)) from the
// hardware base-2 exponential and reciprocal (absolute error ~1e-7, ample for
// a seed; the limits +-1 come out exactly), used in training and inference
__device__ __forceinline__ float mlp_tanh(float x) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(2.8853900817779268f * x));
}

typedef float mlp_floatx16 __attribute__((ext_vector_type(16)));

// MLP forward for the 64 nodes of a wave, one per lane (FP32, tanh hidden
// layers, normalised inputs and outputs); every lane of the wave must take
// part.  The first and the output layer run on the VALU; the 32 x 32 hidden
// layer runs on the matrix cores as Z = W2 H1 for two blocks of 32 nodes
// (v_mfma_f32_32x32x2f32: A[i][k] from lane i + 32 k, B[k][j] from lane
// j + 32 k, lane l receiving rows 8 (r / 4) + 4 (l / 32) + r % 4 of column
// l % 32 in its 16 results r), each lane's h1 values reaching the other half
// of the wave through one lane swap per step.
__device__ __forceinline__ void mlp_eval_wave(const float* P, const float* norm, const float* xin,
                                              float* y) {
  const float* W1 = P;
  const float* b1 = W1 + kMlpHidden * kMlpIn;
  const float* W2 = b1 + kMlpHidden;
  const float* b2 = W2 + kMlpHidden * kMlpHidden;
  const float* W3 = b2 + kMlpHidden;
  const float* b3 = W3 + kMlpOut * kMlpHidden;
  const int l = threadIdx.x & 63, hi = l >> 5, lo = l & 31;
  float x[kMlpIn];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) x[i] = (xin[i] - norm[i]) / norm[3 + i];
  float h1[kMlpHidden];
#pragma unroll
  for (int j = 0; j < kMlpHidden; ++j) {
    float a = b1[j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) a += W1[j * kMlpIn + i] * x[i];
    h1[j] = mlp_tanh(a);
  }
  mlp_floatx16 d0 = {}, d1 = {};
#pragma unroll
  for (int s = 0; s < kMlpHidden / 2; ++s) {
    const float a = W2[lo * kMlpHidden + 2 * s + hi];
    const float give = hi ? h1[2 * s] : h1[2 * s + 1];
    const float recv = __shfl_xor(give, 32, 64);
    const float bb0 = hi ? recv : h1[2 * s];       // h1[2 s + hi] of node lo
    const float bb1 = hi ? h1[2 * s + 1] : recv;   // h1[2 s + hi] of node 32 + lo
    d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bb0, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bb1, d1, 0, 0, 0);
  }
  // output layer: each lane sums its 16 hidden rows, the lane 32 apart the others
  float p0[kMlpOut] = {0.0f, 0.0f}, p1[kMlpOut] = {0.0f, 0.0f};
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = 8 * (r / 4) + 4 * hi + r % 4;
    const float g0 = mlp_tanh(d0[r] + b2[row]), g1 = mlp_tanh(d1[r] + b2[row]);
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
      p0[o] += W3[o * kMlpHidden + row] * g0;
      p1[o] += W3[o * kMlpHidden + row] * g1;
    }
  }
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) {
    const float t0 = p0[o] + __shfl_xor(p0[o], 32, 64);  // node lo
    const float t1 = p1[o] + __shfl_xor(p1[o], 32, 64);  // node 32 + lo
    y[o] = ((hi ? t1 : t0) + b3[o]) * norm[8 + o] + norm[6 + o];
  }
}

struct MlpView {
  const float* params;  // kMlpParams
  const float* norm;    // in_mean[3], in_scale[3], out_mean[2], out_scale[2]
};

struct RainyLaunch {
  int n, nx, ny, n_elements;
  int periodic_y, volume_integral, surface_flux, source;
  double x0, y0, dx, dy;
  double Ja11, Ja22, inv_J;   // uniform Cartesian cells
  double xi[8];               // LGL nodes (n <= 8)
  double inv_w0, inv_wn;      // 1 / w_1, 1 / w_n
  double dsplit[64];          // Dsplit[i * n + m]
  double dhat[64];            // Trixi derivative_dhat[i * n + m]
  const double* u;
  double* out;
  double* k;
  int mode;                   // 0 rhs, 1 RK stage (+ Newton on the new state)
  double t;                   // time of the rhs evaluation (time-dependent sources)
  double rk_a, rk_dt, rk_b;
  // resident time loop: if set, the step's start time and size are read from
  // clock[0] and clock[1] (the stage's time is clock[0] + c_stage clock[1]),
  // and a step of size 0 copies the state unchanged
  const double* clock;
  double c_stage;
  // RK stage, if set: per element the StepsizeCallback sum over directions of
  // the maximum scaled speed of the new (limited) state
  double* speed_partial;
  double newton_tol;
  int newton_max_iter;
  // optional Newton statistics: per wave of a launch (slot = global wave
  // index, fewer than n_elements in both the stage and the Newton kernel)
  // [sum of updates, failures, max updates, fallbacks], accumulated over
  // launches until rainy_launch_stats_reduce collects and clears them
  unsigned long long* stats;
  MlpView mlp;                // seeds the stage limiter's solve when mlp.params is set
  int closed_form_seed;       // else, nonzero: the closed-form guess seeds it
};

bool rainy_launch_stage(const RainyLaunch& L, hipStream_t s);
// out[4] = the statistics summed over the slots [n_slots][4] (max of the
// largest update count); the slots are cleared
void rainy_launch_stats_reduce(unsigned long long* slots, int n_slots, unsigned long long* out,
                               hipStream_t s);
// per-element scaled wave speeds, then their maximum -> out[0]
bool rainy_launch_max_speed(const RainyLaunch& L, double* partial, double* out, hipStream_t s);
// integral of each of the 9 variables -> out[9]
bool rainy_launch_integrals(const RainyLaunch& L, const double* w, double* partial, double* out,
                            hipStream_t s);
// NonlinearSolveDG on the state L.out in place; with `mlp` the guess comes from
// the network and a node whose seeded solve fails is solved again from its
// own u[7:9]; iterations[node] (optional) receives the update count (-1: failed)
void rainy_launch_newton(const RainyLaunch& L, double* u, const MlpView* mlp, int* iterations,
                         hipStream_t s);

// MLP training: per-wave loss and gradient partials of samples
// [offset, offset + count) (through `perm` when given), returning the number
// of partials; their sum in a fixed order; Adam; forward in physical units
int mlp_launch_loss_grad(const float* params, const float* norm, const float* x, const float* y,
                         const int* perm, int64_t offset, int64_t count, float* partial,
                         hipStream_t s);
void mlp_launch_reduce(const float* partial, int blocks, float* grad_loss, hipStream_t s);
void mlp_launch_adam(float* params, float* m, float* v, const float* grad, float lr, float beta1,
                     float beta2, float eps, int step, hipStream_t s);
void mlp_launch_forward(const float* params, const float* norm, const float* x, float* y,
                        int64_t n, hipStream_t s);
// features of a rainy state: x = (rho, internal energy, total water), y = (T, q_v)
void rainy_launch_features(const RainyLaunch& L, const double* u, float* x, float* y,
                           hipStream_t s);

}  // namespace atmo

#endif
