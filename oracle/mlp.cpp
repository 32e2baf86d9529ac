// CPU ORACLE (test infrastructure only): FP64 forward, loss and gradient of the
// 3 -> 32 -> 32 -> 2 tanh network that seeds the rainy-Euler saturation
// solve (no counterpart in the reference; the checker of the library's HIP
// training kernels).  Parameters packed W1[32][3], b1[32], W2[32][32],
// b2[32], W3[2][32], b3[2]; inputs and targets normalised with
// norm = in_mean[3], in_scale[3], out_mean[2], out_scale[2]; loss = mean over
// samples of the summed squared output errors.
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {
constexpr int H = 32, I = 3, O = 2;
constexpr int oW1 = 0, ob1 = oW1 + H * I, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H,
              ob3 = oW3 + O * H, NP = ob3 + O;
}  // namespace

extern "C" void ora_mlp_loss_grad(const float* params, const float* norm, const float* x,
                                  const float* y, long n, double* loss, double* grad) {
  std::vector<double> P(params, params + NP), G(NP, 0.0);
  double L = 0.0;
  for (long s = 0; s < n; ++s) {
    double xi[I], yt[O];
    for (int k = 0; k < I; ++k) xi[k] = ((double)x[3 * s + k] - norm[k]) / norm[3 + k];
    for (int o = 0; o < O; ++o) yt[o] = ((double)y[2 * s + o] - norm[6 + o]) / norm[8 + o];
    double h1[H], h2[H], out[O];
    for (int j = 0; j < H; ++j) {
      double a = P[ob1 + j];
      for (int k = 0; k < I; ++k) a += P[oW1 + j * I + k] * xi[k];
      h1[j] = std::tanh(a);
    }
    for (int j = 0; j < H; ++j) {
      double a = P[ob2 + j];
      for (int k = 0; k < H; ++k) a += P[oW2 + j * H + k] * h1[k];
      h2[j] = std::tanh(a);
    }
    double dy[O];
    for (int o = 0; o < O; ++o) {
      double a = P[ob3 + o];
      for (int k = 0; k < H; ++k) a += P[oW3 + o * H + k] * h2[k];
      out[o] = a;
      const double d = a - yt[o];
      L += d * d / n;
      dy[o] = 2.0 * d / n;
    }
    double d2[H], d1[H];
    for (int j = 0; j < H; ++j) {
      double g = 0.0;
      for (int o = 0; o < O; ++o) g += dy[o] * P[oW3 + o * H + j];
      d2[j] = g * (1.0 - h2[j] * h2[j]);
    }
    for (int k = 0; k < H; ++k) {
      double g = 0.0;
      for (int j = 0; j < H; ++j) g += d2[j] * P[oW2 + j * H + k];
      d1[k] = g * (1.0 - h1[k] * h1[k]);
    }
    for (int o = 0; o < O; ++o) {
      G[ob3 + o] += dy[o];
      for (int k = 0; k < H; ++k) G[oW3 + o * H + k] += dy[o] * h2[k];
    }
    for (int j = 0; j < H; ++j) {
      G[ob2 + j] += d2[j];
      for (int k = 0; k < H; ++k) G[oW2 + j * H + k] += d2[j] * h1[k];
      G[ob1 + j] += d1[j];
      for (int k = 0; k < I; ++k) G[oW1 + j * I + k] += d1[j] * xi[k];
    }
    (void)out;
  }
  *loss = L;
  for (int i = 0; i < NP; ++i) grad[i] = G[i];
}

// the network's outputs y[2 n] in physical units (FP64 forward of FP32 parameters)
extern "C" void ora_mlp_forward(const float* params, const float* norm, const float* x, long n,
                                double* y) {
  std::vector<double> P(params, params + NP);
  for (long s = 0; s < n; ++s) {
    double xi[I], h1[H], h2[H];
    for (int k = 0; k < I; ++k) xi[k] = ((double)x[3 * s + k] - norm[k]) / norm[3 + k];
    for (int j = 0; j < H; ++j) {
      double a = P[ob1 + j];
      for (int k = 0; k < I; ++k) a += P[oW1 + j * I + k] * xi[k];
      h1[j] = std::tanh(a);
    }
    for (int j = 0; j < H; ++j) {
      double a = P[ob2 + j];
      for (int k = 0; k < H; ++k) a += P[oW2 + j * H + k] * h1[k];
      h2[j] = std::tanh(a);
    }
    for (int o = 0; o < O; ++o) {
      double a = P[ob3 + o];
      for (int k = 0; k < H; ++k) a += P[oW3 + o * H + k] * h2[k];
      y[2 * s + o] = a * norm[8 + o] + norm[6 + o];
    }
  }
}
