// gfx950 kernels of the small MLP (3 -> 32 -> 32 -> 2, tanh, FP32) that seeds
// the rainy-Euler saturation solve: batched forward, loss and gradient, and
// Adam.  One wave of 64 lanes takes 64 samples: each lane runs its sample's
// forward and backward pass, the per-sample activations and deltas are
// staged in LDS, and the wave then contracts them into the weight gradients
// (lane t owns a fixed set of weights), writing one partial per wave.  A
// second kernel sums the partials in a fixed order, so a gradient is
// bit-reproducible; Adam then updates the parameters.
#include <hip/hip_runtime.h>

#include "rainy_kernels.hpp"

namespace atmo {

namespace {

constexpr int H = kMlpHidden, I = kMlpIn, O = kMlpOut, P = kMlpParams;
constexpr int oW1 = 0, ob1 = oW1 + H * I, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H,
              ob3 = oW3 + O * H;
constexpr int kPartial = P + 1;  // gradient + loss

}  // namespace

// partial[block][P + 1] for samples [blockIdx.x * 64, +64) of n (norm applied)
__global__ __launch_bounds__(64) void mlp_loss_grad_kernel(const float* params, const float* norm,
                                                           const float* x, const float* y,
                                                           const int* perm, int64_t offset,
                                                           int64_t count, float inv_count,
                                                           float* partial) {
  __shared__ float sP[P];
  __shared__ float sx[64][I];
  __shared__ float sh1[64][H + 1];
  __shared__ float sh2[64][H + 1];
  __shared__ float sd1[64][H + 1];
  __shared__ float sd2[64][H + 1];
  __shared__ float sdy[64][O];
  __shared__ float sloss[64];
  const int t = threadIdx.x;
  for (int i = t; i < P; i += 64) sP[i] = params[i];
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * 64 + t;
  const bool valid = b < count;
  float xi[I] = {0, 0, 0}, yt[O] = {0, 0};
  if (valid) {
    const int64_t src = perm ? (int64_t)perm[offset + b] : offset + b;
#pragma unroll
    for (int k = 0; k < I; ++k) xi[k] = (x[3 * src + k] - norm[k]) / norm[3 + k];
#pragma unroll
    for (int o = 0; o < O; ++o) yt[o] = (y[2 * src + o] - norm[6 + o]) / norm[8 + o];
  }
  float h1[H], h2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = sP[ob1 + j];
#pragma unroll
    for (int k = 0; k < I; ++k) a += sP[oW1 + j * I + k] * xi[k];
    h1[j] = mlp_tanh(a);
  }
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float a = sP[ob2 + j];
#pragma unroll
    for (int k = 0; k < H; ++k) a += sP[oW2 + j * H + k] * h1[k];
    h2[j] = mlp_tanh(a);
  }
  float dy[O], loss = 0.0f;
#pragma unroll
  for (int o = 0; o < O; ++o) {
    float a = sP[ob3 + o];
#pragma unroll
    for (int k = 0; k < H; ++k) a += sP[oW3 + o * H + k] * h2[k];
    const float d = valid ? a - yt[o] : 0.0f;
    loss += d * d;
    dy[o] = 2.0f * d * inv_count;
  }
  // backward through the hidden layers (deltas of the pre-activations)
  float d2[H], d1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float g = 0.0f;
#pragma unroll
    for (int o = 0; o < O; ++o) g += dy[o] * sP[oW3 + o * H + j];
    d2[j] = g * (1.0f - h2[j] * h2[j]);
  }
#pragma unroll 4
  for (int k = 0; k < H; ++k) {
    float g = 0.0f;
#pragma unroll
    for (int j = 0; j < H; ++j) g += d2[j] * sP[oW2 + j * H + k];
    d1[k] = g * (1.0f - h1[k] * h1[k]);
  }
#pragma unroll
  for (int k = 0; k < I; ++k) sx[t][k] = xi[k];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    sh1[t][j] = h1[j];
    sh2[t][j] = h2[j];
    sd1[t][j] = d1[j];
    sd2[t][j] = d2[j];
  }
#pragma unroll
  for (int o = 0; o < O; ++o) sdy[t][o] = dy[o];
  sloss[t] = loss * inv_count;
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * kPartial;
  // W2: lane t owns rows j = t/2 and columns (t%2)*16 .. +16
  {
    const int j = t >> 1, k0 = (t & 1) * 16;
    float acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
    for (int s = 0; s < 64; ++s) {
      const float d = sd2[s][j];
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[c] += d * sh1[s][k0 + c];
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) out[oW2 + j * H + k0 + c] = acc[c];
  }
  // W1 (96), b1 (32), b2 (32), W3 (64), b3 (2): a flat list over the lanes
  for (int idx = t; idx < H * I + H + H + O * H + O; idx += 64) {
    float acc = 0.0f;
    int dst;
    if (idx < H * I) {
      const int j = idx / I, k = idx % I;
      for (int s = 0; s < 64; ++s) acc += sd1[s][j] * sx[s][k];
      dst = oW1 + idx;
    } else if (idx < H * I + H) {
      const int j = idx - H * I;
      for (int s = 0; s < 64; ++s) acc += sd1[s][j];
      dst = ob1 + j;
    } else if (idx < H * I + 2 * H) {
      const int j = idx - H * I - H;
      for (int s = 0; s < 64; ++s) acc += sd2[s][j];
      dst = ob2 + j;
    } else if (idx < H * I + 2 * H + O * H) {
      const int r = idx - H * I - 2 * H, o = r / H, k = r % H;
      for (int s = 0; s < 64; ++s) acc += sdy[s][o] * sh2[s][k];
      dst = oW3 + r;
    } else {
      const int o = idx - H * I - 2 * H - O * H;
      for (int s = 0; s < 64; ++s) acc += sdy[s][o];
      dst = ob3 + o;
    }
    out[dst] = acc;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int s = 0; s < 64; ++s) l += sloss[s];
    out[P] = l;
  }
}

// grad_loss[P + 1] = sum over blocks of the partials, in block order
__global__ void mlp_reduce_kernel(const float* partial, int blocks, float* grad_loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kPartial) return;
  float acc = 0.0f;
  for (int b = 0; b < blocks; ++b) acc += partial[(size_t)b * kPartial + i];
  grad_loss[i] = acc;
}

__global__ void mlp_adam_kernel(float* params, float* m, float* v, const float* grad, float lr,
                                float beta1, float beta2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float g = grad[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * g;
  const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// y[2 n] = network(x[3 n]) in physical units, by the same wave-wide
// evaluation (hidden layer on the matrix cores) that seeds the saturation
// solve; lanes past n evaluate the normalisation means and write nothing
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* params, const float* norm,
                                                          const float* x, float* y, int64_t n) {
  __shared__ float sP[P];
  for (int i = threadIdx.x; i < P; i += blockDim.x) sP[i] = params[i];
  __syncthreads();
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n;
  float xi[I], yo[O];
#pragma unroll
  for (int k = 0; k < I; ++k) xi[k] = valid ? x[3 * q + k] : norm[k];
  mlp_eval_wave(sP, norm, xi, yo);
  if (valid) {
#pragma unroll
    for (int o = 0; o < O; ++o) y[2 * q + o] = yo[o];
  }
}

}  // namespace atmo

namespace atmo {

int mlp_launch_loss_grad(const float* params, const float* norm, const float* x, const float* y,
                         const int* perm, int64_t offset, int64_t count, float* partial,
                         hipStream_t s) {
  const int blocks = (int)((count + 63) / 64);
  hipLaunchKernelGGL(mlp_loss_grad_kernel, dim3(blocks), dim3(64), 0, s, params, norm, x, y, perm,
                     offset, count, 1.0f / (float)count, partial);
  return blocks;
}

void mlp_launch_reduce(const float* partial, int blocks, float* grad_loss, hipStream_t s) {
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s,
                     partial, blocks, grad_loss);
}

void mlp_launch_adam(float* params, float* m, float* v, const float* grad, float lr, float beta1,
                     float beta2, float eps, int step, hipStream_t s) {
  const float bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, m,
                     v, grad, lr, beta1, beta2, eps, bc1, bc2);
}

void mlp_launch_forward(const float* params, const float* norm, const float* x, float* y,
                        int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     params, norm, x, y, n);
}

}  // namespace atmo
