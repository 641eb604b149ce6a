// The cost model's network: a small multilayer perceptron
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
// with MLP_IN inputs, MLP_H hidden units per layer and MLP_OUT outputs (one
// predicted log execute time per plan option). Forward (inference), the
// gradient of the masked squared error over a batch, and an Adam step, as
// hand-written kernels; parameters are one packed fp32 array:
//   W1[H][IN] b1[H] W2[H][H] b2[H] W3[OUT][H] b3[OUT].
// The reference has no counterpart (its method choice is a fixed table,
// src/cuda/heuristics.cu:270-280).
#include <hip/hip_runtime.h>

#include <finufft_amd.h>

#include "common.hpp"

namespace fa {
namespace {

constexpr int IN = FINUFFT_AMD_MLP_IN;
constexpr int H = FINUFFT_AMD_MLP_HIDDEN;
constexpr int OUT = FINUFFT_AMD_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + OUT * H, NPARAM = OB3 + OUT;
static_assert(NPARAM == FINUFFT_AMD_MLP_NPARAM, "parameter layout");
constexpr int MT = 128;

struct Acts {
  float h1[H], h2[H], y[OUT];
};

__device__ __forceinline__ void forward(const float *__restrict__ p, const float *x, Acts &a) {
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float s = p[OB1 + j];
#pragma unroll
    for (int i = 0; i < IN; ++i) s = std::fma(p[OW1 + j * IN + i], x[i], s);
    a.h1[j] = s > 0.f ? s : 0.f;
  }
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float s = p[OB2 + j];
    for (int i = 0; i < H; ++i) s = std::fma(p[OW2 + j * H + i], a.h1[i], s);
    a.h2[j] = s > 0.f ? s : 0.f;
  }
#pragma unroll
  for (int j = 0; j < OUT; ++j) {
    float s = p[OB3 + j];
    for (int i = 0; i < H; ++i) s = std::fma(p[OW3 + j * H + i], a.h2[i], s);
    a.y[j] = s;
  }
}

__global__ void __launch_bounds__(MT)
    mlp_forward_kernel(const float *__restrict__ p, int n, const float *__restrict__ x,
                       float *__restrict__ y) {
  const int b = blockIdx.x * MT + threadIdx.x;
  if (b >= n) return;
  float xv[IN];
#pragma unroll
  for (int i = 0; i < IN; ++i) xv[i] = x[int64_t(b) * IN + i];
  Acts a;
  forward(p, xv, a);
#pragma unroll
  for (int j = 0; j < OUT; ++j) y[int64_t(b) * OUT + j] = a.y[j];
}

// Per sample: loss_b = 0.5 sum_j m_bj (y_bj - t_bj)^2; the batch loss is the
// mean over samples. Gradients are summed into g (zeroed by the caller) with
// global atomics; loss receives the batch loss the same way.
__global__ void __launch_bounds__(MT)
    mlp_grad_kernel(const float *__restrict__ p, int n, const float *__restrict__ x,
                    const float *__restrict__ t, const float *__restrict__ m,
                    float *__restrict__ g, float *__restrict__ loss) {
  const int b = blockIdx.x * MT + threadIdx.x;
  if (b >= n) return;
  const float inv_n = 1.f / float(n);
  float xv[IN];
#pragma unroll
  for (int i = 0; i < IN; ++i) xv[i] = x[int64_t(b) * IN + i];
  Acts a;
  forward(p, xv, a);
  float dy[OUT];
  float l = 0.f;
#pragma unroll
  for (int j = 0; j < OUT; ++j) {
    const float mk = m[int64_t(b) * OUT + j];
    const float e = mk * (a.y[j] - t[int64_t(b) * OUT + j]);
    l = std::fma(0.5f * e, a.y[j] - t[int64_t(b) * OUT + j], l);
    dy[j] = e * inv_n;
  }
  atomicAdd(loss, l * inv_n);
  // layer 3
  float dh2[H];
  for (int i = 0; i < H; ++i) dh2[i] = 0.f;
#pragma unroll
  for (int j = 0; j < OUT; ++j) {
    if (dy[j] == 0.f) continue;
    atomicAdd(g + OB3 + j, dy[j]);
    for (int i = 0; i < H; ++i) {
      atomicAdd(g + OW3 + j * H + i, dy[j] * a.h2[i]);
      dh2[i] = std::fma(p[OW3 + j * H + i], dy[j], dh2[i]);
    }
  }
  // layer 2
  float dh1[H];
  for (int i = 0; i < H; ++i) dh1[i] = 0.f;
  for (int j = 0; j < H; ++j) {
    const float dz = a.h2[j] > 0.f ? dh2[j] : 0.f;
    if (dz == 0.f) continue;
    atomicAdd(g + OB2 + j, dz);
    for (int i = 0; i < H; ++i) {
      atomicAdd(g + OW2 + j * H + i, dz * a.h1[i]);
      dh1[i] = std::fma(p[OW2 + j * H + i], dz, dh1[i]);
    }
  }
  // layer 1
  for (int j = 0; j < H; ++j) {
    const float dz = a.h1[j] > 0.f ? dh1[j] : 0.f;
    if (dz == 0.f) continue;
    atomicAdd(g + OB1 + j, dz);
#pragma unroll
    for (int i = 0; i < IN; ++i) atomicAdd(g + OW1 + j * IN + i, dz * xv[i]);
  }
}

__global__ void mlp_adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                float *__restrict__ m1, float *__restrict__ m2, float lr,
                                float b1, float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a / c1) / (std::sqrt(v / c2) + eps);
}

}  // namespace

void mlp_forward(const float *p, int n, const float *x, float *y, hipStream_t s) {
  if (n <= 0) return;
  mlp_forward_kernel<<<(n + MT - 1) / MT, MT, 0, s>>>(p, n, x, y);
  launch_check("mlp_forward_kernel");
}

void mlp_grad(const float *p, int n, const float *x, const float *t, const float *m, float *g,
              float *loss, hipStream_t s) {
  if (n <= 0) return;
  mlp_grad_kernel<<<(n + MT - 1) / MT, MT, 0, s>>>(p, n, x, t, m, g, loss);
  launch_check("mlp_grad_kernel");
}

void mlp_adam(float *p, const float *g, float *m1, float *m2, int step, float lr,
              hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(step)), c2 = 1.f - std::pow(b2, float(step));
  mlp_adam_kernel<<<(NPARAM + 255) / 256, 256, 0, s>>>(p, g, m1, m2, lr, b1, b2, c1, c2, 1e-8f);
  launch_check("mlp_adam_kernel");
}

}  // namespace fa
