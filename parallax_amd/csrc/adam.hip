// Fused Adam step, gradient zeroing and densify-statistics accumulation.
//
// The engine's applyGradients / adamDelta (engine/native/cpp/src/splat/
// SplatTrainer.cpp:56-65, 218-256): per parameter group a learning rate,
// shared betas and bias corrections; the quaternion is renormalised after its
// update (identity if its norm collapses). Also folds in accumulateStatistics
// (:258-263) for the densify signal. Elementwise over the capacity-laid
// parameter buffer; grid.y selects the group. HBM-bound: 28 B/param of
// compulsory traffic (read p, g, m, v; write p, m, v) plus the 4 B gradient
// clear. Arithmetic is IEEE float without contraction, so with equal
// gradients the update matches the CPU trainer bit for bit.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

__device__ __forceinline__ float adam_delta(float& m, float& v, float g, float lr, const AdamParams& p) {
  m = p.beta1 * m + (1.f - p.beta1) * g;
  v = p.beta2 * v + (1.f - p.beta2) * g * g;
  const float m_hat = m / p.c1;
  const float v_hat = v / p.c2;
  return -lr * m_hat / (__builtin_sqrtf(v_hat) + p.eps);
}

// groups: 0 pos (3), 1 log_scale (3), 2 rot (4, per gaussian), 3 opacity (1),
// 4 dc (3), 5 rest (rs)
__global__ __launch_bounds__(256) void k_adam(AdamParams p, float* __restrict__ params, float* __restrict__ grads,
                                              float* __restrict__ m, float* __restrict__ v, float* __restrict__ sg,
                                              int* __restrict__ vc, float* __restrict__ stat_sg,
                                              int* __restrict__ stat_vc) {
  const int group = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long cap = p.cap, n = p.n;
  if (group == 2) {
    if (i >= n) return;
    const long long o = 6 * cap + 4 * i;
    float q[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float mm = m[o + a], vv = v[o + a];
      q[a] = params[o + a] + adam_delta(mm, vv, grads[o + a], p.lr[2], p);
      m[o + a] = mm;
      v[o + a] = vv;
      grads[o + a] = 0.f;
    }
    // Eigen norm() / normalize() over the (x, y, z, w) coefficients
    const float sq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + q[0] * q[0];
    const float nrm = __builtin_sqrtf(sq);
    if (nrm > 1e-6f) {
#pragma unroll
      for (int a = 0; a < 4; ++a) q[a] = q[a] / nrm;
    } else {
      q[0] = 1.f;
      q[1] = q[2] = q[3] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) params[o + a] = q[a];
    stat_sg[i] += sg[i];
    stat_vc[i] += vc[i];
    sg[i] = 0.f;
    vc[i] = 0;
    return;
  }
  long long off, count;
  switch (group) {
    case 0: off = 0; count = 3 * n; break;
    case 1: off = 3 * cap; count = 3 * n; break;
    case 3: off = 10 * cap; count = n; break;
    case 4: off = 11 * cap; count = 3 * n; break;
    default: off = 14 * cap; count = (long long)p.rest_stride * n; break;
  }
  if (i >= count) return;
  const long long k = off + i;
  float mm = m[k], vv = v[k];
  params[k] += adam_delta(mm, vv, grads[k], p.lr[group], p);
  m[k] = mm;
  v[k] = vv;
  grads[k] = 0.f;
}

}  // namespace

hipError_t launch_adam(const AdamParams& p, float* params, float* grads, float* m, float* v, float* sg, int* vc,
                       float* stat_sg, int* stat_vc, hipStream_t stream) {
  if (p.n <= 0) return hipSuccess;
  long long widest = 4LL * p.n;
  if ((long long)p.rest_stride * p.n > widest) widest = (long long)p.rest_stride * p.n;
  const unsigned blocks = (unsigned)((widest + 255) / 256);
  hipLaunchKernelGGL(k_adam, dim3(blocks, 6), dim3(256), 0, stream, p, params, grads, m, v, sg, vc, stat_sg, stat_vc);
  return hipGetLastError();
}

}  // namespace splat
