// Fused Adam step, gradient zeroing and densify-statistics accumulation.
//
// The engine's applyGradients / adamDelta (engine/native/cpp/src/splat/
// SplatTrainer.cpp:56-65, 218-256): per parameter group a learning rate,
// shared betas and bias corrections; the quaternion is renormalised after its
// update (identity if its norm collapses). Also folds in accumulateStatistics
// (:258-263) for the densify signal. Elementwise over the capacity-laid
// parameter buffer, four elements per thread (16-byte non-temporal accesses
// where the group is aligned: the step streams them once). HBM-bound: 28 B/param of
// compulsory traffic (read p, g, m, v; write p, m, v) plus the 4 B gradient
// clear. Arithmetic is IEEE float without contraction, so with equal
// gradients the update matches the CPU trainer bit for bit.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

__device__ __forceinline__ float adam_delta(float& m, float& v, float g, float lr, const AdamStep& p) {
  m = p.beta1 * m + (1.f - p.beta1) * g;
  v = p.beta2 * v + (1.f - p.beta2) * g * g;
  const float m_hat = m / p.c1;
  const float v_hat = v / p.c2;
  return -lr * m_hat / (__builtin_sqrtf(v_hat) + p.eps);
}

// Groups: 0 pos (3n), 1 log_scale (3n), 2 rot (4n, one quaternion per
// thread), 3 opacity (n), 4 dc (3n), 5 rest (rs n). Each group's elements are
// padded to a multiple of 4 in a virtual index space that one flat grid
// covers with 4 elements per thread, so no workgroup is launched idle. The
// group of a unit is selected with compares, never by indexing a local array
// (which would live in scratch memory).
struct Groups {
  long long cap, n, rs;
  long long vb[6];  // virtual start of groups 1..5, then the end
};

__device__ __forceinline__ long long padded(long long len) { return (len + 3) & ~3LL; }

__device__ __forceinline__ Groups groups_of(const AdamShape& s) {
  Groups G;
  G.cap = s.cap;
  G.n = s.n;
  G.rs = s.rest_stride;
  G.vb[0] = padded(3 * G.n);
  G.vb[1] = G.vb[0] + padded(3 * G.n);
  G.vb[2] = G.vb[1] + 4 * G.n;
  G.vb[3] = G.vb[2] + padded(G.n);
  G.vb[4] = G.vb[3] + padded(3 * G.n);
  G.vb[5] = G.vb[4] + padded(G.rs * G.n);
  return G;
}

struct Unit {
  int g;             // group
  long long local;   // element index within the group
  long long o;       // element index in the parameter buffer
  long long len;     // elements in the group
};

__device__ __forceinline__ Unit unit_of(const Groups& G, long long v0) {
  Unit u;
  const int g = (v0 >= G.vb[0]) + (v0 >= G.vb[1]) + (v0 >= G.vb[2]) + (v0 >= G.vb[3]) + (v0 >= G.vb[4]);
  const long long cap = G.cap, n = G.n;
  long long vbase = 0, off = 0, len = 3 * n;
  if (g == 1) { vbase = G.vb[0]; off = 3 * cap; }
  if (g == 2) { vbase = G.vb[1]; off = 6 * cap; len = 4 * n; }
  if (g == 3) { vbase = G.vb[2]; off = 10 * cap; len = n; }
  if (g == 4) { vbase = G.vb[3]; off = 11 * cap; }
  if (g == 5) { vbase = G.vb[4]; off = 14 * cap; len = G.rs * n; }
  u.g = g;
  u.local = v0 - vbase;
  u.o = off + u.local;
  u.len = len;
  return u;
}

__device__ __forceinline__ float lr_of(const AdamStep& P, int g) {
  float lr = P.lr[0];
  lr = g == 1 ? P.lr[1] : lr;
  lr = g == 2 ? P.lr[2] : lr;
  lr = g == 3 ? P.lr[3] : lr;
  lr = g == 4 ? P.lr[4] : lr;
  lr = g == 5 ? P.lr[5] : lr;
  return lr;
}

// One 4-element unit of the virtual index space.
__device__ __forceinline__ void adam_unit(const Groups& G, const AdamStep& P, long long v0, float* __restrict__ params,
                                          float* __restrict__ grads, float* __restrict__ m, float* __restrict__ v,
                                          float* __restrict__ sg, int* __restrict__ vc, float* __restrict__ stat_sg,
                                          int* __restrict__ stat_vc) {
  const Unit U = unit_of(G, v0);
  const float lr = lr_of(P, U.g);
  const long long o = U.o;
  if (U.g == 2) {
    const long long i = U.local / 4;  // gaussian
    float q[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float mm = m[o + a], vv = v[o + a];
      q[a] = params[o + a] + adam_delta(mm, vv, grads[o + a], lr, P);
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
  const long long count = U.len - U.local;  // elements of this group left from here
  if (count >= 4 && (o & 3) == 0) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v pp = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(params + o));
    const f4v gg = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(grads + o));
    f4v mm = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(m + o));
    f4v vv = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(v + o));
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float ma = mm[a], va = vv[a];
      pp[a] += adam_delta(ma, va, gg[a], lr, P);
      mm[a] = ma;
      vv[a] = va;
    }
    __builtin_nontemporal_store(pp, reinterpret_cast<f4v*>(params + o));
    __builtin_nontemporal_store(mm, reinterpret_cast<f4v*>(m + o));
    __builtin_nontemporal_store(vv, reinterpret_cast<f4v*>(v + o));
    __builtin_nontemporal_store((f4v){0.f, 0.f, 0.f, 0.f}, reinterpret_cast<f4v*>(grads + o));
    return;
  }
  const int todo = count < 4 ? (int)count : 4;
  for (int a = 0; a < todo; ++a) {
    float mm = m[o + a], vv = v[o + a];
    params[o + a] += adam_delta(mm, vv, grads[o + a], lr, P);
    m[o + a] = mm;
    v[o + a] = vv;
    grads[o + a] = 0.f;
  }
}

// One thread per 4-element unit.
__global__ __launch_bounds__(256) void k_adam(AdamShape shape, const AdamStep* __restrict__ step,
                                              const uint32_t* __restrict__ skip, float* __restrict__ params,
                                              float* __restrict__ grads, float* __restrict__ m,
                                              float* __restrict__ v, float* __restrict__ sg, int* __restrict__ vc,
                                              float* __restrict__ stat_sg, int* __restrict__ stat_vc) {
  if (skip && *skip) return;
  const AdamStep P = *step;
  const Groups G = groups_of(shape);
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < G.vb[5] / 4) adam_unit(G, P, 4 * u, params, grads, m, v, sg, vc, stat_sg, stat_vc);
}

}  // namespace

hipError_t launch_adam(const AdamShape& s, const AdamStep* step, const uint32_t* skip_flag, float* params,
                       float* grads, float* m, float* v, float* sg, int* vc, float* stat_sg, int* stat_vc,
                       hipStream_t stream) {
  if (s.n <= 0) return hipSuccess;
  const long long n = s.n;
  const long long virt = ((3 * n + 3) & ~3LL) * 3 + 4 * n + ((n + 3) & ~3LL) + (((long long)s.rest_stride * n + 3) & ~3LL);
  const long long threads = virt / 4;
  const long long blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, stream, s, step, skip_flag, params, grads, m, v,
                     sg, vc, stat_sg, stat_vc);
  return hipGetLastError();
}

}  // namespace splat
