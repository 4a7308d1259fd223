// Adaptive density control on the device: clone / split / prune as stream
// compaction, the Adam-moment remap, opacity reset, and the final prune +
// sort by opacity.
//
// densifyAndPrune (engine/native/cpp/src/splat/SplatDensify.cpp:28-109) with
// the trainer's remapState (SplatTrainer.cpp:267-288):
//   1. classify: per gaussian, clone / split / none (only while below the
//      cap), whether the survivor is pruned, and how many of its additions
//      are prunable;
//   2. exclusive scans give each survivor its new slot (old order kept), each
//      accepted addition its slot after the survivors (parent order), and
//      each split parent its rank;
//   3. scatter: parameters and moments of survivors move to their new slots,
//      additions get zero moments; source_index records old index or -1.
// Additions stop at the cap exactly where the engine's loop breaks.
// The split children's normals come from the engine's single xorshift32
// stream consumed in parent order (6 normals, 12 draws per split parent).
// xorshift32 is linear over GF(2), so the state at draw j is A^j s: the
// host precomputes A^(6*2^i) as 32x32 bit matrices and each split parent
// jumps straight to its own position in the stream.
#include <cmath>
#include <vector>

#include "../../include/splat_hip.h"
#include "splat_context.h"
#include "splat_kernels.h"

using namespace splat;

namespace {

constexpr int kJumpLevels = 28;  // A^(6 * 2^i), i < 28: 2^27 split children

__device__ __forceinline__ float max_scale(const float* ls) {
  const float a = det_exp(ls[0]), b = det_exp(ls[1]), c = det_exp(ls[2]);
  const float ab = (a < b) ? b : a;
  return (ab < c) ? c : ab;
}

struct DensifyArgs {
  long long cap;
  int n;
  int rest_stride;
  int has_room;
  int prune_oversized;
  float threshold;
  float size_threshold;
  float max_scale;
  float prune_opacity;
  float log_div;
};

__device__ __forceinline__ bool prunable(float logit, const float* ls, const DensifyArgs& a) {
  const float opacity = 1.0f / (1.0f + det_exp(-logit));
  if (opacity < a.prune_opacity) return true;
  return a.prune_oversized && max_scale(ls) > a.max_scale;
}

// flags: [0] survivor, [1] non-prunable additions, [2] prunable additions,
// [3] split parent, [4] additions (all)
__global__ __launch_bounds__(256) void k_classify(DensifyArgs a, const float* __restrict__ params,
                                                  const float* __restrict__ sg, const int* __restrict__ vc,
                                                  uint32_t* __restrict__ flags) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  const long long cap = a.cap;
  const float* ls = params + 3 * cap + 3 * (long long)g;
  const float logit = params[10 * cap + g];
  int action = 0;  // 0 none, 1 clone, 2 split
  if (a.has_room) {
    const int seen = vc[g];
    if (seen > 0) {
      const float mean_gradient = sg[g] / (float)seen;
      if (!(mean_gradient < a.threshold)) action = max_scale(ls) <= a.size_threshold ? 1 : 2;
    }
  }
  const bool survivor = action != 2 && !prunable(logit, ls, a);
  uint32_t np = 0, pr = 0;
  if (action == 1) {
    if (prunable(logit, ls, a)) pr = 1; else np = 1;
  } else if (action == 2) {
    const float child_ls[3] = {ls[0] - a.log_div, ls[1] - a.log_div, ls[2] - a.log_div};
    if (prunable(logit, child_ls, a)) pr = 2; else np = 2;
  }
  const long long n = a.n;
  flags[g] = survivor ? 1u : 0u;
  flags[n + g] = np;
  flags[2 * n + g] = pr;
  flags[3 * n + g] = action == 2 ? 1u : 0u;
  flags[4 * n + g] = np + pr;
}

__device__ __forceinline__ uint32_t gf2_apply(const uint32_t* cols, uint32_t s) {
  uint32_t r = 0;
#pragma unroll 8
  for (int b = 0; b < 32; ++b)
    if ((s >> b) & 1u) r ^= cols[b];
  return r;
}

__device__ __forceinline__ uint32_t xorshift(uint32_t& s) {
  uint32_t x = s;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  s = x == 0u ? 1u : x;
  return s;
}

// nextNormal (SplatDensify.cpp:11-24).
__device__ __forceinline__ float next_normal(uint32_t& s) {
  const float u1 = fmaxf(1e-7f, (float)xorshift(s) / 4294967296.f);
  const float u2 = (float)xorshift(s) / 4294967296.f;
  return sqrtf(-2.f * logf(u1)) * cosf(2.f * 3.14159265358979f * u2);
}

// offs: [0] survivor offsets, [1] accepted-addition ranks, [2] split ranks,
// [3] addition sequence positions; scal: [0] S, [1] K, [2] cut, [3] pruned
__global__ __launch_bounds__(256) void k_scatter(DensifyArgs a, const float* __restrict__ params,
                                                 const float* __restrict__ m, const float* __restrict__ v,
                                                 float* __restrict__ out_p, float* __restrict__ out_m,
                                                 float* __restrict__ out_v, const uint32_t* __restrict__ flags,
                                                 const uint32_t* __restrict__ offs, const uint32_t* __restrict__ scal,
                                                 const uint32_t* __restrict__ jump, uint32_t seed,
                                                 int* __restrict__ source) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  const long long cap = a.cap, n = a.n;
  const int rs = a.rest_stride;
  const uint32_t S = scal[0], K = scal[1];
  const long long group_off[6] = {0, 3 * cap, 6 * cap, 10 * cap, 11 * cap, 14 * cap};
  const int group_w[6] = {3, 3, 4, 1, 3, rs};
  if (flags[g]) {
    const long long d = offs[g];
    for (int q = 0; q < 6; ++q)
      for (int k = 0; k < group_w[q]; ++k) {
        const long long src = group_off[q] + (long long)group_w[q] * g + k;
        const long long dst = group_off[q] + (long long)group_w[q] * d + k;
        out_p[dst] = params[src];
        out_m[dst] = m[src];
        out_v[dst] = v[src];
      }
    source[d] = g;
  }
  const uint32_t np = flags[n + g];
  if (np == 0) return;
  const uint32_t rank0 = offs[n + g];
  const bool split = flags[3 * n + g] != 0;
  float R[3][3], sc[3];
  if (split) {
    const float* rp = params + 6 * cap + 4 * (long long)g;
    float qw = rp[0], qx = rp[1], qy = rp[2], qz = rp[3];
    const float sq = qx * qx + qy * qy + qz * qz + qw * qw;
    if (sq > 0.f) {
      const float s = __builtin_sqrtf(sq);
      qx = qx / s; qy = qy / s; qz = qz / s; qw = qw / s;
    }
    const float tx = 2.f * qx, ty = 2.f * qy, tz = 2.f * qz;
    const float twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const float txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const float tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    R[0][0] = 1.f - (tyy + tzz); R[0][1] = txy - twz; R[0][2] = txz + twy;
    R[1][0] = txy + twz; R[1][1] = 1.f - (txx + tzz); R[1][2] = tyz - twx;
    R[2][0] = txz - twy; R[2][1] = tyz + twx; R[2][2] = 1.f - (txx + tyy);
    const float* ls = params + 3 * cap + 3 * (long long)g;
    sc[0] = det_exp(ls[0]);
    sc[1] = det_exp(ls[1]);
    sc[2] = det_exp(ls[2]);
  }
  for (uint32_t c = 0; c < np; ++c) {
    const uint32_t r = rank0 + c;
    if (r >= K) break;
    const long long d = (long long)S + r;
    for (int q = 0; q < 6; ++q)
      for (int k = 0; k < group_w[q]; ++k) {
        const long long src = group_off[q] + (long long)group_w[q] * g + k;
        const long long dst = group_off[q] + (long long)group_w[q] * d + k;
        out_p[dst] = params[src];
        out_m[dst] = 0.f;
        out_v[dst] = 0.f;
      }
    if (split) {
      // this child's place in the stream: 6 normals per child, 2 children
      // per split parent, in parent order
      uint32_t state = seed;
      uint32_t jumps = 2u * offs[2 * n + g] + c;
      for (int i = 0; i < kJumpLevels && jumps; ++i, jumps >>= 1)
        if (jumps & 1u) state = gf2_apply(jump + 32 * i, state);
      float sample[3];
      sample[2] = next_normal(state) * sc[2];
      sample[1] = next_normal(state) * sc[1];
      sample[0] = next_normal(state) * sc[0];
      for (int i = 0; i < 3; ++i)
        out_p[3 * d + i] = params[3 * (long long)g + i] + (R[i][0] * sample[0] + R[i][1] * sample[1] + R[i][2] * sample[2]);
      const float* ls = params + 3 * cap + 3 * (long long)g;
      for (int i = 0; i < 3; ++i) out_p[3 * cap + 3 * d + i] = ls[i] - a.log_div;
    }
    source[d] = -1;
  }
}

// the addition at sequence position `cut` is the last accepted one when the
// cap is reached; prunable additions before it are counted
__global__ __launch_bounds__(256) void k_cut(DensifyArgs a, const uint32_t* __restrict__ flags,
                                             const uint32_t* __restrict__ offs, uint32_t* __restrict__ scal) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  const long long n = a.n;
  const uint32_t K = scal[1];
  const uint32_t np = flags[n + g];
  if (np == 0 || K == 0) return;
  const uint32_t rank0 = offs[n + g];
  if (K - 1 >= rank0 && K - 1 < rank0 + np) scal[2] = offs[3 * n + g] + (K - 1 - rank0);
}

__global__ __launch_bounds__(256) void k_count_pruned(DensifyArgs a, const uint32_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ offs,
                                                      uint32_t* __restrict__ scal) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  const long long n = a.n;
  uint32_t count = 0;
  // originals: not split, not a survivor
  if (!flags[g] && !flags[3 * n + g]) count += 1;
  const uint32_t pr = flags[2 * n + g];
  if (pr) {
    const uint32_t cut = scal[2];
    const uint32_t pos = offs[3 * n + g];
    for (uint32_t c = 0; c < pr; ++c)
      if (pos + c < cut) count += 1;
  }
  if (count) atomicAdd(&scal[3], count);
}

__global__ void k_reset_opacity(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                                long long cap, int n, float ceiling) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const long long k = 10 * cap + g;
  const float logit = params[k];
  params[k] = (ceiling < logit) ? ceiling : logit;
  m[k] = 0.f;
  v[k] = 0.f;
}

// prune + sortByOpacity: keys sort ascending to descending logit; culled get
// the largest key so they land last
__global__ void k_finish_keys(const float* __restrict__ params, long long cap, int n, float min_opacity,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* __restrict__ keep) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float logit = params[10 * cap + g];
  const float* ls = params + 3 * cap + 3 * (long long)g;
  const float mx = fmaxf(fmaxf(ls[0], ls[1]), ls[2]);
  const float largest = det_exp(mx);
  const float opacity = 1.0f / (1.0f + det_exp(-logit));
  const bool ok = !(opacity < min_opacity) && (largest == largest) && !(largest > 1e3f) && largest < INFINITY;
  const uint32_t b = __float_as_uint(logit);
  const uint32_t asc = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  keys[g] = ok ? ~asc : 0xFFFFFFFFu;
  vals[g] = (uint32_t)g;
  keep[g] = ok ? 1u : 0u;
}

__global__ void k_gather(const float* __restrict__ params, float* __restrict__ out, long long cap, int count,
                         int rest_stride, const uint32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const long long g = order[i];
  const long long group_off[6] = {0, 3 * cap, 6 * cap, 10 * cap, 11 * cap, 14 * cap};
  const int group_w[6] = {3, 3, 4, 1, 3, rest_stride};
  for (int q = 0; q < 6; ++q)
    for (int k = 0; k < group_w[q]; ++k)
      out[group_off[q] + (long long)group_w[q] * i + k] = params[group_off[q] + (long long)group_w[q] * g + k];
}

// Column j of A (xorshift32 as a GF(2) matrix) is A e_j.
uint32_t xorshift_host(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}

void mat_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {  // out = a b (column form)
  for (int j = 0; j < 32; ++j) {
    uint32_t r = 0;
    for (int k = 0; k < 32; ++k)
      if ((b[j] >> k) & 1u) r ^= a[k];
    out[j] = r;
  }
}

uint32_t mat_apply(const uint32_t* a, uint32_t s) {
  uint32_t r = 0;
  for (int b = 0; b < 32; ++b)
    if ((s >> b) & 1u) r ^= a[b];
  return r;
}

// table[i] = A^(6 * 2^i); also the host jump of `steps` draws.
std::vector<uint32_t> jump_table() {
  std::vector<uint32_t> A(32), P(32), T(32);
  for (int j = 0; j < 32; ++j) A[j] = xorshift_host(1u << j);
  // P = A^6
  P = A;
  for (int k = 1; k < 6; ++k) {
    mat_mul(A.data(), P.data(), T.data());
    P = T;
  }
  std::vector<uint32_t> table(32 * kJumpLevels);
  for (int i = 0; i < kJumpLevels; ++i) {
    std::copy(P.begin(), P.end(), table.begin() + 32 * i);
    mat_mul(P.data(), P.data(), T.data());
    P = T;
  }
  return table;
}

}  // namespace

#define CKD(expr, what)                                          \
  do {                                                           \
    hipError_t _e = (expr);                                      \
    if (_e != hipSuccess) {                                      \
      ctx->error = std::string(what) + ": " + hipGetErrorString(_e); \
      return SPLAT_ERROR_DEVICE;                                 \
    }                                                            \
  } while (0)

extern "C" {

splat_status splat_densify(splat_context* ctx, const splat_densify_params* p, float extent, int32_t iteration,
                           uint32_t* rng_state, int32_t* source_index, int32_t* counts) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) {
    ctx->error = "densify: no capacity reserved";
    return SPLAT_ERROR_STATE;
  }
  if (!p || !rng_state) {
    ctx->error = "densify: null params or rng_state";
    return SPLAT_ERROR_INVALID_ARGUMENT;
  }
  CKD(hipSetDevice(ctx->device), "set device");
  const int n = ctx->n;
  if (counts) counts[0] = counts[1] = counts[2] = 0;
  if (n == 0) return SPLAT_OK;
  const long long cap = ctx->cap;
  const long long floats = cap * (14 + ctx->rs);
  DensifyArgs a{};
  a.cap = cap;
  a.n = n;
  a.rest_stride = ctx->rs;
  a.has_room = n < p->max_gaussians ? 1 : 0;
  a.prune_oversized = iteration > p->opacity_reset_interval ? 1 : 0;
  a.threshold = p->densify_grad_threshold;
  a.size_threshold = p->percent_dense * extent;
  a.max_scale = p->prune_scale_fraction * extent;
  a.prune_opacity = p->prune_opacity;
  a.log_div = std::log(p->split_scale_divider);
  if (*rng_state == 0u) *rng_state = 1u;

  CKD(ctx->dens_flags.reserve(sizeof(uint32_t) * 5 * (size_t)n), "alloc");
  CKD(ctx->dens_a.reserve(sizeof(uint32_t) * 4 * (size_t)n), "alloc");
  CKD(ctx->dens_b.reserve(sizeof(uint32_t) * 64), "alloc");
  CKD(ctx->dens_c.reserve(sizeof(uint32_t) * 32 * kJumpLevels), "alloc");
  CKD(ctx->dens_src.reserve(sizeof(int) * (size_t)cap), "alloc");
  CKD(ctx->scratch_params.reserve(sizeof(float) * floats), "alloc");
  CKD(ctx->scratch_m.reserve(sizeof(float) * floats), "alloc");
  CKD(ctx->scratch_v.reserve(sizeof(float) * floats), "alloc");
  CKD(ctx->ensure_sort_workspace(std::max<long long>(n, 256LL * 4096)), "alloc");
  static const std::vector<uint32_t> table = jump_table();
  CKD(hipMemcpyAsync(ctx->dens_c.ptr, table.data(), 4 * table.size(), hipMemcpyHostToDevice, ctx->stream), "upload");

  const unsigned blocks = (unsigned)((n + 255) / 256);
  uint32_t* flags = ctx->dens_flags.u();
  uint32_t* offs = ctx->dens_a.u();
  uint32_t* scal = ctx->dens_b.u();
  CKD(launch_fill_u32(scal, 0u, 16, ctx->stream), "clear");
  hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(256), 0, ctx->stream, a, ctx->params.f(), ctx->stat_sg.f(),
                     ctx->stat_vc.i(), flags);
  CKD(hipGetLastError(), "classify");
  // survivors, accepted-addition ranks, split ranks, addition positions
  const int scan_src[4] = {0, 1, 3, 4};
  for (int k = 0; k < 4; ++k)
    CKD(launch_exclusive_scan(flags + (long long)scan_src[k] * n, nullptr, offs + (long long)k * n, n, scal + 8 + k,
                              ctx->sort_ws.scan, ctx->stream),
        "scan");
  uint32_t totals[4];
  CKD(hipMemcpyAsync(totals, scal + 8, sizeof(totals), hipMemcpyDeviceToHost, ctx->stream), "read");
  CKD(hipStreamSynchronize(ctx->stream), "sync");
  const uint32_t S = totals[0], total_np = totals[1], splits = totals[2];
  const long long room = std::max(0LL, (long long)p->max_gaussians - (long long)S);
  const uint32_t K = (uint32_t)std::min<long long>(total_np, room);
  if ((long long)S + K > cap) {
    ctx->error = "densify: the result exceeds the reserved capacity";
    return SPLAT_ERROR_CAPACITY;
  }
  // cut: position of the last accepted addition if the cap is reached
  const uint32_t host_scal[3] = {S, K, (S + (long long)K >= p->max_gaussians) ? 0u : 0xFFFFFFFFu};
  CKD(hipMemcpyAsync(scal, host_scal, sizeof(host_scal), hipMemcpyHostToDevice, ctx->stream), "upload");
  if (host_scal[2] == 0u)
    hipLaunchKernelGGL(k_cut, dim3(blocks), dim3(256), 0, ctx->stream, a, flags, offs, scal);
  hipLaunchKernelGGL(k_count_pruned, dim3(blocks), dim3(256), 0, ctx->stream, a, flags, offs, scal);
  CKD(hipGetLastError(), "count");
  hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, ctx->stream, a, ctx->params.f(), ctx->m.f(), ctx->v.f(),
                     ctx->scratch_params.f(), ctx->scratch_m.f(), ctx->scratch_v.f(), flags, offs, scal,
                     ctx->dens_c.u(), *rng_state, ctx->dens_src.i());
  CKD(hipGetLastError(), "scatter");
  const int n1 = (int)(S + K);
  std::swap(ctx->params, ctx->scratch_params);
  std::swap(ctx->m, ctx->scratch_m);
  std::swap(ctx->v, ctx->scratch_v);
  CKD(launch_fill_f32(ctx->grads.f(), 0.f, floats, ctx->stream), "clear");
  CKD(launch_fill_f32(ctx->sg.f(), 0.f, cap, ctx->stream), "clear");
  CKD(launch_fill_u32(ctx->vc.u(), 0u, cap, ctx->stream), "clear");
  ctx->grads_clear = true;
  CKD(launch_fill_f32(ctx->stat_sg.f(), 0.f, cap, ctx->stream), "clear");
  CKD(launch_fill_u32(ctx->stat_vc.u(), 0u, cap, ctx->stream), "clear");
  uint32_t result[4];
  CKD(hipMemcpyAsync(result, scal, sizeof(result), hipMemcpyDeviceToHost, ctx->stream), "read");
  if (source_index && n1 > 0)
    CKD(hipMemcpyAsync(source_index, ctx->dens_src.ptr, sizeof(int) * (size_t)n1, hipMemcpyDeviceToHost, ctx->stream),
        "read");
  CKD(hipStreamSynchronize(ctx->stream), "sync");
  // the stream advances 12 draws per split parent
  uint32_t state = *rng_state;
  uint64_t steps = 2ull * splits;  // in units of 6 draws
  for (int i = 0; i < kJumpLevels && steps; ++i, steps >>= 1)
    if (steps & 1ull) state = mat_apply(table.data() + 32 * i, state);
  *rng_state = state;
  if (counts) {
    counts[0] = 0;
    counts[1] = (int32_t)splits;
    counts[2] = (int32_t)result[3];
  }
  // clones = parents with one addition that are not splits
  if (counts) {
    std::vector<uint32_t> h((size_t)n * 2);
    CKD(hipMemcpy(h.data(), flags + (long long)3 * n, sizeof(uint32_t) * 2 * n, hipMemcpyDeviceToHost), "read");
    int clones = 0;
    for (int g = 0; g < n; ++g)
      if (!h[g] && h[(size_t)n + g] == 1) ++clones;
    counts[0] = clones;
  }
  ctx->n = n1;
  ctx->forward_done = ctx->loss_done = false;
  return SPLAT_OK;
}

splat_status splat_reset_opacity(splat_context* ctx, float ceiling) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) {
    ctx->error = "reset_opacity: no capacity reserved";
    return SPLAT_ERROR_STATE;
  }
  CKD(hipSetDevice(ctx->device), "set device");
  if (ctx->n > 0) {
    hipLaunchKernelGGL(k_reset_opacity, dim3((ctx->n + 255) / 256), dim3(256), 0, ctx->stream, ctx->params.f(),
                       ctx->m.f(), ctx->v.f(), ctx->cap, ctx->n, ceiling);
    CKD(hipGetLastError(), "reset opacity");
  }
  CKD(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_prune_and_sort(splat_context* ctx, float min_opacity) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) {
    ctx->error = "prune_and_sort: no capacity reserved";
    return SPLAT_ERROR_STATE;
  }
  CKD(hipSetDevice(ctx->device), "set device");
  const int n = ctx->n;
  if (n == 0) return SPLAT_OK;
  const long long cap = ctx->cap, floats = cap * (14 + ctx->rs);
  CKD(ctx->keys.reserve(4 * (size_t)n), "alloc");
  CKD(ctx->vals.reserve(4 * (size_t)n), "alloc");
  CKD(ctx->keys_alt.reserve(4 * (size_t)n), "alloc");
  CKD(ctx->vals_alt.reserve(4 * (size_t)n), "alloc");
  CKD(ctx->counts.reserve(4 * (size_t)n), "alloc");
  CKD(ctx->scalars.reserve(64), "alloc");
  CKD(ctx->scratch_params.reserve(sizeof(float) * floats), "alloc");
  CKD(ctx->ensure_sort_workspace(std::max<long long>(n, 256LL * 4096)), "alloc");
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_finish_keys, dim3(blocks), dim3(256), 0, ctx->stream, ctx->params.f(), cap, n, min_opacity,
                     ctx->keys.u(), ctx->vals.u(), ctx->counts.u());
  CKD(hipGetLastError(), "finish keys");
  CKD(ctx->offsets.reserve(4 * (size_t)n), "alloc");
  CKD(launch_exclusive_scan(ctx->counts.u(), nullptr, ctx->offsets.u(), n, ctx->scalars.u(), ctx->sort_ws.scan,
                            ctx->stream),
      "scan");
  bool alt = false;
  CKD(launch_radix_sort(ctx->keys.u(), ctx->vals.u(), ctx->keys_alt.u(), ctx->vals_alt.u(), n, 0, 32, ctx->sort_ws,
                        &alt, ctx->stream),
      "sort");
  uint32_t kept = 0;
  CKD(hipMemcpyAsync(&kept, ctx->scalars.u(), 4, hipMemcpyDeviceToHost, ctx->stream), "read");
  CKD(hipStreamSynchronize(ctx->stream), "sync");
  const uint32_t* order = alt ? ctx->vals_alt.u() : ctx->vals.u();
  if (kept > 0) {
    hipLaunchKernelGGL(k_gather, dim3((kept + 255) / 256), dim3(256), 0, ctx->stream, ctx->params.f(),
                       ctx->scratch_params.f(), cap, (int)kept, ctx->rs, order);
    CKD(hipGetLastError(), "gather");
  }
  std::swap(ctx->params, ctx->scratch_params);
  CKD(hipStreamSynchronize(ctx->stream), "sync");
  ctx->n = (int)kept;
  ctx->forward_done = ctx->loss_done = false;
  return SPLAT_OK;
}

}  // extern "C"
