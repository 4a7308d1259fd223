// Forward and adjoint compositing over 16x16 tiles.
//
// One 256-thread workgroup per (view, 16x16 tile) of the whole batch, one
// thread per pixel; all views of a batch go through one launch. Each
// workgroup walks its tile's near-to-far list in batches of 256 splats staged
// in LDS as packed float4 records (one coalesced record fetch per splat per
// tile, broadcast LDS reads per pixel).
//
// Forward: the engine's rasterizeSplats pixel loop (engine/native/cpp/src/
// splat/SplatRaster.cpp:277-313) in IEEE float with no contraction and the
// deterministic exponential, so every skip and termination decision -- and
// therefore the image -- matches the CPU checker bit for bit. It also stores
// the last contributor per pixel, so the adjoint needs no replay.
//
// Adjoint: rasterizeSplatsBackward's pixel half (:346-433). Walks each
// pixel's list back to front from its last contributor, recomputing the
// transmittance by division (nothing per splat and pixel is stored), and
// accumulates per splat dL/dcolour, dL/dopacity, dL/dmean, |dL/dmean| and
// dL/dconic. Per-pixel terms are summed across the wave64 by a transposed
// DPP butterfly (16 values in ~4 exchange levels, every lane ending with one
// value's wave total), the four waves' sums are combined in LDS in a fixed
// order, and
// one workgroup result per (splat, tile) pair then goes out either as a
// hardware float atomic (atomic mode) or as a plain store into that pair's
// own slot for a later fixed-order per-gaussian sum (deterministic mode).
// VALU-bound; the figure of merit is (pixel, splat) evaluations per second.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr int kBatch = 256;   // forward: splats staged per LDS batch
constexpr int kBatchB = 128;  // adjoint: smaller, so its per-wave partials fit 5 workgroups per CU
constexpr int kVals = 11;  // dcolour 3, dopacity, dmean 2, |dmean|, dconic 3, touched

__device__ __forceinline__ int find_view(const ViewDesc* views, int n_views, int tile) {
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].tile_base <= tile) v = k;
  return v;
}

__global__ __launch_bounds__(256) void k_forward(RasterParams p, const ViewDesc* __restrict__ views,
                                                 const uint32_t* __restrict__ ranges,
                                                 const uint32_t* __restrict__ pair_vals,
                                                 const SplatRec* __restrict__ rec, float* __restrict__ rgb,
                                                 float* __restrict__ T_out, int* __restrict__ last_out) {
  __shared__ float4 s_a[kBatch];  // mx, my, ca, cb
  __shared__ float4 s_b[kBatch];  // cc, opacity, r, g
  __shared__ float s_c[kBatch];   // b
  const int tile = blockIdx.x;
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  const int px = tx * kTile + (threadIdx.x & 15), py = ty * kTile + (threadIdx.x >> 4);
  const bool inside = px < V.width && py < V.height;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
  float T = 1.f, acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  int last = 0;
  bool done = !inside;
  for (uint32_t start = begin; start < end; start += kBatch) {
    if (__syncthreads_count(done) == (int)blockDim.x) break;
    const uint32_t e = start + threadIdx.x;
    if (e < end) {
      const SplatRec* r = rec + pair_vals[e];
      const float4 a = *reinterpret_cast<const float4*>(&r->mx);
      const float4 b = *reinterpret_cast<const float4*>(&r->cc);
      s_a[threadIdx.x] = a;
      s_b[threadIdx.x] = b;
      s_c[threadIdx.x] = r->cbl;
    }
    __syncthreads();
    const int cnt = (int)min((uint32_t)kBatch, end - start);
    if (!done) {
      for (int j = 0; j < cnt; ++j) {
        const float4 a = s_a[j];
        const float4 b = s_b[j];
        const float dx = fx - a.x, dy = fy - a.y;
        const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
        if (power > 0.f) continue;
        const float raw = b.y * det_exp(power);
        const float alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
        if (alpha < p.min_alpha) continue;
        const float w = T * alpha;
        acc0 += w * b.z;
        acc1 += w * b.w;
        acc2 += w * s_c[j];
        T *= 1.f - alpha;
        last = (int)(start + j + 1 - begin);
        if (T < p.cutoff) {
          done = true;
          break;
        }
      }
    }
    __syncthreads();
  }
  if (inside) {
    const long long pixel = V.pixel_base + (long long)py * V.width + px;
    rgb[3 * pixel + 0] = acc0 + T * p.bg[0];
    rgb[3 * pixel + 1] = acc1 + T * p.bg[1];
    rgb[3 * pixel + 2] = acc2 + T * p.bg[2];
    T_out[pixel] = T;
    last_out[pixel] = last;
  }
}

// Cross-lane exchange with the partner whose lane index differs in bit
// kBit (0..3), through DPP: quad_perm for bits 0 and 1, row_half_mirror then
// quad_perm [3,2,1,0] for bit 2, row_ror:8 for bit 3.
template <int kCtrl>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}
template <int kBit>
__device__ __forceinline__ float xchg(float v) {
  if constexpr (kBit == 0) return dpp<0xB1>(v);
  else if constexpr (kBit == 1) return dpp<0x4E>(v);
  else if constexpr (kBit == 2) return dpp<0x1B>(dpp<0x141>(v));
  else return dpp<0x128>(v);
}

// Sum 16 per-lane values over the wave64 in one transposed butterfly: each
// level exchanges half of the values a lane holds with its partner and keeps
// the other half, so 16 values -> 1 in four DPP levels (lane bits 0-3), then
// two cross-row steps. Every lane ends with the wave total of one value: slot
// bitrev4(lane & 15). The summation order is fixed, so it is deterministic.
__device__ __forceinline__ float butterfly16(const float (&v)[16], int lane) {
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4, b3 = lane & 8;
  float w[8], x[4], y[2];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float send = b0 ? v[k] : v[k + 8];
    const float keep = b0 ? v[k + 8] : v[k];
    w[k] = keep + xchg<0>(send);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float send = b1 ? w[k] : w[k + 4];
    const float keep = b1 ? w[k + 4] : w[k];
    x[k] = keep + xchg<1>(send);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const float send = b2 ? x[k] : x[k + 2];
    const float keep = b2 ? x[k + 2] : x[k];
    y[k] = keep + xchg<2>(send);
  }
  const float send = b3 ? y[0] : y[1];
  const float keep = b3 ? y[1] : y[0];
  float z = keep + xchg<3>(send);
  z += __shfl_xor(z, 16, 64);
  z += __shfl_xor(z, 32, 64);
  return z;
}

__device__ __forceinline__ int butterfly_slot(int lane) {
  return ((lane & 1) << 3) | ((lane & 2) << 1) | ((lane & 4) >> 1) | ((lane & 8) >> 3);
}

__global__ __launch_bounds__(256) void k_backward(RasterParams p, const ViewDesc* __restrict__ views,
                                                  const uint32_t* __restrict__ ranges,
                                                  const uint32_t* __restrict__ pair_vals,
                                                  const SplatRec* __restrict__ rec,
                                                  const float* __restrict__ dL_drgb, const float* __restrict__ T_in,
                                                  const int* __restrict__ last_in, float* __restrict__ screen,
                                                  const uint32_t* __restrict__ pair_base, float* __restrict__ partial,
                                                  int deterministic) {
  __shared__ float4 s_a[kBatchB];
  __shared__ float4 s_b[kBatchB];
  __shared__ float s_c[kBatchB];
  __shared__ uint32_t s_dst[kBatchB];          // slot (atomic) or pair id (deterministic)
  __shared__ float s_part[4][kBatchB][kVals];  // per-wave sums
  __shared__ int s_maxlast;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int my_slot = butterfly_slot(lane);
  const bool writer = lane < 16 && my_slot < kVals;
  const int tile = blockIdx.x;
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  const int px = tx * kTile + (tid & 15), py = ty * kTile + (tid >> 4);
  const bool inside = px < V.width && py < V.height;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
  float dL0 = 0.f, dL1 = 0.f, dL2 = 0.f, T_final = 1.f;
  int lastp = 0;
  if (inside) {
    const long long pixel = V.pixel_base + (long long)py * V.width + px;
    dL0 = dL_drgb[3 * pixel];
    dL1 = dL_drgb[3 * pixel + 1];
    dL2 = dL_drgb[3 * pixel + 2];
    T_final = T_in[pixel];
    lastp = last_in[pixel];
  }
  if (tid == 0) s_maxlast = 0;
  __syncthreads();
  if (lastp > 0) atomicMax(&s_maxlast, lastp);
  __syncthreads();
  const uint32_t stop = begin + (uint32_t)s_maxlast;
  // Pairs past every pixel's last contributor get zero partials.
  if (deterministic) {
    for (uint32_t e = stop + tid; e < end; e += blockDim.x) {
      const uint32_t slot = pair_vals[e];
      const SplatRec& r = rec[slot];
      const uint32_t pid = pair_base[slot] + (uint32_t)((ty - r.y0) * (r.x1 - r.x0 + 1) + (tx - r.x0));
      float* dst = partial + (long long)pid * kScreenStride;
#pragma unroll
      for (int k = 0; k < kScreenStride; ++k) dst[k] = 0.f;
    }
  }
  const float bg_term = p.bg[0] * dL0 + p.bg[1] * dL1 + p.bg[2] * dL2;
  float T = T_final;
  float suf0 = 0.f, suf1 = 0.f, suf2 = 0.f;
  float last_alpha = 0.f, lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;
  uint32_t hi = stop;
  while (hi > begin) {
    const uint32_t lo = (hi - begin) > (uint32_t)kBatchB ? hi - kBatchB : begin;
    const int cnt = (int)(hi - lo);
    __syncthreads();
    if (tid < cnt) {
      const uint32_t e = lo + tid;
      const uint32_t slot = pair_vals[e];
      const SplatRec* r = rec + slot;
      s_a[tid] = *reinterpret_cast<const float4*>(&r->mx);
      s_b[tid] = *reinterpret_cast<const float4*>(&r->cc);
      s_c[tid] = r->cbl;
      if (deterministic)
        s_dst[tid] = pair_base[slot] + (uint32_t)((ty - r->y0) * (r->x1 - r->x0 + 1) + (tx - r->x0));
      else
        s_dst[tid] = slot;
    }
    __syncthreads();
    for (int j = cnt - 1; j >= 0; --j) {
      const uint32_t e = lo + j;
      bool contrib = inside && (int)(e - begin) < lastp;
      float val[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) val[k] = 0.f;
      const float4 a = s_a[j];
      const float4 b = s_b[j];
      // the forward's decisions, recomputed with its exact float operations
      float dx = 0.f, dy = 0.f, G = 0.f, raw = 0.f, alpha = 0.f;
      if (contrib) {
        dx = fx - a.x;
        dy = fy - a.y;
        const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
        if (power > 0.f) {
          contrib = false;
        } else {
          G = det_exp(power);
          raw = b.y * G;
          alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
          if (alpha < p.min_alpha) contrib = false;
        }
      }
      if (contrib) {
#pragma clang fp contract(fast)
        const float c2 = s_c[j];
        const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
        T = T * inv;
        const float weight = alpha * T;
        val[0] = weight * dL0;
        val[1] = weight * dL1;
        val[2] = weight * dL2;
        suf0 = last_alpha * lc0 + (1.f - last_alpha) * suf0;
        suf1 = last_alpha * lc1 + (1.f - last_alpha) * suf1;
        suf2 = last_alpha * lc2 + (1.f - last_alpha) * suf2;
        lc0 = b.z;
        lc1 = b.w;
        lc2 = c2;
        float dL_dalpha = ((b.z - suf0) * dL0 + (b.w - suf1) * dL1 + (c2 - suf2) * dL2) * T;
        last_alpha = alpha;
        dL_dalpha -= T_final * inv * bg_term;
        if (raw < kMaxAlpha) {
          val[3] = G * dL_dalpha;
          const float dL_dweight = b.y * dL_dalpha;
          const float gx = G * dx, gy = G * dy;
          const float mx = dL_dweight * (-gx * a.z - gy * a.w);
          const float my = dL_dweight * (-gy * b.x - gx * a.w);
          val[4] = -mx;
          val[5] = -my;
          val[6] = __builtin_sqrtf(mx * mx + my * my);
          val[7] = -0.5f * dL_dweight * gx * dx;
          val[8] = -dL_dweight * gx * dy;
          val[9] = -0.5f * dL_dweight * gy * dy;
          val[10] = 1.f;
        }
      }
      float z = 0.f;
      if (__ballot(contrib) != 0ull) z = butterfly16(val, lane);
      if (writer) s_part[wave][j][my_slot] = z;
    }
    __syncthreads();
    if (tid < cnt) {
      float sum[kVals];
#pragma unroll
      for (int k = 0; k < kVals; ++k) sum[k] = ((s_part[0][tid][k] + s_part[1][tid][k]) + s_part[2][tid][k]) + s_part[3][tid][k];
      if (deterministic) {
        float* dst = partial + (long long)s_dst[tid] * kScreenStride;
#pragma unroll
        for (int k = 0; k < kVals; ++k) dst[k] = sum[k];
        dst[kVals] = 0.f;
      } else {
        float* dst = screen + (long long)s_dst[tid] * kScreenStride;
#pragma unroll
        for (int k = 0; k < kVals; ++k)
          if (sum[k] != 0.f) unsafeAtomicAdd(dst + k, sum[k]);
      }
    }
    hi = lo;
  }
}

}  // namespace

hipError_t launch_forward(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                          const uint32_t* pair_vals, const SplatRec* rec, float* rgb, float* T, int* last,
                          hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_forward, dim3(p.total_tiles), dim3(256), 0, stream, p, views, ranges, pair_vals, rec, rgb, T,
                     last);
  return hipGetLastError();
}

hipError_t launch_backward(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                           const uint32_t* pair_vals, const SplatRec* rec, const float* dL_drgb, const float* T,
                           const int* last, float* screen, const uint32_t* pair_base, float* partial,
                           bool deterministic, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_backward, dim3(p.total_tiles), dim3(256), 0, stream, p, views, ranges, pair_vals, rec, dL_drgb,
                     T, last, screen, pair_base, partial, deterministic ? 1 : 0);
  return hipGetLastError();
}

}  // namespace splat
