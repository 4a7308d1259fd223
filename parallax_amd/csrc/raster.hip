// Forward and adjoint compositing over 16x16 tiles.
//
// One 256-thread workgroup per (view, 16x16 tile) of the whole batch; each of
// its four waves owns one 8x8 quadrant of the tile, one lane per pixel. All
// views of a batch go through one launch, and workgroups are mapped to tiles
// so that each XCD works through one contiguous run of tiles.
//
// A wave skips every splat of its tile's list whose cull box (the box of the
// ellipse where alpha >= min_alpha can hold, computed by the projection)
// misses its quadrant's pixel centres. Such a splat is skipped by every pixel
// of the quadrant anyway, so nothing changes but the work: about half of the
// (wave, splat) steps of a 16x16 tile list fall away at 1080p.
//
// Forward: the engine's rasterizeSplats pixel loop (engine/native/cpp/src/
// splat/SplatRaster.cpp:277-313) in IEEE float with no contraction and the
// deterministic exponential, so every skip and termination decision -- and
// therefore the image -- matches the CPU checker bit for bit. It also stores
// the last contributor per pixel, so the adjoint needs no replay. Each wave
// streams its tile's list in chunks of 64 splats: one lane per splat fetches
// the 48-byte record, tests the cull box, and the survivors are compacted in
// order into the wave's own LDS slice (ballot + mbcnt). Waves never wait on
// each other, and a wave stops as soon as all of its 64 pixels terminate.
//
// Adjoint: rasterizeSplatsBackward's pixel half (:346-433). Walks each
// pixel's list back to front from its last contributor, recomputing the
// transmittance by division (nothing per splat and pixel is stored), and
// accumulates per splat dL/dcolour, dL/dopacity, dL/dmean, |dL/dmean| and
// dL/dconic. The workgroup stages the list in batches of 128 records and
// marks, per splat, the quadrants whose cull box it meets below that
// quadrant's last contributor; each wave then visits only its marked splats,
// in reverse order. Per-pixel terms are summed across the wave64 by a
// transposed DPP butterfly (16 values in ~4 exchange levels, every lane
// ending with one value's wave total), the marked waves' sums are combined in
// LDS in a fixed order, and one workgroup result per (splat, tile) pair then
// goes out either as a hardware float atomic (atomic mode) or as a plain
// store into that pair's own slot for a later fixed-order per-gaussian sum
// (deterministic mode). VALU-bound; the figure of merit is (pixel, splat)
// evaluations per second.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr int kChunk = 64;    // forward: splats fetched per wave step
constexpr int kBatchB = 128;  // adjoint: splats staged per workgroup batch
constexpr int kVals = 11;  // dcolour 3, dopacity, dmean 2, |dmean|, dconic 3, touched

__device__ __forceinline__ int find_view(const ViewDesc* views, int n_views, int tile) {
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].tile_base <= tile) v = k;
  return v;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int lane_rank(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The quadrant of a tile that wave `wave` owns, as pixel-centre bounds clipped
// to the image. Returns false when the quadrant lies outside the image.
struct Quad {
  int x0, y0;
  float bx0, bx1, by0, by1;
};
__device__ __forceinline__ bool quad_of(const ViewDesc& V, int tx, int ty, int wave, Quad* q) {
  q->x0 = tx * kTile + 8 * (wave & 1);
  q->y0 = ty * kTile + 8 * (wave >> 1);
  if (q->x0 >= V.width || q->y0 >= V.height) return false;
  q->bx0 = (float)q->x0 + 0.5f;
  q->by0 = (float)q->y0 + 0.5f;
  q->bx1 = (float)min(q->x0 + 7, V.width - 1) + 0.5f;
  q->by1 = (float)min(q->y0 + 7, V.height - 1) + 0.5f;
  return true;
}
__device__ __forceinline__ bool box_hits(float mx, float my, float ex, float ey, const Quad& q) {
  return mx + ex >= q.bx0 && mx - ex <= q.bx1 && my + ey >= q.by0 && my - ey <= q.by1;
}

__global__ __launch_bounds__(256) void k_forward(RasterParams p, const ViewDesc* __restrict__ views,
                                                 const uint32_t* __restrict__ ranges,
                                                 const uint32_t* __restrict__ pair_vals,
                                                 const SplatRec* __restrict__ rec, float* __restrict__ rgb,
                                                 float* __restrict__ T_out, int* __restrict__ last_out) {
  __shared__ float4 s_a[4][kChunk];  // mx, my, ca, cb
  __shared__ float4 s_b[4][kChunk];  // cc, opacity, r, g
  __shared__ float2 s_c[4][kChunk];  // b, list index (int bits)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = xcd_remap(blockIdx.x, p.total_tiles);
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  Quad q;
  if (!quad_of(V, tx, ty, wave, &q)) return;  // no barrier in this kernel: waves are independent
  const int px = q.x0 + (lane & 7), py = q.y0 + (lane >> 3);
  const bool inside = px < V.width && py < V.height;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
  float T = 1.f, acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  int last = 0;
  bool done = !inside;
  for (uint32_t base = begin; base < end; base += kChunk) {
    if (__ballot(!done) == 0ull) break;
    const uint32_t e = base + lane;
    bool hit = false;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a;
    if (e < end) {
      const float4* r = reinterpret_cast<const float4*>(rec + pair_vals[e]);
      a = r[0];
      b = r[1];
      c = r[2];
      hit = box_hits(a.x, a.y, c.y, c.z, q);
    }
    const uint64_t m = __ballot(hit);
    if (hit) {
      const int pos = lane_rank(m);
      s_a[wave][pos] = a;
      s_b[wave][pos] = b;
      s_c[wave][pos] = make_float2(c.x, __int_as_float((int)(e - begin)));
    }
    wave_sync_lds();
    const int cnt = __popcll(m);
    if (!done) {
      for (int j = 0; j < cnt; ++j) {
        const float4 sa = s_a[wave][j];
        const float4 sb = s_b[wave][j];
        const float dx = fx - sa.x, dy = fy - sa.y;
        const float power = -0.5f * (sa.z * dx * dx + sb.x * dy * dy) - sa.w * dx * dy;
        if (power > 0.f) continue;
        const float raw = sb.y * det_exp_nonpos(power);
        const float alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
        if (alpha < p.min_alpha) continue;
        const float2 sc = s_c[wave][j];
        const float w = T * alpha;
        acc0 += w * sb.z;
        acc1 += w * sb.w;
        acc2 += w * sc.x;
        T *= 1.f - alpha;
        last = __float_as_int(sc.y) + 1;
        if (T < p.cutoff) {
          done = true;
          break;
        }
      }
    }
    wave_sync_lds();
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
                                                  uint32_t p_cap, int deterministic) {
  __shared__ float4 s_a[kBatchB];
  __shared__ float4 s_b[kBatchB];
  __shared__ float s_c[kBatchB];
  __shared__ uint32_t s_dst[kBatchB];          // slot (atomic) or pair id (deterministic)
  __shared__ uint32_t s_mask[kBatchB];         // quadrants that visit the splat
  __shared__ float s_part[4][kBatchB][kVals];  // per-wave sums (valid where s_mask says)
  __shared__ int s_stop[4];                    // per-quadrant last contributor
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int my_slot = butterfly_slot(lane);
  const bool writer = lane < 16 && my_slot < kVals;
  const int tile = xcd_remap(blockIdx.x, p.total_tiles);
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  // the four quadrants' cull bounds, for the staging threads
  Quad quads[4];
  bool qvalid[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) qvalid[k] = quad_of(V, tx, ty, k, &quads[k]);
  const int px = tx * kTile + 8 * (wave & 1) + (lane & 7), py = ty * kTile + 8 * (wave >> 1) + (lane >> 3);
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
  if (tid < 4) s_stop[tid] = 0;
  __syncthreads();
  if (lastp > 0) atomicMax(&s_stop[wave], lastp);
  __syncthreads();
  const int stops[4] = {s_stop[0], s_stop[1], s_stop[2], s_stop[3]};
  const uint32_t stop = begin + (uint32_t)max(max(stops[0], stops[1]), max(stops[2], stops[3]));
  // Pairs past every pixel's last contributor get zero partials.
  if (deterministic) {
    for (uint32_t e = stop + tid; e < end; e += blockDim.x) {
      const uint32_t slot = pair_vals[e];
      const SplatRec& r = rec[slot];
      const uint32_t pid = pair_base[slot] + (uint32_t)((ty - r.y0) * (r.x1 - r.x0 + 1) + (tx - r.x0));
      if (pid >= p_cap) continue;  // only when the step's capacity check has flagged it
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
      const float4* r4 = reinterpret_cast<const float4*>(r);
      const float4 a = r4[0], b = r4[1], c = r4[2];
      s_a[tid] = a;
      s_b[tid] = b;
      s_c[tid] = c.x;
      const int idx = (int)(e - begin);
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (qvalid[k] && idx < stops[k] && box_hits(a.x, a.y, c.y, c.z, quads[k])) m |= 1u << k;
      s_mask[tid] = m;
      if (deterministic)
        s_dst[tid] = pair_base[slot] + (uint32_t)((ty - r->y0) * (r->x1 - r->x0 + 1) + (tx - r->x0));
      else
        s_dst[tid] = slot;
    }
    __syncthreads();
    // this wave's splats of the batch, visited from the back
    uint64_t mine[2];
    mine[0] = __ballot(lane < cnt && ((s_mask[lane] >> wave) & 1u));
    mine[1] = __ballot(lane + 64 < cnt && ((s_mask[lane + 64] >> wave) & 1u));
#pragma unroll
    for (int h = 1; h >= 0; --h) {
      uint64_t bits = mine[h];
      while (bits) {
        const int bit = 63 - __builtin_clzll(bits);
        bits &= ~(1ull << bit);
        const int j = 64 * h + bit;
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
            G = det_exp_nonpos(power);
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
            val[6] = __builtin_amdgcn_sqrtf(mx * mx + my * my);  // densify statistic only: hardware sqrt
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
    }
    __syncthreads();
    if (tid < cnt) {
      const uint32_t m = s_mask[tid];
      float sum[kVals];
#pragma unroll
      for (int k = 0; k < kVals; ++k) sum[k] = 0.f;
      // fixed order over the quadrants that visited the splat
#pragma unroll
      for (int w = 0; w < 4; ++w)
        if ((m >> w) & 1u) {
#pragma unroll
          for (int k = 0; k < kVals; ++k) sum[k] += s_part[w][tid][k];
        }
      if (deterministic) {
        if (s_dst[tid] < p_cap) {
          float* dst = partial + (long long)s_dst[tid] * kScreenStride;
#pragma unroll
          for (int k = 0; k < kVals; ++k) dst[k] = sum[k];
          dst[kVals] = 0.f;
        }
      } else if (m) {
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
                           uint32_t p_cap, bool deterministic, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_backward, dim3(p.total_tiles), dim3(256), 0, stream, p, views, ranges, pair_vals, rec, dL_drgb,
                     T, last, screen, pair_base, partial, p_cap, deterministic ? 1 : 0);
  return hipGetLastError();
}

}  // namespace splat
