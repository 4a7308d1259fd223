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
#include <algorithm>

#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

// det_exp_nonpos(min(power, 0)), the compositing falloff: one med3 clamps the
// argument to [-80, 0]. For every non-NaN power (+-inf, +-0 and -80 included)
// this is the value of the two-step clamp; the exponential of -0 and of +0 are
// both exactly 1. A NaN power is outside that claim: med3 then returns the
// min of the other two operands (-80, so alpha ~ 0 and the splat is skipped)
// where the CPU form gives alpha 0.99. A NaN power needs a non-finite mean or
// conic, i.e. non-finite parameters, for which the engine's own binning is
// undefined (floor of NaN converted to int, SplatRaster.cpp:104-151).
__device__ __forceinline__ float falloff(float power) {
  const float e = splat_det::exp_core(__builtin_amdgcn_fmed3f(power, -80.0f, 0.0f));
  return power <= -80.0f ? 0.0f : e;
}

// The falloff without its power <= -80 test: for power <= -80 it gives
// e^-80 where falloff gives 0. Either way alpha = opacity * G <= 1.9e-35, below
// any min_alpha >= 1e-30, so a compositing step that goes on to test
// alpha < min_alpha skips the splat at that pixel with either value, and G is
// read nowhere else. With smaller min_alpha the loops use falloff itself.
//
// Nor is the clamp at 0 needed: a step that keeps the splat has power <= 0,
// where max(power, -80) is the two-sided clamp (NaN power, too, becomes -80,
// as with med3). Where power > 0 the pixel is skipped and G is garbage that
// every use masks out by a select (never a multiply by zero).
__device__ __forceinline__ float falloff_above_cut(float power) {
  return splat_det::exp_core(fmaxf(power, -80.0f));
}

// Lane masks (64-bit, in scalar registers) of per-lane float and integer
// compares, and selects by such a mask. The compositing loops combine their
// decisions as masks: a ballot of a combined condition, or a select on one,
// would copy the mask through a vector register first.
constexpr int kFcmpOlt = 4;   // ordered less than
constexpr int kFcmpUge = 11;  // unordered or greater or equal: !(a < b)
constexpr int kFcmpUle = 13;  // unordered or less or equal: !(a > b)
constexpr int kIcmpSlt = 40;  // signed less than
__device__ __forceinline__ float select_lanes(uint64_t lanes, float if_set, float if_clear) {
  float r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(if_clear), "v"(if_set), "s"(lanes));
  return r;
}
__device__ __forceinline__ int select_lanes(uint64_t lanes, int if_set, int if_clear) {
  int r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(if_clear), "v"(if_set), "s"(lanes));
  return r;
}
__device__ __forceinline__ float keep_lanes(uint64_t lanes, float x) {
  float r;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(x), "s"(lanes));
  return r;
}

constexpr int kChunk = 64;    // forward: splats fetched per wave step

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

// Whether any point of the rectangle [x0, x1] x [y0, y1] (pixel centres of a
// quadrant) can lie in the splat's alpha >= min_alpha ellipse
// Q(d) = a dx^2 + 2 b dx dy + c dy^2 <= q, q = 2 ln(opacity / min_alpha).
// The minimum of the positive-definite Q over the rectangle is 0 when the mean
// is inside, else on an edge, where each edge's minimiser is the clamped
// stationary point. Conservative: the continuous minimum is at most the value
// at any pixel centre, q is raised by 1% + 1e-3 (the hardware log, the float
// minimum and the compositing loop's own float Q all err far less for
// kappa = ac / det <= 1000), and a more elongated splat is always kept (the
// cull box alone decides). A splat it rejects is skipped by every pixel of the
// quadrant anyway, so results do not change.
__device__ __forceinline__ bool ellipse_meets_rect(float a, float b, float c, float opacity, float min_alpha,
                                                   float x0, float x1, float y0, float y1) {
  const float det = a * c - b * b;
  if (!(det > 0.f) || a * c > 1000.f * det) return true;
  const float q = 2.f * 0.69314718f * (__builtin_amdgcn_logf(opacity) - __builtin_amdgcn_logf(min_alpha));
  const float q_hi = q * 1.01f + 1e-3f;
  if (x0 <= 0.f && x1 >= 0.f && y0 <= 0.f && y1 >= 0.f) return q_hi >= 0.f;
  const float ia = __builtin_amdgcn_rcpf(a), ic = __builtin_amdgcn_rcpf(c);
  float best = __builtin_inff();
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float xe = e ? x1 : x0, ye = e ? y1 : y0;
    const float yc = fminf(fmaxf(-b * xe * ic, y0), y1);
    const float xc = fminf(fmaxf(-b * ye * ia, x0), x1);
    best = fminf(best, a * xe * xe + 2.f * b * xe * yc + c * yc * yc);
    best = fminf(best, a * xc * xc + 2.f * b * xc * ye + c * ye * ye);
  }
  return best <= q_hi;
}

// The pixel-centre bounds of a tile's two quadrant columns and two quadrant
// rows, as wave-uniform values in scalar registers (the float conversions run
// on the vector unit; readfirstlane moves the results back to SGPRs so they
// do not hold vector registers across the compositing loop).
struct TileQuads {
  float x0[2], x1[2], y0[2], y1[2];
  bool xv[2], yv[2];  // column / row inside the image
};
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ TileQuads tile_quads(const ViewDesc& V, int tx, int ty) {
  TileQuads t;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int x0 = tx * kTile + 8 * c, y0 = ty * kTile + 8 * c;
    t.xv[c] = x0 < V.width;
    t.yv[c] = y0 < V.height;
    t.x0[c] = uniform_f((float)x0 + 0.5f);
    t.x1[c] = uniform_f((float)min(x0 + 7, V.width - 1) + 0.5f);
    t.y0[c] = uniform_f((float)y0 + 0.5f);
    t.y1[c] = uniform_f((float)min(y0 + 7, V.height - 1) + 0.5f);
  }
  return t;
}
// Bit k set: the cull box meets quadrant k (k = 2 * row + column).
__device__ __forceinline__ uint32_t quad_hits(float mx, float my, float ex, float ey, const TileQuads& t) {
  const bool c0 = t.xv[0] && mx + ex >= t.x0[0] && mx - ex <= t.x1[0];
  const bool c1 = t.xv[1] && mx + ex >= t.x0[1] && mx - ex <= t.x1[1];
  const bool r0 = t.yv[0] && my + ey >= t.y0[0] && my - ey <= t.y1[0];
  const bool r1 = t.yv[1] && my + ey >= t.y0[1] && my - ey <= t.y1[1];
  return (uint32_t)(r0 && c0) | (uint32_t)(r0 && c1) << 1 | (uint32_t)(r1 && c0) << 2 | (uint32_t)(r1 && c1) << 3;
}

// A staged list entry of the forward: the compositing fields of its record
// and its list position + 1 (one 48-byte slot, so the loads share one address).
struct alignas(16) FwdStaged {
  float4 a;  // mx, my, conic a, conic b
  float4 b;  // conic c, opacity, r, g
  float c;   // b
  int pos;   // list position + 1
};

// One wave's forward over one 8x8 quadrant (0..3, row-major) of one tile. The
// LDS slice is the wave's own. kLowCut: the falloff with its power <= -80
// test (min_alpha below 1e-30).
template <bool kLowCut>
__device__ __forceinline__ void forward_quad(const RasterParams& p, const ViewDesc* __restrict__ views,
                                             const uint32_t* __restrict__ ranges,
                                             const uint32_t* __restrict__ pair_vals,
                                             const SplatRec* __restrict__ rec, float* __restrict__ rgb,
                                             float* __restrict__ T_out, int* __restrict__ last_out, int tile,
                                             int quadrant, FwdStaged* s_st) {
  const int lane = threadIdx.x & 63;
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  Quad q;
  if (!quad_of(V, tx, ty, quadrant, &q)) return;
  const int px = q.x0 + (lane & 7), py = q.y0 + (lane >> 3);
  const bool inside = px < V.width && py < V.height;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const uint32_t begin = ranges[2 * tile], end = ranges[2 * tile + 1];
  float T = 1.f, acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  int last = 0;
  // terminated pixels as a lane mask (pixels outside the image start there)
  uint64_t done = __ballot(!inside);
  for (uint32_t base = begin; base < end; base += kChunk) {
    if (done == ~0ull) break;
    const uint32_t e = base + lane;
    bool hit = false;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a;
    if (e < end) {
      const float4* r = reinterpret_cast<const float4*>(rec + pair_vals[e]);
      a = r[0];
      b = r[1];
      c = r[2];
      hit = box_hits(a.x, a.y, c.y, c.z, q) &&
            ellipse_meets_rect(a.z, a.w, b.x, b.y, p.min_alpha, q.bx0 - a.x, q.bx1 - a.x, q.by0 - a.y, q.by1 - a.y);
    }
    const uint64_t m = __ballot(hit);
    if (hit) {
      const int pos = lane_rank(m);
      s_st[pos].a = a;
      s_st[pos].b = b;
      s_st[pos].c = c.x;
      s_st[pos].pos = (int)(e - begin) + 1;
    }
    wave_sync_lds();
    const int cnt = __builtin_amdgcn_readfirstlane(__popcll(m));  // wave-uniform
    // Branch-free over the chunk: every lane evaluates every splat and keeps
    // the result only where the engine's loop would (not yet terminated,
    // power <= 0, alpha >= min_alpha); a kept step is the same float
    // operations on the same operands, a dropped one adds an exact zero.
    // (the counter passes through readfirstlane so that loop strength
    // reduction cannot turn it into a vector down-counter with a carry)
    for (int j = 0; j < cnt; j = __builtin_amdgcn_readfirstlane(j + 1)) {
      const float4 sa = s_st[j].a;
      const float4 sb = s_st[j].b;
      const float sc = s_st[j].c;
      const int spos = s_st[j].pos;
      const float dx = fx - sa.x, dy = fy - sa.y;
      const float power = -0.5f * (sa.z * dx * dx + sb.x * dy * dy) - sa.w * dx * dy;
      const float raw = sb.y * (kLowCut ? falloff(power) : falloff_above_cut(power));
      const float alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
      const uint64_t use = ~done & __builtin_amdgcn_fcmpf(power, 0.f, kFcmpUle) &
                           __builtin_amdgcn_fcmpf(alpha, p.min_alpha, kFcmpUge);
      const float w = keep_lanes(use, T * alpha);
      acc0 += w * sb.z;
      acc1 += w * sb.w;
      acc2 += w * sc;
      T = select_lanes(use, T * (1.f - alpha), T);
      last = select_lanes(use, spos, last);
      done |= use & __builtin_amdgcn_fcmpf(T, p.cutoff, kFcmpOlt);
      // leave once all the wave's pixels have terminated (a scalar compare
      // of the lane mask after every splat)
      if (done == ~0ull) break;
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

// Work units of the forward launch shapes. kShape 0: a 256-thread workgroup
// per 16x16 tile, its waves on the four quadrants. kShape 1: a one-wave
// workgroup per quadrant (8x8). kShape 2: a 256-thread workgroup per 32x8
// strip -- the top or bottom halves of two side-by-side tiles (units of tiles
// in odd columns are empty). The waves never wait for each other, so every
// shape gives the same image; they differ in scheduling and cache sharing.
template <int kShape>
struct FwdShape;
template <>
struct FwdShape<0> {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ int units(int tiles) { return tiles; }
  static __device__ __forceinline__ bool wave_work(const ViewDesc* views, int n_views, int unit, int wave, int* tile,
                                                   int* quadrant) {
    (void)views;
    (void)n_views;
    *tile = unit;
    *quadrant = wave;
    return true;
  }
};
template <>
struct FwdShape<1> {
  static constexpr int kThreads = 64;
  static __device__ __forceinline__ int units(int tiles) { return 4 * tiles; }
  static __device__ __forceinline__ bool wave_work(const ViewDesc* views, int n_views, int unit, int wave, int* tile,
                                                   int* quadrant) {
    (void)views;
    (void)n_views;
    (void)wave;
    *tile = unit >> 2;
    *quadrant = unit & 3;
    return true;
  }
};
template <>
struct FwdShape<2> {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ int units(int tiles) { return 2 * tiles; }
  static __device__ __forceinline__ bool wave_work(const ViewDesc* views, int n_views, int unit, int wave, int* tile,
                                                   int* quadrant) {
    const int left = unit >> 1, half = unit & 1;
    const ViewDesc& V = views[find_view(views, n_views, left)];
    const int tx = (left - V.tile_base) % V.cols;
    if (tx & 1) return false;                          // units start on even columns
    if ((wave >> 1) && tx + 1 >= V.cols) return false;  // no right neighbour
    *tile = left + (wave >> 1);
    *quadrant = 2 * half + (wave & 1);
    return true;
  }
};

// kQueue false: one workgroup per unit, units spread over the XCDs in
// contiguous runs. kQueue true: a persistent grid; each workgroup takes the
// next unit from a counter over `order` (tiles, longest list first), so the
// longest tiles start first and the launch ends on short ones.
template <int kShape, bool kQueue, bool kLowCut>
__global__ __launch_bounds__(FwdShape<kShape>::kThreads) void k_forward(
    RasterParams p, const ViewDesc* __restrict__ views, const uint32_t* __restrict__ ranges,
    const uint32_t* __restrict__ pair_vals, const SplatRec* __restrict__ rec, float* __restrict__ rgb,
    float* __restrict__ T_out, int* __restrict__ last_out, const uint32_t* __restrict__ order,
    uint32_t* __restrict__ counter) {
  using Sh = FwdShape<kShape>;
  constexpr int kWaves = Sh::kThreads / 64;
  __shared__ FwdStaged s_st[kWaves][kChunk];
  __shared__ int s_unit;
  const int wave = threadIdx.x >> 6;
  const int n_units = Sh::units(p.total_tiles);
  if (!kQueue) {
    const int unit = xcd_remap(blockIdx.x, n_units);
    int tile, quadrant;
    if (Sh::wave_work(views, p.n_views, unit, wave, &tile, &quadrant))
      forward_quad<kLowCut>(p, views, ranges, pair_vals, rec, rgb, T_out, last_out, tile, quadrant, s_st[wave]);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) s_unit = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int idx = s_unit;
    __syncthreads();
    if (idx >= n_units) break;
    // the idx-th unit: tiles longest first, a tile's units together
    const int per = n_units / p.total_tiles;
    const int unit = (int)order[idx / per] * per + idx % per;
    int tile, quadrant;
    if (Sh::wave_work(views, p.n_views, unit, wave, &tile, &quadrant))
      forward_quad<kLowCut>(p, views, ranges, pair_vals, rec, rgb, T_out, last_out, tile, quadrant, s_st[wave]);
  }
}

// DPP move within a row of 16 lanes (quad_perm controls only here).
template <int kCtrl>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}

// The adjoint's per-splat reduction goes through LDS: each lane writes its 10
// accumulated terms into a [term][lane] array, then lanes 4i..4i+3 each load
// a quarter of term i's row (16 lanes, four 16-byte loads), sum it in a fixed
// tree, and combine the four quarters with two quad DPP adds (commutative, so
// all four lanes hold the same bits). Lane 4i then holds term i's wave total.
// Rows are 68 floats apart, which keeps both the row writes and the
// quarter-row loads free of LDS bank conflicts.
constexpr int kAcc = 10;       // dcolour 3, dopacity, dmean 2, |dmean|, dconic 3
constexpr int kRedPitch = 68;  // floats per term row

struct Quarter {
  float4 q0, q1, q2, q3;
};
__device__ __forceinline__ Quarter lds_quarter_load(const float* s_red, int lane) {
  const int i = min(lane >> 2, kAcc - 1);  // lanes 40..63 repeat term 9's loads; their sums are not used
  const float4* src = reinterpret_cast<const float4*>(s_red + i * kRedPitch + (lane & 3) * 16);
  return Quarter{src[0], src[1], src[2], src[3]};
}
__device__ __forceinline__ float quarter_sum(const Quarter& r) {
  const float t0 = (r.q0.x + r.q0.y) + (r.q0.z + r.q0.w), t1 = (r.q1.x + r.q1.y) + (r.q1.z + r.q1.w);
  const float t2 = (r.q2.x + r.q2.y) + (r.q2.z + r.q2.w), t3 = (r.q3.x + r.q3.y) + (r.q3.z + r.q3.w);
  float t = (t0 + t1) + (t2 + t3);
  t = t + dpp<0xB1>(t);  // quad_perm [1,0,3,2]
  t = t + dpp<0x4E>(t);  // quad_perm [2,3,0,1]
  return t;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// A staged list entry of the adjoint: the compositing fields of its record
// (one 48-byte slot, so the three loads share one address).
struct alignas(16) Staged {
  float4 a;  // mx, my, conic a, conic b
  float4 b;  // conic c, opacity, r, g
  float c;   // b
};

// The lane's four pixels (one per 8x8 quadrant of the tile) in the engine's
// backward walk (SplatRaster.cpp:346-433): dL/drgb, the background term, the
// transmittance and the suffix colour, kept folded (the value the next
// contributor's step reads, updated as soon as this one is done) so the
// previous contributor's alpha and colour need no registers.
struct AdjPixels {
  float dL0[4], dL1[4], dL2[4], tb[4], T[4], suf0[4], suf1[4], suf2[4];
  int lastp[4];
};

// Splat (a, b, cblue) at list index idx over quadrant k's pixels: the forward's
// decisions recomputed with its exact float operations, then the gradient
// terms added into acc. Every lane computes and non-contributing lanes add
// exact zeros; a quadrant with no contributing lane returns early. Returns
// whether any lane contributed; adds the pixels that shaped the splat (alpha
// below the clamp) to `shaped`.
// kLowCut: the falloff with its power <= -80 test (min_alpha below 1e-30).
//
// The falloff's exponent -0.5 (a dx^2 + c dy^2) - b dx dy is evaluated with
// the forward's operations: a dx dx and b dx depend on the pixel's column
// only and c dy dy on its row, and a lane's four pixels share two columns and
// two rows, so those products come in per splat (SplatCols, SplatRows).
struct SplatCols {
  float dx[2], axx[2], bx[2];  // per pixel column of the lane: dx, (a dx) dx, b dx
};
struct SplatRows {
  float dy[2], cyy[2];  // per pixel row of the lane: dy, (c dy) dy
};
template <int k, bool kLowCut>
__device__ __forceinline__ bool adjoint_quad(float (&acc)[kAcc], uint32_t& shaped, AdjPixels& s, const float4& a,
                                             const float4& b, float cblue, int idx, const SplatCols& X,
                                             const SplatRows& Y, float min_alpha) {
  const float dx = X.dx[k & 1], dy = Y.dy[k >> 1];
  const float power = -0.5f * (X.axx[k & 1] + Y.cyy[k >> 1]) - X.bx[k & 1] * dy;
  const float G = kLowCut ? falloff(power) : falloff_above_cut(power);
  const float raw = b.y * G;
  const float alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
  // the pixels that contribute: idx < last, !(power > 0), !(alpha < min_alpha)
  const uint64_t cmask = __builtin_amdgcn_sicmp(idx, s.lastp[k], kIcmpSlt) &
                         __builtin_amdgcn_fcmpf(power, 0.f, kFcmpUle) &
                         __builtin_amdgcn_fcmpf(alpha, min_alpha, kFcmpUge);
  if (cmask == 0ull) return false;
  {
  // the engine's expressions (SplatRaster.cpp:403-428) arranged for fused
  // multiply-adds: alpha_c is alpha where this pixel contributes and 0
  // elsewhere, so a dropped step changes nothing; the suffix colour
  // alpha c + (1 - alpha) suf is suf + alpha_c (c - suf), reusing dL/dalpha's
  // differences
#pragma clang fp contract(fast)
  // the contributing pixels whose alpha is below the clamp shape the splat
  const uint64_t smask = cmask & __builtin_amdgcn_fcmpf(raw, kMaxAlpha, kFcmpOlt);
  const float alpha_c = keep_lanes(cmask, alpha);
  const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
  const float Tn = s.T[k] * inv;
  const float weight = alpha_c * Tn;
  const float d0 = b.z - s.suf0[k], d1 = b.w - s.suf1[k], d2 = cblue - s.suf2[k];
  const float dL_dalpha = (d0 * s.dL0[k] + d1 * s.dL1[k] + d2 * s.dL2[k]) * Tn - s.tb[k] * inv;
  const float gd = keep_lanes(smask, G) * dL_dalpha;  // dL/dopacity term
  const float pw = b.y * gd;                            // dL/dpower
  const float px = pw * dx, py = pw * dy;
  const float u = px * a.z + py * a.w;  // -dL/dmean_x
  const float w = py * b.x + px * a.w;  // -dL/dmean_y
  acc[0] += weight * s.dL0[k];
  acc[1] += weight * s.dL1[k];
  acc[2] += weight * s.dL2[k];
  acc[3] += gd;
  acc[4] += u;
  acc[5] += w;
  acc[6] += __builtin_amdgcn_sqrtf(u * u + w * w);  // densify statistic: hardware sqrt
  acc[7] -= px * dx;  // dL/dconic xx and yy are halved after the wave sum
  acc[8] -= px * dy;
  acc[9] -= py * dy;
  shaped += (uint32_t)__popcll(smask);
  s.T[k] = select_lanes(cmask, Tn, s.T[k]);
  s.suf0[k] += alpha_c * d0;
  s.suf1[k] += alpha_c * d1;
  s.suf2[k] += alpha_c * d2;
  }
  return true;
}

// acc = 0 from the LDS zero block (volatile-free loads kept in place by an
// empty asm that clobbers memory, so they are neither hoisted nor merged).
__device__ __forceinline__ void load_zeros(float (&acc)[kAcc], const f32x4* s_zero, const f32x2& s_zero2) {
  asm volatile("" ::: "memory");
  const f32x4 z0 = s_zero[0], z1 = s_zero[1];
  const f32x2 z2 = s_zero2;
  acc[0] = z0.x, acc[1] = z0.y, acc[2] = z0.z, acc[3] = z0.w;
  acc[4] = z1.x, acc[5] = z1.y, acc[6] = z1.z, acc[7] = z1.w;
  acc[8] = z2.x, acc[9] = z2.y;
}

// The adjoint of one 16x16 tile by one wave: each lane owns four pixels, one
// per 8x8 quadrant (lane & 7, lane >> 3 within each), so a splat's terms over
// the whole tile are reduced once per wave and the four quadrants never meet
// at a barrier. The tile's list is walked back to front from the tile's last
// contributor in chunks of kBB entries: the lanes stage the chunk's records
// in LDS with a 4-bit mask of the quadrants whose cull box they meet below
// that quadrant's last contributor; then, per marked splat from the back,
// each marked quadrant's pixels are evaluated (a wave-uniform branch per
// quadrant), the lane adds its pixels' terms, the wave sums them through LDS,
// and lanes 0, 4, .., 40 issue the splat's result: up to 11 hardware float
// atomics (atomic mode), or 16 lanes one 64-byte partial store
// (deterministic mode, summation order fixed by the code alone).
//
// Five waves per SIMD (at most 96 VGPRs, one value spilled outside the splat
// loop): the per-pixel chains are long and dependent, and the fifth wave
// hides more of their latency than the registers a fourth-wave budget would
// allow (a pipelined LDS reduction, 118 VGPRs at four waves, was 4% slower).
// kDet selects the deterministic records at compile time, so the atomic
// form's register allocation carries none of their code.
template <int kBB, bool kQueue, bool kLowCut, bool kDet>
__global__ __launch_bounds__(64, 5) void k_backward(RasterParams p, const ViewDesc* __restrict__ views,
                                                 const uint32_t* __restrict__ ranges,
                                                 const uint32_t* __restrict__ pair_vals,
                                                 const SplatRec* __restrict__ rec, const float* __restrict__ dL_drgb,
                                                 const float* __restrict__ T_in, const int* __restrict__ last_in,
                                                 float* __restrict__ screen, const uint32_t* __restrict__ pair_base,
                                                 float* __restrict__ partial, uint8_t* __restrict__ visit,
                                                 uint32_t p_cap, const uint32_t* __restrict__ epoch,
                                                 const uint32_t* __restrict__ order, uint32_t* __restrict__ counter) {
  static_assert(kBB == 64 || kBB == 128, "chunks of one or two entries per lane");
  static_assert(kScreenStride >= kAcc + 1 && kPartialStride == 16, "record slots");
  constexpr int kHalves = kBB / 64;
  __shared__ Staged s_rec[kBB];
  __shared__ __attribute__((aligned(16))) float s_red[kAcc * kRedPitch];
  __shared__ f32x4 s_zero[2];  // the accumulators' initial zeros
  __shared__ f32x2 s_zero2;
  __shared__ int s_unit;
  const int lane = threadIdx.x;
  if (lane < 2) s_zero[lane] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (lane == 2) s_zero2 = f32x2{0.f, 0.f};
  wave_sync_lds();
  // after the reduction lane 4i holds term i (i < 10); lane 40 takes the
  // shaped-pixel count (slot 10), which is known in a scalar register
  const int my_slot = lane >> 2;
  const bool slot_lane = (lane & 3) == 0;
  const bool writer = slot_lane && my_slot <= kAcc;
  // terms 7 and 9 (dL/dconic xx, yy) are summed without their factor 1/2,
  // applied to the total (exact: a power of two)
  const float slot_scale = (my_slot == 7 || my_slot == 9) ? 0.5f : 1.f;
  // deterministic records: slots 0-10 the terms, 11 this pass's tag, 12-15 zero
  const float rec_pad = my_slot == 11 ? __uint_as_float(kDet ? *epoch : 0u) : 0.f;
  // and beside each record its visit byte, the tag's low byte
  const uint8_t visit_tag = (uint8_t)(kDet ? *epoch : 0u);
  for (int iter = 0;; ++iter) {
    int tile;
    if (kQueue) {
      if (lane == 0) s_unit = (int)atomicAdd(counter, 1u);
      wave_sync_lds();
      const int idx = __builtin_amdgcn_readfirstlane(s_unit);
      wave_sync_lds();
      if (idx >= p.total_tiles) break;
      tile = (int)order[idx];
    } else {
      if (iter > 0) break;
      tile = xcd_remap(blockIdx.x, p.total_tiles);
    }
    const ViewDesc& V = views[find_view(views, p.n_views, tile)];
    const int local = tile - V.tile_base;
    const int tx = local % V.cols, ty = local / V.cols;
    const TileQuads tq = tile_quads(V, tx, ty);
    const uint32_t begin = ranges[2 * tile];  // the walk starts at the last contributor, before the list's end
    // pixel centres: quadrant 0's, +8 (exact) for the others
    const float fx0 = (float)(tx * kTile + (lane & 7)) + 0.5f, fy0 = (float)(ty * kTile + (lane >> 3)) + 0.5f;
    AdjPixels s;
    int stops[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int px = tx * kTile + 8 * (k & 1) + (lane & 7), py = ty * kTile + 8 * (k >> 1) + (lane >> 3);
      s.dL0[k] = s.dL1[k] = s.dL2[k] = 0.f;
      float T_final = 1.f;
      s.lastp[k] = 0;
      if (px < V.width && py < V.height) {
        const long long pixel = V.pixel_base + (long long)py * V.width + px;
        s.dL0[k] = dL_drgb[3 * pixel];
        s.dL1[k] = dL_drgb[3 * pixel + 1];
        s.dL2[k] = dL_drgb[3 * pixel + 2];
        T_final = T_in[pixel];
        s.lastp[k] = last_in[pixel];
      }
      s.tb[k] = T_final * (p.bg[0] * s.dL0[k] + p.bg[1] * s.dL1[k] + p.bg[2] * s.dL2[k]);
      s.T[k] = T_final;
      s.suf0[k] = s.suf1[k] = s.suf2[k] = 0.f;
      int m = s.lastp[k];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
      stops[k] = __builtin_amdgcn_readfirstlane(m);
    }
    // pairs past every pixel's last contributor are not visited: their
    // partials keep an older tag and read as zeros
    const uint32_t stop = begin + (uint32_t)max(max(stops[0], stops[1]), max(stops[2], stops[3]));
    // the accumulators start as zeros loaded from LDS (a broadcast read: no
    // vector ALU instruction, and a value the compiler cannot fold into the
    // first quadrant's adds, which would cost a copy of all ten on every path
    // that skips that quadrant)
    float acc[kAcc];
    load_zeros(acc, s_zero, s_zero2);
    for (uint32_t hi = stop; hi > begin;) {
      const uint32_t lo = hi - min((uint32_t)kBB, hi - begin);
      const int cnt = (int)(hi - lo);
      // stage the chunk: records to LDS; each entry's quadrant mask goes into
      // four wave ballots (scalar registers) and its destination stays in the
      // staging lane's register, so the compositing loop reads neither from LDS
      uint64_t qbits[kHalves][4];
      uint32_t dst_reg[kHalves];
#pragma unroll
      for (int h = 0; h < kHalves; ++h) {
        const int j = lane + 64 * h;
        uint32_t m = 0u, dst = 0u;
        if (j < cnt) {
          const uint32_t e = lo + j;
          const uint32_t slot = pair_vals[e];
          const SplatRec* r = rec + slot;
          const float4* r4 = reinterpret_cast<const float4*>(r);
          const float4 a = r4[0], b = r4[1], c = r4[2];
          s_rec[j].a = a;
          s_rec[j].b = b;
          s_rec[j].c = c.x;
          const int idx = (int)(e - begin);
          m = quad_hits(a.x, a.y, c.y, c.z, tq);
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (idx >= stops[k] ||
                ((m >> k) & 1u && !ellipse_meets_rect(a.z, a.w, b.x, b.y, p.min_alpha, tq.x0[k & 1] - a.x,
                                                      tq.x1[k & 1] - a.x, tq.y0[k >> 1] - a.y, tq.y1[k >> 1] - a.y)))
              m &= ~(1u << k);
          dst = kDet ? pair_base[slot] + (uint32_t)((ty - r->y0) * (r->x1 - r->x0 + 1) + (tx - r->x0)) : slot;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) qbits[h][k] = __ballot((m >> k) & 1u);
        dst_reg[h] = dst;
      }
      wave_sync_lds();
#pragma unroll
      for (int h = kHalves - 1; h >= 0; --h) {
        uint64_t bits = qbits[h][0] | qbits[h][1] | qbits[h][2] | qbits[h][3];
        while (bits) {
          const int bit = 63 - __builtin_clzll(bits);
          bits &= ~(1ull << bit);
          const int j = 64 * h + bit;
          const int idx = (int)(lo + j - begin);
          const uint32_t m = (uint32_t)((qbits[h][0] >> bit) & 1u) | (uint32_t)((qbits[h][1] >> bit) & 1u) << 1 |
                             (uint32_t)((qbits[h][2] >> bit) & 1u) << 2 | (uint32_t)((qbits[h][3] >> bit) & 1u) << 3;
          const float4 a = s_rec[j].a;
          const float4 b = s_rec[j].b;
          const float cblue = s_rec[j].c;
          SplatCols X;
          SplatRows Y;
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            X.dx[c] = (c ? fx0 + 8.f : fx0) - a.x;
            X.axx[c] = a.z * X.dx[c] * X.dx[c];
            X.bx[c] = a.w * X.dx[c];
            Y.dy[c] = (c ? fy0 + 8.f : fy0) - a.y;
            Y.cyy[c] = b.x * Y.dy[c] * Y.dy[c];
          }
          uint32_t shaped = 0u;  // pixels that shaped the splat, from ballots: a scalar
          bool any = false;
          if (m & 1u) any = adjoint_quad<0, kLowCut>(acc, shaped, s, a, b, cblue, idx, X, Y, p.min_alpha) | any;
          if (m & 2u) any = adjoint_quad<1, kLowCut>(acc, shaped, s, a, b, cblue, idx, X, Y, p.min_alpha) | any;
          if (m & 4u) any = adjoint_quad<2, kLowCut>(acc, shaped, s, a, b, cblue, idx, X, Y, p.min_alpha) | any;
          if (m & 8u) any = adjoint_quad<3, kLowCut>(acc, shaped, s, a, b, cblue, idx, X, Y, p.min_alpha) | any;
          float z = 0.f;
          if (any) {
#pragma unroll
            for (int i = 0; i < kAcc; ++i) s_red[i * kRedPitch + lane] = acc[i];
            wave_sync_lds();
            z = quarter_sum(lds_quarter_load(s_red, lane)) * slot_scale;
            z = my_slot == kAcc ? (float)shaped : z;  // an exact integer
            wave_sync_lds();  // the next splat's writes follow these loads
            load_zeros(acc, s_zero, s_zero2);  // a splat that added nothing left them at zero
          }
          // lanes 0, 4, .., 40 hold the splat's totals: atomics (nonzero terms
          // only), or the partial's 64-byte store from lanes 0, 4, .., 60
          // (slots 11-15 the tag and zero padding, so the store fills its
          // whole memory sector) and the visit byte from lane 0
          const uint32_t dsti = (uint32_t)__builtin_amdgcn_readlane((int)dst_reg[h], bit);
          if constexpr (kDet) {
            if (slot_lane && dsti < p_cap) {
              partial[(long long)dsti * kPartialStride + my_slot] = my_slot <= kAcc ? z : rec_pad;
              if (lane == 0) visit[dsti] = visit_tag;
            }
          } else if (any && writer && z != 0.f) {
            unsafeAtomicAdd(screen + (long long)dsti * kScreenStride + my_slot, z);
          }
        }
      }
      wave_sync_lds();  // the next chunk overwrites the staging
      hi = lo;
    }
  }
}

// Tiles by list length, longest first (key: 0xFFFF - min(length, 0xFFFF)).
__global__ __launch_bounds__(256) void k_tile_order_keys(const uint32_t* __restrict__ ranges, int total_tiles,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total_tiles) return;
  const uint32_t len = ranges[2 * t + 1] - ranges[2 * t];
  keys[t] = 0xFFFFu - (len < 0xFFFFu ? len : 0xFFFFu);
  vals[t] = (uint32_t)t;
}

}  // namespace

namespace {

template <int kShape>
int forward_units(int tiles) {
  return kShape == 0 ? tiles : (kShape == 1 ? 4 * tiles : 2 * tiles);
}

// Resident workgroups per CU of a kernel (for persistent grids).
template <typename K>
int resident_blocks(K kernel, int threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
  return n;
}

template <int kShape, bool kQueue, bool kLowCut>
hipError_t forward_variant(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                           const uint32_t* pair_vals, const SplatRec* rec, float* rgb, float* T, int* last,
                           const uint32_t* order, uint32_t* counter, int cus, hipStream_t stream) {
  constexpr int kThreads = kShape == 1 ? 64 : 256;
  const int units = forward_units<kShape>(p.total_tiles);
  static const int per_cu = resident_blocks(k_forward<kShape, kQueue, kLowCut>, kThreads);
  const int blocks = kQueue ? std::min(units, per_cu * cus) : units;
  hipLaunchKernelGGL((k_forward<kShape, kQueue, kLowCut>), dim3(blocks), dim3(kThreads), 0, stream, p, views, ranges,
                     pair_vals, rec, rgb, T, last, order, counter);
  return hipGetLastError();
}

template <int kBB, bool kQueue, bool kLowCut>
hipError_t backward_variant(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                            const uint32_t* pair_vals, const SplatRec* rec, const float* dL_drgb, const float* T,
                            const int* last, float* screen, const uint32_t* pair_base, float* partial,
                            uint8_t* visit, uint32_t p_cap, bool deterministic, const uint32_t* epoch,
                            const uint32_t* order,
                            uint32_t* counter, int cus, hipStream_t stream) {
  // the two forms have the same launch bounds and LDS, so the same residency
  static const int per_cu = resident_blocks(k_backward<kBB, kQueue, kLowCut, false>, 64);
  const int blocks = kQueue ? std::min(p.total_tiles, per_cu * cus) : p.total_tiles;
  if (deterministic)
    hipLaunchKernelGGL((k_backward<kBB, kQueue, kLowCut, true>), dim3(blocks), dim3(64), 0, stream, p, views, ranges,
                       pair_vals, rec, dL_drgb, T, last, screen, pair_base, partial, visit, p_cap, epoch, order,
                       counter);
  else
    hipLaunchKernelGGL((k_backward<kBB, kQueue, kLowCut, false>), dim3(blocks), dim3(64), 0, stream, p, views, ranges,
                       pair_vals, rec, dL_drgb, T, last, screen, pair_base, partial, visit, p_cap, epoch, order,
                       counter);
  return hipGetLastError();
}

}  // namespace

const char* raster_shape_name(int shape) {
  return shape == 0 ? "16x16" : (shape == 1 ? "8x8" : "32x8");
}

hipError_t launch_tile_order_keys(const uint32_t* ranges, int total_tiles, uint32_t* keys, uint32_t* vals,
                                  hipStream_t stream) {
  if (total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_order_keys, dim3((unsigned)((total_tiles + 255) / 256)), dim3(256), 0, stream, ranges,
                     total_tiles, keys, vals);
  return hipGetLastError();
}

hipError_t launch_forward(const RasterParams& p, const RasterConfig& cfg, const ViewDesc* views,
                          const uint32_t* ranges, const uint32_t* pair_vals, const SplatRec* rec, float* rgb, float* T,
                          int* last, const uint32_t* order, uint32_t* counter, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  if (cfg.fwd_queue) SPLAT_HIP_CHECK(launch_fill_u32(counter, 0u, 1, stream));
  const int pb = cfg.cus;
  // the falloff's power <= -80 test is needed only for min_alpha below the
  // alpha of e^-80 (falloff_above_cut)
  const bool low_cut = !(p.min_alpha >= 1e-30f);
#define SPLAT_FWD(SH, Q, LC) \
  forward_variant<SH, Q, LC>(p, views, ranges, pair_vals, rec, rgb, T, last, order, counter, pb, stream)
  switch ((low_cut ? 6 : 0) + cfg.fwd_shape * 2 + (cfg.fwd_queue ? 1 : 0)) {
    case 0: return SPLAT_FWD(0, false, false);
    case 1: return SPLAT_FWD(0, true, false);
    case 2: return SPLAT_FWD(1, false, false);
    case 3: return SPLAT_FWD(1, true, false);
    case 4: return SPLAT_FWD(2, false, false);
    case 5: return SPLAT_FWD(2, true, false);
    case 6: return SPLAT_FWD(0, false, true);
    case 7: return SPLAT_FWD(0, true, true);
    case 8: return SPLAT_FWD(1, false, true);
    case 9: return SPLAT_FWD(1, true, true);
    case 10: return SPLAT_FWD(2, false, true);
    case 11: return SPLAT_FWD(2, true, true);
    default: return hipErrorInvalidValue;
  }
#undef SPLAT_FWD
}

hipError_t launch_backward(const RasterParams& p, const RasterConfig& cfg, const ViewDesc* views,
                           const uint32_t* ranges, const uint32_t* pair_vals, const SplatRec* rec,
                           const float* dL_drgb, const float* T, const int* last, float* screen,
                           const uint32_t* pair_base, float* partial, uint8_t* visit, uint32_t p_cap,
                           bool deterministic, const uint32_t* epoch, const uint32_t* order, uint32_t* counter,
                           hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  if (deterministic && (partial == nullptr || visit == nullptr)) return hipErrorInvalidValue;
  if (cfg.bwd_queue) SPLAT_HIP_CHECK(launch_fill_u32(counter, 0u, 1, stream));
  const int pb = cfg.cus;
  // the falloff's power <= -80 test is needed only for min_alpha below the
  // alpha of e^-80 (see adjoint_quad)
  const bool low_cut = !(p.min_alpha >= 1e-30f);
#define SPLAT_BWD(BB, Q, LC)                                                                                     \
  backward_variant<BB, Q, LC>(p, views, ranges, pair_vals, rec, dL_drgb, T, last, screen, pair_base, partial, \
                              visit, p_cap, deterministic, epoch, order, counter, pb, stream)
  switch ((low_cut ? 4 : 0) + (cfg.bwd_batch == 64 ? 0 : 2) + (cfg.bwd_queue ? 1 : 0)) {
    case 0: return SPLAT_BWD(64, false, false);
    case 1: return SPLAT_BWD(64, true, false);
    case 2: return SPLAT_BWD(128, false, false);
    case 3: return SPLAT_BWD(128, true, false);
    case 4: return SPLAT_BWD(64, false, true);
    case 5: return SPLAT_BWD(64, true, true);
    case 6: return SPLAT_BWD(128, false, true);
    default: return SPLAT_BWD(128, true, true);
  }
#undef SPLAT_BWD
}

}  // namespace splat
