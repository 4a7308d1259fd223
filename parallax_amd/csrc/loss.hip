// L1 + D-SSIM photometric loss and its gradient.
//
// The engine's photometricLoss / meanSsim / blur (engine/native/cpp/src/splat/
// SplatRaster.cpp:580-738): L = (1-w) mean|x-y| + w (1 - mean SSIM) with an
// 11-tap sigma 1.5 Gaussian window applied as a zero-padded separable blur
// (rows first, then columns), C1 = 0.01^2, C2 = 0.03^2, means over 3 * pixels.
// The gradient of the SSIM term is itself a blur of five per-pixel partial
// planes (the blur is self-adjoint on a zero-padded domain):
//   dSSIM/dx = B(t_mu) + 2x B(t_var) - 2 B(t_var mu_x) + y B(t_cov) - B(t_cov mu_y)
// and, B being linear, the three planes that carry no per-pixel factor are
// summed before the blur: A = t_mu - 2 t_var mu_x - t_cov mu_y, so three
// planes per channel are written, read and blurred instead of five
// (dSSIM/dx = B(A) + 2x B(t_var) + y B(t_cov)).
//
// Both passes walk down the image. A work unit is a 128-column x 86-row block
// of one view; its waves walk the block's rows top to bottom, from 5 rows
// above it to 5 rows below (zero outside the image):
//   - the row blur of a row reads each lane's 11-tap window from that row,
//     staged in the wave's own LDS slice (no workgroup barrier);
//   - the column blur keeps, per lane, the last 12 row-blurred values of each
//     plane in registers (a 12-slot ring, the row loop unrolled by 12 so
//     every slot index is static) and blurs the last 11 as each new row
//     arrives, so no row is blurred twice within a unit.
// stats: three waves per half unit (64 columns), one per channel, one lane
//        per column: the five statistics planes (x, y, xx, yy, xy) per pixel,
//        then the SSIM value, the three partial planes (to HBM) and the L1
//        terms; per-half-unit L1 and SSIM sums in a fixed order.
// grad:  three waves per half unit, one per channel, one lane per column:
//        the three partial planes of the channel are blurred and combined
//        into dL/drgb; the three channels of 12 output rows are assembled in
//        LDS and written as whole interleaved rows.
// The blur taps run in tap order -5..5 and the per-pixel SSIM expressions
// are the engine's, with its divisions by c d, c^2 d and c d^2 taken as
// products of two hardware reciprocals (1 ulp each); the loss and dL/drgb
// are held to the float64 evaluation of the engine's loss.
// Compulsory HBM traffic per pixel: 24 B read and 36 B written by the first
// pass, 60 B read and 12 B written by the second, plus 10 halo rows per 86.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr int kR = 5;
constexpr int kTaps = 2 * kR + 1;         // 11
// Ring slots and rows per unrolled group: 12, a multiple of both prefetch
// depths below, so the queue's registers line up at the group loop's back
// edge and no load in flight has to be waited for there.
constexpr int kRing = 12;
constexpr int kUnitW = 128;               // columns per unit
constexpr int kUnitH = 86;                // output rows per unit: 96 input rows, eight groups of 12
constexpr int kStatWaves = 3;             // stats: one wave per channel over half a unit (64 columns)
constexpr int kGradWaves = 3;             // grad: one wave per channel over half a unit (64 columns)
constexpr int kGradPitch = 128;           // 74 staged columns per grad wave (two per lane written)
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

// Work units of a view: 128-column x 78-row blocks. The statistics pass runs
// one workgroup per half unit (64 columns); its partial sums and the views'
// lt_base count half units.
__host__ __device__ __forceinline__ int loss_units(int width, int height) {
  return ((width + kUnitW - 1) / kUnitW) * ((height + kUnitH - 1) / kUnitH);
}
__host__ __device__ __forceinline__ int loss_halves(int width, int height) { return 2 * loss_units(width, height); }

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

struct Unit {
  int block;     // index into the block sums (statistics pass: half units)
  int sx, by;    // first column, first output row
  int rows;      // output rows (<= kUnitH)
  const ViewDesc* V;
};

// The block's half unit (64 columns). With `reverse` (the gradient pass)
// the blocks take the half units in the reverse of the statistics pass's
// order: the partial planes written last are then read first, while the
// memory-side cache may still hold them.
__device__ __forceinline__ Unit unit_of(const LossParams& p, const ViewDesc* views, bool reverse) {
  Unit u;
  const int n = p.total_tiles;
  u.block = xcd_remap(blockIdx.x, n);
  if (reverse) u.block = n - 1 - u.block;
  int v = 0;
  for (int k = 1; k < p.n_views; ++k)
    if (views[k].lt_base <= u.block) v = k;
  const ViewDesc& V = views[v];
  const int local = u.block - V.lt_base;
  const int unit = local >> 1;
  const int strips = (V.width + kUnitW - 1) / kUnitW;
  u.sx = (unit % strips) * kUnitW + 64 * (local & 1);
  u.by = (unit / strips) * kUnitH;
  u.rows = min(kUnitH, V.height - u.by);
  u.V = &V;
  return u;
}

// Value of channel c of an interleaved RGB image at (row, col), zero outside.
// The load is unconditional, from the nearest pixel inside the image, so a
// row's loads can be issued rows ahead with no branch around them.
__device__ __forceinline__ float pixel_or_zero(const float* img, int width, int height, int row, int col, int c) {
  const bool in = row >= 0 && row < height && col >= 0 && col < width;
  const int rr = min(max(row, 0), height - 1), cc = min(max(col, 0), width - 1);
  const float v = img[3 * ((long long)rr * width + cc) + c];
  return in ? v : 0.f;
}

// Rows a wave's loads run ahead of its compositing of them: enough rows in
// flight to cover an HBM round trip.
constexpr int kStatAhead = 4;
constexpr int kGradAhead = 4;

// A raw buffer over `bytes` bytes at `base`: stores at offsets past the end
// are dropped by the hardware, so a lane is switched off by giving it the
// offset kOff, with no branch around the store.
constexpr uint32_t kOff = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void store_or_drop(__amdgpu_buffer_rsrc_t r, uint32_t offset, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)offset, 0, 0);
}
// A load of the float at byte `lane_offset` + `row_offset`: the hardware's
// range check covers the per-lane offset only (a lane offset of kOff reads
// 0), so the wave-uniform row offset must stay inside the buffer.
__device__ __forceinline__ float load_or_zero(__amdgpu_buffer_rsrc_t r, uint32_t lane_offset, uint32_t row_offset) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_offset, (int)row_offset, 0));
}

// ---------------------------------------------------------------- stats --
// The loop body of a row is free of branches (loads from clamped addresses,
// selects, stores the hardware drops), so the compiler can keep the rows'
// loads in flight across it.
template <bool kSsim>
__global__ __launch_bounds__(kStatWaves * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_loss_stats(LossParams p, const ViewDesc* __restrict__ views,
                                                                const float* __restrict__ rgb,
                                                                const float* __restrict__ targets,
                                                                float* __restrict__ dL, float* __restrict__ terms,
                                                                long long total_px, float* __restrict__ block_sums) {
  // two staged rows per wave (74 columns; every lane writes two slots of each)
  __shared__ float s_x[kStatWaves][2 * 128];
  __shared__ float s_y[kStatWaves][2 * 128];
  __shared__ float s_red[2][kStatWaves];
  const Unit U = unit_of(p, views, false);
  const ViewDesc& V = *U.V;
  // the wave index read from lane 0 is known to be uniform: buffer
  // descriptors and plane bases built from it stay in scalar registers
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = wave;                            // channel
  const int x0 = U.sx;                           // first output column
  const int col = x0 + lane;                     // this lane's column
  // the view's fields in registers (the buffer stores below are opaque, so
  // fields read through V would be re-read after each of them)
  const int W = V.width, H = V.height, by = U.by, rows = U.rows;
  const bool col_ok = col < W;
  const long long pixels = (long long)W * H;
  const float inverse = 1.f / (float)(pixels * 3);
  const float* img = rgb + 3 * V.pixel_base;
  const float* tgt = targets + V.target_base;
  float* sx_ = s_x[wave];
  float* sy_ = s_y[wave];
  float l1 = 0.f, ssim = 0.f;

  if (!kSsim) {
    // L1 only: dL/drgb = sign(x - y) / (3 W H) per pixel, written here
    const __amdgpu_buffer_rsrc_t rdl = buffer_of(dL + 3 * V.pixel_base, (uint32_t)(pixels * 12));
    for (int o = 0; o < rows; ++o) {
      const int row = by + o;
      const float x = pixel_or_zero(img, W, H, row, col, c);
      const float y = pixel_or_zero(tgt, W, H, row, col, c);
      const float difference = x - y;
      l1 += col_ok ? fabsf(difference) : 0.f;
      const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
      store_or_drop(rdl, col_ok ? (uint32_t)((3 * ((long long)row * W + col) + c) * 4) : kOff,
                    sign * inverse);
    }
  } else {
    // staged columns x0 - 5 + lane and x0 + 59 + lane of an input row, loaded
    // kStatAhead rows before the row is staged (q[0] is the next row to stage)
    struct Row {
      float x0, y0, x1, y1;
    };
    Row q[kStatAhead];
    // lanes read columns x0 - 5 + lane and, the first ten of them, x0 + 59 +
    // lane: the 74 columns the windows use (kOff outside the image, or
    // unused: no memory access); rows outside it read row 0 and are zeroed
    // by a uniform select
    const __amdgpu_buffer_rsrc_t rimg = buffer_of(img, (uint32_t)(pixels * 12));
    const __amdgpu_buffer_rsrc_t rtgt = buffer_of(tgt, (uint32_t)(pixels * 12));
    const int cola = x0 - kR + lane, colb = cola + 64;
    const uint32_t la = cola >= 0 && cola < W ? (uint32_t)((3 * cola + c) * 4) : kOff;
    const uint32_t lb = lane < 2 * kR && colb < W ? (uint32_t)((3 * colb + c) * 4) : kOff;
    auto fetch = [&](int r) {
      const int row = by - kR + r;
      const bool row_ok = row >= 0 && row < H;
      const uint32_t ro = row_ok ? (uint32_t)(row * W * 12) : 0u;
      const uint32_t a = row_ok ? la : kOff, b = row_ok ? lb : kOff;  // rows outside read 0
      Row f;
      f.x0 = load_or_zero(rimg, a, ro);
      f.y0 = load_or_zero(rtgt, a, ro);
      f.x1 = load_or_zero(rimg, b, ro);
      f.y1 = load_or_zero(rtgt, b, ro);
      return f;
    };
    const uint32_t plane_bytes = (uint32_t)(pixels * 4);
    const __amdgpu_buffer_rsrc_t rt0 = buffer_of(terms + (long long)(3 * c) * total_px + V.pixel_base, plane_bytes);
    const __amdgpu_buffer_rsrc_t rt1 = buffer_of(terms + (long long)(3 * c + 1) * total_px + V.pixel_base, plane_bytes);
    const __amdgpu_buffer_rsrc_t rt2 = buffer_of(terms + (long long)(3 * c + 2) * total_px + V.pixel_base, plane_bytes);
    float ring[5][kRing];  // row-blurred x, y, xx, yy, xy of the last 12 rows

    // one input row: stage it, take the lane's window, add the lane's own
    // pixel to L1 (input rows 5 .. rows + 4 are the unit's own rows), and
    // row-blur the five statistics into ring slot `slot`
    auto stage = [&](int r, const Row& f) {
      float* bx = sx_ + (r & 1) * 128;
      float* by_ = sy_ + (r & 1) * 128;
      bx[lane] = f.x0;
      by_[lane] = f.y0;
      bx[64 + lane] = f.x1;
      by_[64 + lane] = f.y1;
    };
    auto row_in = [&](int r, int slot) {
      // one scheduling region per row: the prefetch queue, not the
      // scheduler, decides how far loads run ahead
      __builtin_amdgcn_sched_barrier(0);
      // row r was staged (in buffer r & 1) a step ago: its writes are long
      // done, and the stage of row r + 1 below goes to the other buffer,
      // whose row r - 1 was read a step ago
      wave_sync_lds();
      float xw[kTaps], yw[kTaps];
#pragma unroll
      for (int i = 0; i < kTaps; ++i) {
        xw[i] = sx_[(r & 1) * 128 + lane + i];
        yw[i] = sy_[(r & 1) * 128 + lane + i];
      }
      stage(r + 1, q[0]);
#pragma unroll
      for (int k = 0; k + 1 < kStatAhead; ++k) q[k] = q[k + 1];
      q[kStatAhead - 1] = fetch(r + 1 + kStatAhead);  // past the band: clamped loads, never staged
      const bool own = col_ok && r >= kR && r < rows + kR;
      l1 += own ? fabsf(xw[kR] - yw[kR]) : 0.f;
      // k x and k y once per tap: the x, y sums are their sums, and the
      // xx, yy, xy sums are one fused multiply-add each per tap
      float h[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < kTaps; ++i) {
        const float kx = p.kernel[i] * xw[i], ky = p.kernel[i] * yw[i];
        h[0] += kx;
        h[1] += ky;
        h[2] = __builtin_fmaf(kx, xw[i], h[2]);
        h[3] = __builtin_fmaf(ky, yw[i], h[3]);
        h[4] = __builtin_fmaf(kx, yw[i], h[4]);
      }
#pragma unroll
      for (int pl = 0; pl < 5; ++pl) ring[pl][slot] = h[pl];
    };

    stage(0, fetch(0));
#pragma unroll
    for (int k = 0; k < kStatAhead; ++k) q[k] = fetch(k + 1);
    // groups of 12 input rows, row r into ring slot r % 12; input rows 10 ..
    // rows + 9 complete output rows 0 .. rows - 1, rows past the unit (the
    // last band of a view) are computed from zeros and masked
    const int groups = (rows + 2 * kR + kRing - 1) / kRing;
    for (int g = 0; g < groups; ++g) {
#pragma unroll
      for (int j = 0; j < kRing; ++j) {
        const int r = g * kRing + j;
        row_in(r, j);
        const int o = r - 2 * kR;  // the output row this input row completes
        if (o < 0) continue;       // the first group's first ten rows only fill the ring
        const bool out_ok = col_ok && o < rows;
        // the column blur of the ring, oldest row first
        float st[5];
#pragma unroll
        for (int pl = 0; pl < 5; ++pl) {
          float s = 0.f;
#pragma unroll
          for (int i = 0; i < kTaps; ++i) s = __builtin_fmaf(p.kernel[i], ring[pl][(j + kRing - 2 * kR + i) % kRing], s);
          st[pl] = s;
        }
        const float mu_x = st[0], mu_y = st[1], exx = st[2], eyy = st[3], exy = st[4];
        const float var_x = exx - mu_x * mu_x;
        const float var_y = eyy - mu_y * mu_y;
        const float cov = exy - mu_x * mu_y;
        const float a = 2.f * mu_x * mu_y + kC1;
        const float b = 2.f * cov + kC2;
        const float cc = mu_x * mu_x + mu_y * mu_y + kC1;
        const float d = var_x + var_y + kC2;
        // cc >= C1 and d ~ C2 > 0: two hardware reciprocals (1 ulp) stand in
        // for the engine's five divisions
        const float rc = __builtin_amdgcn_rcpf(cc), rd = __builtin_amdgcn_rcpf(d);
        const float rcd = rc * rd;
        const float value = (a * b) * rcd;  // the pixel's SSIM
        ssim += out_ok ? value : 0.f;
        const float term_mu = inverse * (2.f * mu_y * b * rcd - 2.f * mu_x * value * rc);
        const float term_var = inverse * (-value * rd);
        const float term_cov = inverse * (2.f * a * rcd);
        const uint32_t off = out_ok ? (uint32_t)(((by + o) * W + col) * 4) : kOff;
        store_or_drop(rt0, off, (term_mu - 2.f * term_var * mu_x) - term_cov * mu_y);
        store_or_drop(rt1, off, term_var);
        store_or_drop(rt2, off, term_cov);
      }
      // the three waves read the same interleaved RGB lines (each takes one
      // channel): kept within a group of rows of each other, the lines one
      // wave brings into the cache are still there for the others. A bare
      // barrier: nothing is exchanged through memory here.
      __builtin_amdgcn_s_barrier();
    }
  }
  // fixed-order sums: wave butterflies, then the waves in index order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    l1 += __shfl_xor(l1, off, 64);
    ssim += __shfl_xor(ssim, off, 64);
  }
  if (lane == 0) {
    s_red[0][wave] = l1;
    s_red[1][wave] = ssim;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
#pragma unroll
    for (int k = 0; k < kStatWaves; ++k) {
      a += s_red[0][k];
      b += s_red[1][k];
    }
    block_sums[2 * U.block] = a;
    block_sums[2 * U.block + 1] = b;
  }
}

// ----------------------------------------------------------------- grad --
// Three waves per half unit, one per channel, one lane per column: the
// three partial planes of the channel are row-blurred from a 74-column LDS
// slice, column-blurred from a 12-row register ring, and combined into
// dL/drgb; the three channels of 12 output rows are assembled in LDS and
// written as whole interleaved rows.
__global__ __launch_bounds__(kGradWaves * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_loss_grad(LossParams p, const ViewDesc* __restrict__ views,
                                                               const float* __restrict__ rgb,
                                                               const float* __restrict__ targets,
                                                               const float* __restrict__ terms, long long total_px,
                                                               float* __restrict__ dL) {
  __shared__ __align__(16) float s_t[kGradWaves][3][kGradPitch];  // 74 staged columns; every lane writes two
  __shared__ __align__(16) float s_dl[kRing][64 * 3];              // one group of output rows, interleaved
  const Unit U = unit_of(p, views, true);
  const ViewDesc& V = *U.V;
  const int lane = threadIdx.x & 63;
  const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave = channel, wave-uniform
  const int col = U.sx + lane;                                     // this lane's column
  // the view's fields in registers (the buffer stores below are opaque)
  const int W = V.width, H = V.height, by = U.by, sx = U.sx, rows = U.rows;
  const long long pixels = (long long)W * H;
  const float inverse = 1.f / (float)(pixels * 3);
  const float w = p.ssim_weight;
  const float* img = rgb + 3 * V.pixel_base;
  const float* tgt = targets + V.target_base;
  const float* plane0 = terms + (long long)(3 * c) * total_px + V.pixel_base;
  const int valid_floats = 3 * min(64, W - sx);
  const __amdgpu_buffer_rsrc_t rdl = buffer_of(dL + 3 * V.pixel_base, (uint32_t)(pixels * 12));
  float (*st)[kGradPitch] = s_t[c];

  // for input row r: the three planes at staged columns sx - 5 + lane and
  // sx + 59 + lane, and x and y of this lane's pixel in output row r - 10;
  // loaded kGradAhead rows before they are used (clamped loads, no branches)
  struct Row {
    float v[3][2];
    float ox, oy;
  };
  Row q[kGradAhead];
  const uint32_t plane_bytes = (uint32_t)(pixels * 4);
  const __amdgpu_buffer_rsrc_t rp0 = buffer_of(plane0, plane_bytes);
  const __amdgpu_buffer_rsrc_t rp1 = buffer_of(plane0 + total_px, plane_bytes);
  const __amdgpu_buffer_rsrc_t rp2 = buffer_of(plane0 + 2 * total_px, plane_bytes);
  const __amdgpu_buffer_rsrc_t rimg = buffer_of(img, (uint32_t)(pixels * 12));
  const __amdgpu_buffer_rsrc_t rtgt = buffer_of(tgt, (uint32_t)(pixels * 12));
  // plane columns sx - 5 + lane and, the first ten lanes, sx + 59 + lane: the
  // 74 the windows use (kOff outside the image or past them: no memory access)
  uint32_t lp[2];
  {
    const int cola = sx - kR + lane, colb = cola + 64;
    lp[0] = cola >= 0 && cola < W ? (uint32_t)(cola * 4) : kOff;
    lp[1] = lane < 2 * kR && colb < W ? (uint32_t)(colb * 4) : kOff;
  }
  const uint32_t lo = col < W ? (uint32_t)((3 * col + c) * 4) : kOff;  // this lane's pixel, channel c
  auto fetch = [&](int r) {
    Row f;
    const int row = by - kR + r;
    const bool row_ok = row >= 0 && row < H;
    const uint32_t ro = row_ok ? (uint32_t)(row * W * 4) : 0u;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t a = row_ok ? lp[k] : kOff;  // rows outside read 0
      f.v[0][k] = load_or_zero(rp0, a, ro);
      f.v[1][k] = load_or_zero(rp1, a, ro);
      f.v[2][k] = load_or_zero(rp2, a, ro);
    }
    const int orow = by + r - 2 * kR;
    const bool orow_ok = orow >= 0 && orow < H;
    const uint32_t oro = orow_ok ? (uint32_t)(orow * W * 12) : 0u;
    const uint32_t a = orow_ok ? lo : kOff;
    f.ox = load_or_zero(rimg, a, oro);
    f.oy = load_or_zero(rtgt, a, oro);
    return f;
  };
  float ring[3][kRing];
  float ox = 0.f, oy = 0.f;  // x and y of the lane's pixel in the current output row

  // one input row: stage the planes, row-blur the lane's column into ring
  // slot `slot`, and take the output row's x and y
  auto row_in = [&](int r, int slot) {
    __builtin_amdgcn_sched_barrier(0);  // one scheduling region per row
    wave_sync_lds();
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      st[pl][lane] = q[0].v[pl][0];
      st[pl][64 + lane] = q[0].v[pl][1];
    }
    ox = q[0].ox;
    oy = q[0].oy;
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k + 1 < kGradAhead; ++k) q[k] = q[k + 1];
    q[kGradAhead - 1] = fetch(r + kGradAhead);  // past the band: clamped loads, never used
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      float h = 0.f;
#pragma unroll
      for (int i = 0; i < kTaps; ++i) h = __builtin_fmaf(p.kernel[i], st[pl][lane + i], h);
      ring[pl][slot] = h;
    }
  };
#pragma unroll
  for (int k = 0; k < kGradAhead; ++k) q[k] = fetch(k);
  // groups of 12 input rows, row r into ring slot r % 12; input row r
  // completes output row r - 10, staged in LDS slot (r - 10) % 12
  const int groups = (rows + 2 * kR + kRing - 1) / kRing;
  for (int g = 0; g < groups; ++g) {
#pragma unroll
    for (int j = 0; j < kRing; ++j) {
      const int r = g * kRing + j;
      row_in(r, j);
      const int o = r - 2 * kR;
      if (o < 0) continue;  // the first group's first ten rows only fill the ring
      float bl[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < kTaps; ++i) s = __builtin_fmaf(p.kernel[i], ring[pl][(j + kRing - 2 * kR + i) % kRing], s);
        bl[pl] = s;
      }
      float d = bl[0];
      d += 2.f * ox * bl[1];
      d += oy * bl[2];
      const float difference = ox - oy;
      const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
      float g_ = (1.f - w) * sign * inverse;
      g_ -= w * d;
      s_dl[o % kRing][3 * lane + c] = g_;
    }
    // all three channels of the output rows this group completed are staged:
    // write those inside the half unit as whole interleaved rows, one float
    // per thread
    __syncthreads();
    const int o_first = max(0, g * kRing - 2 * kR), o_end = min(rows, g * kRing + kRing - 2 * kR);
    const int e = (int)threadIdx.x;
    for (int o = o_first; o < o_end; ++o) {
      const int row_float = 3 * ((by + o) * W + sx);
      store_or_drop(rdl, e < valid_floats ? (uint32_t)((row_float + e) * 4) : kOff, s_dl[o % kRing][e]);
    }
    __syncthreads();
  }
}

// One workgroup per view: each thread sums a strided run of the view's
// half-unit sums in order, then a fixed-shape tree combines the 256 partials, so
// the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_loss_reduce(const ViewDesc* __restrict__ views,
                                                     const float* __restrict__ block_sums, float* __restrict__ out) {
  __shared__ float s_l1[256], s_ss[256];
  const int v = blockIdx.x;
  const ViewDesc& V = views[v];
  const int units = loss_halves(V.width, V.height);
  float l1 = 0.f, ssim = 0.f;
  for (int t = threadIdx.x; t < units; t += blockDim.x) {
    l1 += block_sums[2 * (V.lt_base + t)];
    ssim += block_sums[2 * (V.lt_base + t) + 1];
  }
  s_l1[threadIdx.x] = l1;
  s_ss[threadIdx.x] = ssim;
  __syncthreads();
  for (int width = 128; width > 0; width >>= 1) {
    if ((int)threadIdx.x < width) {
      s_l1[threadIdx.x] += s_l1[threadIdx.x + width];
      s_ss[threadIdx.x] += s_ss[threadIdx.x + width];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * v] = s_l1[0];
    out[2 * v + 1] = s_ss[0];
  }
}

}  // namespace

int loss_tiles(int width, int height) { return loss_halves(width, height); }

hipError_t launch_loss(const LossParams& p, const ViewDesc* views, const float* rgb, const float* targets,
                       float* dL, float* terms, long long total_px, float* block_sums, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  if (!(p.ssim_weight > 0.f)) {
    hipLaunchKernelGGL(k_loss_stats<false>, dim3(p.total_tiles), dim3(kStatWaves * 64), 0, stream, p, views, rgb,
                       targets, dL, terms, total_px, block_sums);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_loss_stats<true>, dim3(p.total_tiles), dim3(kStatWaves * 64), 0, stream, p, views, rgb, targets,
                     dL, terms, total_px, block_sums);
  SPLAT_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_loss_grad, dim3(p.total_tiles), dim3(kGradWaves * 64), 0, stream, p, views, rgb, targets,
                     terms, total_px, dL);
  return hipGetLastError();
}

hipError_t launch_loss_reduce(const ViewDesc* views, int n_views, const float* block_sums, float* out,
                              hipStream_t stream) {
  if (n_views <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loss_reduce, dim3(n_views), dim3(256), 0, stream, views, block_sums, out);
  return hipGetLastError();
}

}  // namespace splat
