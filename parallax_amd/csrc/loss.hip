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
// Two passes over 64x16-pixel loss tiles of every view of the batch, one
// 256-thread workgroup each, one channel at a time:
//   stats: the image/target window (74x26, zero outside the image) goes to
//          LDS; one row pass over the 26 rows blurs all five statistics
//          planes (x, y, xx, yy, xy) from a single read of each run's window,
//          then a column pass per plane down to the tile's 16 rows, every
//          thread sliding an 11-tap window over a run of 4 outputs in
//          registers; then
//          per pixel the SSIM value, the three partial planes (to HBM), and
//          the L1 terms; per-tile L1 and SSIM sums.
//   grad:  per partial plane, its 74x26 window to LDS (two buffers, so a
//          plane costs two workgroup barriers), the same two passes, and the
//          blurred planes combined into dL/drgb.
// Each blur tap is one fused multiply-add in tap order -5..5 and the loss
// value's expressions are the engine's, as the CPU checker evaluates them;
// dL/drgb differs from the checker's only by the rounding of the plane sum A.
// Compulsory HBM traffic per pixel: 24 B read and 36 B written by the first
// pass, 60 B read and 12 B written by the second (halos hit in cache).
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr int kR = 5;
constexpr int kLW = 64, kLH = 16;         // loss tile, pixels
constexpr int kWinW = kLW + 2 * kR;       // 74: window columns
constexpr int kWinH = kLH + 2 * kR;       // 26: window rows
constexpr int kRun = 4;                   // outputs per thread along a blur
constexpr int kTaps = kRun + 2 * kR;      // 14 window values per run
constexpr int kPitch = kWinW + 1;         // 75: odd, conflict-free row walks
constexpr int kHPitch = kLW + 1;          // 65
constexpr int kWinPerThread = (kWinH * kWinW + 255) / 256;  // window values each of 256 threads fetches
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

__device__ __forceinline__ int find_view_lt(const ViewDesc* views, int n_views, int lt) {
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].lt_base <= lt) v = k;
  return v;
}

// out[k] = sum_i kernel[i] * win[k + i] for k < kRun, one fused multiply-add
// per tap in tap order.
__device__ __forceinline__ void blur_run(const float (&win)[kTaps], const float* kernel, float (&out)[kRun]) {
#pragma unroll
  for (int k = 0; k < kRun; ++k) {
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 2 * kR + 1; ++i) s = __builtin_fmaf(kernel[i], win[k + i], s);
    out[k] = s;
  }
}

struct LossTile {
  int ox, oy;  // first pixel
  const ViewDesc* V;
};

__device__ __forceinline__ LossTile loss_tile_of(const LossParams& p, const ViewDesc* views) {
  const int lt = xcd_remap(blockIdx.x, p.total_tiles);
  const int v = find_view_lt(views, p.n_views, lt);
  const ViewDesc& V = views[v];
  const int local = lt - V.lt_base;
  const int ltcols = (V.width + kLW - 1) / kLW;
  LossTile t;
  t.ox = (local % ltcols) * kLW;
  t.oy = (local / ltcols) * kLH;
  t.V = &V;
  return t;
}

// Where this thread's window elements (tid + 256 r, row-major over the
// 74 x 26 window) come from: the pixel offset within the view (0 where the
// element lies outside the image or past the window), and a bit per element
// that is set when it lies inside both.
struct WindowMap {
  int off[kWinPerThread];
  uint32_t inside;
};

__device__ __forceinline__ WindowMap window_map(const LossTile& L, const ViewDesc& V) {
  WindowMap m;
  m.inside = 0u;
#pragma unroll
  for (int r = 0; r < kWinPerThread; ++r) {
    const int idx = (int)threadIdx.x + r * 256;
    const int hy = idx / kWinW, hx = idx % kWinW;
    const int qx = L.ox + hx - kR, qy = L.oy + hy - kR;
    const bool in = idx < kWinH * kWinW && qx >= 0 && qx < V.width && qy >= 0 && qy < V.height;
    m.off[r] = in ? qy * V.width + qx : 0;
    m.inside |= in ? 1u << r : 0u;
  }
  return m;
}

// Row pass of one plane over the window's 26 rows into hbuf[26][65]: plane 0
// blurs a, 1 blurs b, 2 a*a, 3 b*b, 4 a*b (a, b: LDS windows).
__device__ __forceinline__ void row_pass(int plane, const float (*a)[kPitch], const float (*b)[kPitch],
                                         const float* kernel, float (*hbuf)[kHPitch]) {
  // a wave covers 4 window rows x 16 runs: with both pitches odd mod 4 the
  // 64 lanes of every LDS access land in 64 different banks
  for (int r = (int)(threadIdx.x >> 4); r < kWinH; r += (int)(blockDim.x >> 4)) {
    const int j0 = (int)(threadIdx.x & 15) * kRun;
    float win[kTaps], o[kRun];
#pragma unroll
    for (int i = 0; i < kTaps; ++i) {
      const float x = a[r][j0 + i];
      if (plane == 0) {
        win[i] = x;
      } else {
        const float y = b[r][j0 + i];
        win[i] = plane == 1 ? y : (plane == 2 ? x * x : (plane == 3 ? y * y : x * y));
      }
    }
    blur_run(win, kernel, o);
#pragma unroll
    for (int k = 0; k < kRun; ++k) hbuf[r][j0 + k] = o[k];
  }
}

// The row passes of all five statistics planes at once: each run's window of
// a and b is read from LDS once.
__device__ __forceinline__ void row_pass5(const float (*a)[kPitch], const float (*b)[kPitch], const float* kernel,
                                          float (*hbuf)[kWinH][kHPitch]) {
  for (int r = (int)(threadIdx.x >> 4); r < kWinH; r += (int)(blockDim.x >> 4)) {
    const int j0 = (int)(threadIdx.x & 15) * kRun;
    float x[kTaps], y[kTaps], win[kTaps], o[kRun];
#pragma unroll
    for (int i = 0; i < kTaps; ++i) {
      x[i] = a[r][j0 + i];
      y[i] = b[r][j0 + i];
    }
#pragma unroll
    for (int pl = 0; pl < 5; ++pl) {
#pragma unroll
      for (int i = 0; i < kTaps; ++i)
        win[i] = pl == 0 ? x[i] : (pl == 1 ? y[i] : (pl == 2 ? x[i] * x[i] : (pl == 3 ? y[i] * y[i] : x[i] * y[i])));
      blur_run(win, kernel, o);
#pragma unroll
      for (int k = 0; k < kRun; ++k) hbuf[pl][r][j0 + k] = o[k];
    }
  }
}

// Column pass: this thread's 4 outputs (column j, rows r0..r0+3 of the tile).
__device__ __forceinline__ void col_pass(const float (*hbuf)[kHPitch], const float* kernel, int j, int r0,
                                         float (&out)[kRun]) {
  float win[kTaps];
#pragma unroll
  for (int i = 0; i < kTaps; ++i) win[i] = hbuf[r0 + i][j];
  blur_run(win, kernel, out);
}

__global__ __launch_bounds__(256) void k_loss_stats(LossParams p, const ViewDesc* __restrict__ views,
                                                    const float* __restrict__ rgb,
                                                    const float* __restrict__ targets, float* __restrict__ dL,
                                                    float* __restrict__ terms, long long total_px,
                                                    float* __restrict__ block_sums) {
  __shared__ float s_x[kWinH][kPitch];
  __shared__ float s_y[kWinH][kPitch];
  __shared__ float s_h[5][kWinH][kHPitch];
  __shared__ float s_red[2][4];
  const LossTile L = loss_tile_of(p, views);
  const ViewDesc& V = *L.V;
  const int tid = threadIdx.x;
  const int j = tid & (kLW - 1), r0 = (tid / kLW) * kRun;  // this thread's column and rows
  const long long pixels = (long long)V.width * V.height;
  const float inverse = 1.f / (float)(pixels * 3);
  const float w = p.ssim_weight;
  const float* img = rgb + 3 * V.pixel_base;
  const float* tgt = targets + V.target_base;
  const int gx = L.ox + j;
  float l1 = 0.f, ssim = 0.f;
  // each channel's image/target window, prefetched into registers while the
  // previous channel is blurred
  const WindowMap win = window_map(L, V);
  float px_[kWinPerThread], py_[kWinPerThread];
  auto fetch = [&](int c) {
#pragma unroll
    for (int r = 0; r < kWinPerThread; ++r) {
      const float x = img[3 * win.off[r] + c], y = tgt[3 * win.off[r] + c];
      const bool in = (win.inside >> r) & 1u;
      px_[r] = in ? x : 0.f;
      py_[r] = in ? y : 0.f;
    }
  };
  fetch(0);
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int r = 0; r < kWinPerThread; ++r) {
      const int idx = tid + r * 256;
      if (idx < kWinH * kWinW) {
        s_x[idx / kWinW][idx % kWinW] = px_[r];
        s_y[idx / kWinW][idx % kWinW] = py_[r];
      }
    }
    __syncthreads();
    if (c + 1 < 3) fetch(c + 1);
    float st[5][kRun];
    if (w > 0.f) {
      row_pass5(s_x, s_y, p.kernel, s_h);
      __syncthreads();
#pragma unroll
      for (int pl = 0; pl < 5; ++pl) col_pass(s_h[pl], p.kernel, j, r0, st[pl]);
    }
#pragma unroll
    for (int k = 0; k < kRun; ++k) {
      const int gy = L.oy + r0 + k;
      if (gx >= V.width || gy >= V.height) continue;
      const long long pixel = V.pixel_base + (long long)gy * V.width + gx;
      const float x = s_x[r0 + k + kR][j + kR], y = s_y[r0 + k + kR][j + kR];
      const float difference = x - y;
      l1 += fabsf(difference);
      if (w > 0.f) {
        const float mu_x = st[0][k], mu_y = st[1][k], exx = st[2][k], eyy = st[3][k], exy = st[4][k];
        const float var_x = exx - mu_x * mu_x;
        const float var_y = eyy - mu_y * mu_y;
        const float cov = exy - mu_x * mu_y;
        const float a = 2.f * mu_x * mu_y + kC1;
        const float b = 2.f * cov + kC2;
        const float cc = mu_x * mu_x + mu_y * mu_y + kC1;
        const float d = var_x + var_y + kC2;
        ssim += (a * b) / (cc * d);
        const float term_mu = inverse * (2.f * mu_y * b / (cc * d) - 2.f * mu_x * a * b / (cc * cc * d));
        const float term_var = inverse * (-(a * b) / (cc * d * d));
        const float term_cov = inverse * (2.f * a / (cc * d));
        float* tp = terms + (long long)(c * 3) * total_px + pixel;
        tp[0] = (term_mu - 2.f * term_var * mu_x) - term_cov * mu_y;
        tp[total_px] = term_var;
        tp[2 * total_px] = term_cov;
      } else {
        const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
        dL[3 * pixel + c] = sign * inverse;
      }
    }
    __syncthreads();
  }
  // fixed-order workgroup sums
  float a = l1, b = ssim;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if ((tid & 63) == 0) {
    s_red[0][tid >> 6] = a;
    s_red[1][tid >> 6] = b;
  }
  __syncthreads();
  if (tid == 0) {
    const int lt = xcd_remap(blockIdx.x, p.total_tiles);
    block_sums[2 * lt] = ((s_red[0][0] + s_red[0][1]) + s_red[0][2]) + s_red[0][3];
    block_sums[2 * lt + 1] = ((s_red[1][0] + s_red[1][1]) + s_red[1][2]) + s_red[1][3];
  }
}

__global__ __launch_bounds__(256) void k_loss_grad(LossParams p, const ViewDesc* __restrict__ views,
                                                   const float* __restrict__ rgb, const float* __restrict__ targets,
                                                   const float* __restrict__ terms, long long total_px,
                                                   float* __restrict__ dL) {
  // plane q stages through buffer q & 1: the barrier after a plane's window
  // store also orders the previous plane's column reads before this plane's
  // row pass overwrites the other buffer
  __shared__ float s_t[2][kWinH][kPitch];
  __shared__ float s_h[2][kWinH][kHPitch];
  const LossTile L = loss_tile_of(p, views);
  const ViewDesc& V = *L.V;
  const int tid = threadIdx.x;
  const int j = tid & (kLW - 1), r0 = (tid / kLW) * kRun;
  const long long pixels = (long long)V.width * V.height;
  const float inverse = 1.f / (float)(pixels * 3);
  const float w = p.ssim_weight;
  const float* img = rgb + 3 * V.pixel_base;
  const float* tgt = targets + V.target_base;
  const int gx = L.ox + j;
  // the 9 partial planes' windows, each prefetched into registers while the
  // previous one is blurred; the window's element offsets (0 outside the
  // image, then masked) are the same for every plane, so the loads need no
  // branches
  const WindowMap win = window_map(L, V);
  float pre[kWinPerThread];
  auto fetch = [&](int q) {
    const float* plane = terms + (long long)q * total_px + V.pixel_base;
#pragma unroll
    for (int r = 0; r < kWinPerThread; ++r) {
      const float x = plane[win.off[r]];
      pre[r] = (win.inside >> r) & 1u ? x : 0.f;
    }
  };
  fetch(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float d[kRun], xs[kRun], ys[kRun];
#pragma unroll
    for (int k = 0; k < kRun; ++k) {
      const int gy = L.oy + r0 + k;
      xs[k] = ys[k] = 0.f;
      if (gx < V.width && gy < V.height) {
        const long long qq = (long long)gy * V.width + gx;
        xs[k] = img[3 * qq + c];
        ys[k] = tgt[3 * qq + c];
      }
    }
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      const int q = 3 * c + pl, buf = q & 1;
#pragma unroll
      for (int r = 0; r < kWinPerThread; ++r) {
        const int idx = tid + r * 256;
        if (idx < kWinH * kWinW) s_t[buf][idx / kWinW][idx % kWinW] = pre[r];
      }
      __syncthreads();
      if (q + 1 < kSsimPlanes) fetch(q + 1);
      row_pass(0, s_t[buf], s_t[buf], p.kernel, s_h[buf]);
      __syncthreads();
      float bl[kRun];
      col_pass(s_h[buf], p.kernel, j, r0, bl);
#pragma unroll
      for (int k = 0; k < kRun; ++k) {
        if (pl == 0) d[k] = bl[k];
        else if (pl == 1) d[k] += 2.f * xs[k] * bl[k];
        else d[k] += ys[k] * bl[k];
      }
    }
#pragma unroll
    for (int k = 0; k < kRun; ++k) {
      const int gy = L.oy + r0 + k;
      if (gx >= V.width || gy >= V.height) continue;
      const float difference = xs[k] - ys[k];
      const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
      float g = (1.f - w) * sign * inverse;
      g -= w * d[k];
      dL[3 * (V.pixel_base + (long long)gy * V.width + gx) + c] = g;
    }
  }
}

// One workgroup per view: each thread sums a strided run of the view's tile
// sums in tile order, then a fixed-shape tree combines the 256 partials, so
// the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_loss_reduce(const ViewDesc* __restrict__ views,
                                                     const float* __restrict__ block_sums, float* __restrict__ out) {
  __shared__ float s_l1[256], s_ss[256];
  const int v = blockIdx.x;
  const ViewDesc& V = views[v];
  const int tiles = ((V.width + kLW - 1) / kLW) * ((V.height + kLH - 1) / kLH);
  float l1 = 0.f, ssim = 0.f;
  for (int t = threadIdx.x; t < tiles; t += blockDim.x) {
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

int loss_tiles(int width, int height) { return ((width + kLW - 1) / kLW) * ((height + kLH - 1) / kLH); }

hipError_t launch_loss(const LossParams& p, const ViewDesc* views, const float* rgb, const float* targets,
                       float* dL, float* terms, long long total_px, float* block_sums, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loss_stats, dim3(p.total_tiles), dim3(256), 0, stream, p, views, rgb, targets, dL, terms,
                     total_px, block_sums);
  SPLAT_HIP_CHECK(hipGetLastError());
  if (!(p.ssim_weight > 0.f)) return hipSuccess;
  hipLaunchKernelGGL(k_loss_grad, dim3(p.total_tiles), dim3(256), 0, stream, p, views, rgb, targets, terms, total_px,
                     dL);
  return hipGetLastError();
}

hipError_t launch_loss_reduce(const ViewDesc* views, int n_views, const float* block_sums, float* out,
                              hipStream_t stream) {
  if (n_views <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loss_reduce, dim3(n_views), dim3(256), 0, stream, views, block_sums, out);
  return hipGetLastError();
}

}  // namespace splat
