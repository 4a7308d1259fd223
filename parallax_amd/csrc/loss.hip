// Fused L1 + D-SSIM photometric loss and its gradient.
//
// The engine's photometricLoss / meanSsim / blur (engine/native/cpp/src/splat/
// SplatRaster.cpp:580-738): L = (1-w) mean|x-y| + w (1 - mean SSIM) with an
// 11-tap sigma 1.5 Gaussian window applied as a zero-padded separable blur,
// C1 = 0.01^2, C2 = 0.03^2, means over 3 * pixels. The gradient of the SSIM
// term is itself a blur of per-pixel partials (the blur is self-adjoint on a
// zero-padded domain).
//
// Two passes over 16x16 output tiles of every view of the batch, one 256-thread
// workgroup per tile, each staging a 26x26 halo tile per channel in LDS:
//   pass 1: the 5 statistics planes (x, y, xx, yy, xy) blurred horizontally
//           then vertically in LDS, the SSIM map, the L1 gradient, and the five
//           per-pixel partial planes written to HBM; per-tile L1 and SSIM sums.
//   pass 2: the five partial planes blurred the same way and combined into
//           dL/drgb.
// Tap order, zero padding and every expression follow the engine's, so
// dL/drgb matches the CPU checker to float rounding of identical operations.
// HBM-bound: pass 1 reads 24 B and writes 72 B per pixel, pass 2 reads
// 84 B and writes 12 B per pixel (halo re-reads hit in L2).
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr int kR = 5;
constexpr int kHalo = kTile + 2 * kR;  // 26
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

__device__ __forceinline__ int find_view(const ViewDesc* views, int n_views, int tile) {
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].tile_base <= tile) v = k;
  return v;
}

// Blur `planes` LDS planes of kHalo x kHalo (zero outside the image) into
// one value per plane at this thread's pixel: horizontal pass over all halo
// rows into `hbuf`, then vertical. Tap order -5..5 as the engine's blur.
template <int kPlanes>
__device__ __forceinline__ void blur_tile(const float (*src)[kHalo][kHalo], float (*hbuf)[kHalo][kTile],
                                          const float* kernel, float* out) {
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kPlanes * kHalo * kTile; idx += blockDim.x) {
    const int pl = idx / (kHalo * kTile);
    const int rem = idx % (kHalo * kTile);
    const int row = rem / kTile, col = rem % kTile;
    float sum = 0.f;
#pragma unroll
    for (int i = 0; i < 2 * kR + 1; ++i) sum += kernel[i] * src[pl][row][col + i];
    hbuf[pl][row][col] = sum;
  }
  __syncthreads();
  const int ly = tid >> 4, lx = tid & 15;
#pragma unroll
  for (int pl = 0; pl < kPlanes; ++pl) {
    float sum = 0.f;
#pragma unroll
    for (int i = 0; i < 2 * kR + 1; ++i) sum += kernel[i] * hbuf[pl][ly + i][lx];
    out[pl] = sum;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_loss_forward(LossParams p, const ViewDesc* __restrict__ views,
                                                      const float* __restrict__ rgb, const float* __restrict__ target,
                                                      float* __restrict__ dL, float* __restrict__ terms,
                                                      long long total_px, float* __restrict__ block_sums) {
  __shared__ float s_src[5][kHalo][kHalo];
  __shared__ float s_h[5][kHalo][kTile];
  __shared__ float s_red[2][4];
  const int tile = xcd_remap(blockIdx.x, p.total_tiles);
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  const int tid = threadIdx.x;
  const int px = tx * kTile + (tid & 15), py = ty * kTile + (tid >> 4);
  const bool inside = px < V.width && py < V.height;
  const long long pixels = (long long)V.width * V.height;
  const float inverse = 1.f / (float)(pixels * 3);
  const float scale = inverse;
  const float w = p.ssim_weight;
  float l1 = 0.f, ssim = 0.f;
  const long long pixel = V.pixel_base + (long long)py * V.width + px;
  for (int c = 0; c < 3; ++c) {
    for (int idx = tid; idx < kHalo * kHalo; idx += blockDim.x) {
      const int hy = idx / kHalo, hx = idx % kHalo;
      const int gx = tx * kTile + hx - kR, gy = ty * kTile + hy - kR;
      float x = 0.f, y = 0.f;
      if (gx >= 0 && gx < V.width && gy >= 0 && gy < V.height) {
        const long long q = V.pixel_base + (long long)gy * V.width + gx;
        x = rgb[3 * q + c];
        y = target[3 * q + c];
      }
      s_src[0][hy][hx] = x;
      s_src[1][hy][hx] = y;
      s_src[2][hy][hx] = x * x;
      s_src[3][hy][hx] = y * y;
      s_src[4][hy][hx] = x * y;
    }
    __syncthreads();
    float st[5];
    blur_tile<5>(s_src, s_h, p.kernel, st);
    if (inside) {
      const float x = rgb[3 * pixel + c], y = target[3 * pixel + c];
      const float difference = x - y;
      l1 += fabsf(difference);
      const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
      dL[3 * pixel + c] = (1.f - w) * sign * inverse;
      if (w > 0.f) {
        const float mu_x = st[0], mu_y = st[1], exx = st[2], eyy = st[3], exy = st[4];
        const float var_x = exx - mu_x * mu_x;
        const float var_y = eyy - mu_y * mu_y;
        const float cov = exy - mu_x * mu_y;
        const float a = 2.f * mu_x * mu_y + kC1;
        const float b = 2.f * cov + kC2;
        const float cc = mu_x * mu_x + mu_y * mu_y + kC1;
        const float d = var_x + var_y + kC2;
        ssim += (a * b) / (cc * d);
        const float term_mu = scale * (2.f * mu_y * b / (cc * d) - 2.f * mu_x * a * b / (cc * cc * d));
        const float term_var = scale * (-(a * b) / (cc * d * d));
        const float term_cov = scale * (2.f * a / (cc * d));
        float* tp = terms + (long long)(c * 5) * total_px + pixel;
        tp[0] = term_mu;
        tp[total_px] = term_var;
        tp[2 * total_px] = term_var * mu_x;
        tp[3 * total_px] = term_cov;
        tp[4 * total_px] = term_cov * mu_y;
      }
    }
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
    block_sums[2 * tile] = ((s_red[0][0] + s_red[0][1]) + s_red[0][2]) + s_red[0][3];
    block_sums[2 * tile + 1] = ((s_red[1][0] + s_red[1][1]) + s_red[1][2]) + s_red[1][3];
  }
}

__global__ __launch_bounds__(256) void k_loss_backward(LossParams p, const ViewDesc* __restrict__ views,
                                                       const float* __restrict__ rgb,
                                                       const float* __restrict__ target,
                                                       const float* __restrict__ terms, long long total_px,
                                                       float* __restrict__ dL) {
  __shared__ float s_src[5][kHalo][kHalo];
  __shared__ float s_h[5][kHalo][kTile];
  const int tile = xcd_remap(blockIdx.x, p.total_tiles);
  const int v = find_view(views, p.n_views, tile);
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  const int tid = threadIdx.x;
  const int px = tx * kTile + (tid & 15), py = ty * kTile + (tid >> 4);
  const bool inside = px < V.width && py < V.height;
  const long long pixel = V.pixel_base + (long long)py * V.width + px;
  for (int c = 0; c < 3; ++c) {
    for (int idx = tid; idx < kHalo * kHalo; idx += blockDim.x) {
      const int hy = idx / kHalo, hx = idx % kHalo;
      const int gx = tx * kTile + hx - kR, gy = ty * kTile + hy - kR;
      const bool in = gx >= 0 && gx < V.width && gy >= 0 && gy < V.height;
      const long long q = V.pixel_base + (long long)gy * V.width + gx;
#pragma unroll
      for (int k = 0; k < 5; ++k) s_src[k][hy][hx] = in ? terms[(long long)(c * 5 + k) * total_px + q] : 0.f;
    }
    __syncthreads();
    float bl[5];
    blur_tile<5>(s_src, s_h, p.kernel, bl);
    if (inside) {
      const float x = rgb[3 * pixel + c], y = target[3 * pixel + c];
      float d = 0.f;
      d += bl[0];
      d += 2.f * x * bl[1];
      d -= 2.f * bl[2];
      d += y * bl[3];
      d -= bl[4];
      dL[3 * pixel + c] -= p.ssim_weight * d;
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
  const int tiles = V.cols * V.rows;
  float l1 = 0.f, ssim = 0.f;
  for (int t = threadIdx.x; t < tiles; t += blockDim.x) {
    l1 += block_sums[2 * (V.tile_base + t)];
    ssim += block_sums[2 * (V.tile_base + t) + 1];
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

hipError_t launch_loss_forward(const LossParams& p, const ViewDesc* views, const float* rgb, const float* target,
                               float* dL, float* terms, long long total_px, float* block_sums, hipStream_t stream) {
  if (p.total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loss_forward, dim3(p.total_tiles), dim3(256), 0, stream, p, views, rgb, target, dL, terms,
                     total_px, block_sums);
  return hipGetLastError();
}

hipError_t launch_loss_backward(const LossParams& p, const ViewDesc* views, const float* rgb, const float* target,
                                const float* terms, long long total_px, float* dL, hipStream_t stream) {
  if (p.total_tiles <= 0 || !(p.ssim_weight > 0.f)) return hipSuccess;
  hipLaunchKernelGGL(k_loss_backward, dim3(p.total_tiles), dim3(256), 0, stream, p, views, rgb, target, terms,
                     total_px, dL);
  return hipGetLastError();
}

hipError_t launch_loss_reduce(const ViewDesc* views, int n_views, const float* block_sums, float* out,
                              hipStream_t stream) {
  if (n_views <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loss_reduce, dim3(n_views), dim3(256), 0, stream, views, block_sums, out);
  return hipGetLastError();
}

}  // namespace splat
