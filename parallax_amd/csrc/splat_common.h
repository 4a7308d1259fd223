// Device-side layouts and constants shared by the splat kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "splat_exp.h"

namespace splat {

constexpr int kTile = 16;             // logical tile edge (SplatRaster.cpp:18)
constexpr int kSuper = 8;             // super-tile edge, in tiles (binning only)
constexpr float kMaxAlpha = 0.99f;    // SplatRaster.cpp:22
constexpr float kSlack = 1.3f;        // SplatRaster.cpp:26
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr int kScreenStride = 12;     // floats per (view, splat) screen accumulator
// floats per deterministic (tile, splat) partial: 64-byte records, each one
// whole memory sector, so their scattered stores need no read-modify-write
constexpr int kPartialStride = 16;
constexpr int kMaxBatchViews = 64;

// One camera of a batch, in the float form the projection uses
// (SplatRaster.cpp:206-212): pose cast per view, camera centre -R^T t
// evaluated in double then cast, tan of the half field of view.
struct ViewDesc {
  float W[9];
  float t[3];
  float C[3];
  float fx, fy, cx, cy;
  float thx, thy;
  int width, height;
  int cols, rows;      // 16x16 tiles
  int tile_base;       // first global tile id of this view in the batch
  int st_base;         // first global super-tile id (8x8 tiles) of this view
  int scols, srows;    // super-tiles
  int lt_base;         // first loss block (64x78 pixels) of this view
  long long pixel_base;  // first pixel of this view in the batch image buffers
  long long target_base;  // first float of this view's target image in the target store
};

// A projected splat of one (view, gaussian) slot; 64 bytes. Written by the
// projection for every slot; `flags` bit 31 marks a projected (not culled)
// splat. The compositing kernels read the first 48 bytes as three float4.
struct alignas(16) SplatRec {
  float mx, my;        // mean in pixels
  float ca, cb, cc;    // conic (inverse 2D covariance: xx, xy, yy)
  float opacity;       // sigmoid(logit)
  float cr, cg, cbl;   // SH colour, clamped at zero
  float ex, ey;        // half extents (pixels) of a box holding every pixel centre
                       // where alpha >= min_alpha can hold; +inf: no bound
  float depth;
  float radius;        // 3 sigma of the major axis, pixels
  uint32_t flags;      // bits 0-2 colour clamped, bits 3-4 frustum clamped, bit 31 projected
  int16_t x0, x1, y0, y1;  // 16x16 tile rect, inclusive
};
static_assert(sizeof(SplatRec) == 64, "SplatRec must be 64 bytes");

constexpr uint32_t kProjected = 0x80000000u;

__host__ __device__ __forceinline__ float det_exp(float x) { return splat_det::exp(x); }
// det_exp for the compositing loop's falloff, whose argument is <= 0.
__host__ __device__ __forceinline__ float det_exp_nonpos(float x) { return splat_det::exp_nonpos(x); }

// Workgroups go to the 8 XCDs of the device round-robin. Give each XCD one
// contiguous run of the n work items (tiles), so items that share data --
// neighbouring tiles share splats and image halos -- meet in the same L2.
// A bijection of [0, n) for any n.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int kXcds = 8;
  const int xcd = b % kXcds, k = b / kXcds;
  const int per = n / kXcds, extra = n % kXcds;
  return xcd * per + min(xcd, extra) + k;
}

// Depth in an unsigned key that sorts like the float (depth >= 0.01 > 0).
__device__ __forceinline__ uint32_t depth_key(float d) { return __float_as_uint(d); }

// A slot's tile rect packed into 32 bits for the binning, where every view of
// the batch has at most kPackedTiles tile columns and rows (4096 pixels):
// x0 | x1 << 8 | y0 << 16 | y1 << 24, inclusive; kNoRect (x0 > x1) for a
// slot that enters no tile.
constexpr int kPackedTiles = 256;
constexpr uint32_t kNoRect = 0x000000FFu;
__host__ __device__ __forceinline__ uint32_t pack_rect(int x0, int x1, int y0, int y1) {
  return (uint32_t)x0 | (uint32_t)x1 << 8 | (uint32_t)y0 << 16 | (uint32_t)y1 << 24;
}
__host__ __device__ __forceinline__ int rect_x0(uint32_t r) { return (int)(r & 255u); }
__host__ __device__ __forceinline__ int rect_x1(uint32_t r) { return (int)((r >> 8) & 255u); }
__host__ __device__ __forceinline__ int rect_y0(uint32_t r) { return (int)((r >> 16) & 255u); }
__host__ __device__ __forceinline__ int rect_y1(uint32_t r) { return (int)(r >> 24); }
// tiles of a packed rect
__host__ __device__ __forceinline__ uint32_t rect_tiles(uint32_t r) {
  if (rect_x0(r) > rect_x1(r)) return 0u;
  return (uint32_t)((rect_x1(r) - rect_x0(r) + 1) * (rect_y1(r) - rect_y0(r) + 1));
}
// super-tiles (kSuper x kSuper tiles) of a packed rect
__host__ __device__ __forceinline__ uint32_t rect_supers(uint32_t r) {
  if (rect_x0(r) > rect_x1(r)) return 0u;
  return (uint32_t)((rect_x1(r) / kSuper - rect_x0(r) / kSuper + 1) * (rect_y1(r) / kSuper - rect_y0(r) / kSuper + 1));
}

}  // namespace splat

#define SPLAT_HIP_CHECK(expr)                                   \
  do {                                                          \
    hipError_t _e = (expr);                                     \
    if (_e != hipSuccess) return _e;                            \
  } while (0)
