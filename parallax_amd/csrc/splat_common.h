// Device-side layouts and constants shared by the splat kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "splat_exp.h"

namespace splat {

constexpr int kTile = 16;             // logical tile edge (SplatRaster.cpp:18)
constexpr float kMaxAlpha = 0.99f;    // SplatRaster.cpp:22
constexpr float kSlack = 1.3f;        // SplatRaster.cpp:26
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr int kScreenStride = 12;     // floats per (view, splat) screen accumulator
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
  int pad0;
  long long pixel_base;  // first pixel of this view in the batch image buffers
};

// A projected splat of one (view, gaussian) slot; 64 bytes. Written by the
// projection for every slot; `flags` bit 31 marks a projected (not culled)
// splat.
struct alignas(16) SplatRec {
  float mx, my;        // mean in pixels
  float ca, cb, cc;    // conic (inverse 2D covariance: xx, xy, yy)
  float opacity;       // sigmoid(logit)
  float cr, cg, cbl;   // SH colour, clamped at zero
  float depth;
  float radius;        // 3 sigma of the major axis, pixels
  uint32_t flags;      // bits 0-2 colour clamped, bits 3-4 frustum clamped, bit 31 projected
  int16_t x0, x1, y0, y1;  // 16x16 tile rect, inclusive
  float cov_xx, cov_yy;    // unused by the kernels; kept for diagnostics
};
static_assert(sizeof(SplatRec) == 64, "SplatRec must be 64 bytes");

constexpr uint32_t kProjected = 0x80000000u;

__host__ __device__ __forceinline__ float det_exp(float x) { return splat_det::exp(x); }

// Depth in an unsigned key that sorts like the float (depth >= 0.01 > 0).
__device__ __forceinline__ uint32_t depth_key(float d) { return __float_as_uint(d); }

}  // namespace splat

#define SPLAT_HIP_CHECK(expr)                                   \
  do {                                                          \
    hipError_t _e = (expr);                                     \
    if (_e != hipSuccess) return _e;                            \
  } while (0)
