// splat_loss_noalloc: the photometric loss of one image held in the caller's
// device memory, with every intermediate in one fixed device array of this
// file (`scratch`), so the call allocates nothing.
//
// It runs the kernels splat_loss runs (k_loss_stats, k_loss_grad,
// k_loss_reduce in loss.hip) on one view whose pixels start at the caller's
// pointers, and finishes the loss value on the device with the float
// operations splat_loss evaluates on the host:
//   loss = (1 - w) * (l1 * inv) + w * (1 - ssim * inv),  inv = 1 / (3 W H)
// (or l1 * inv when w <= 0), so both entry points give the same bits.
//
// Scratch layout (one array per device, held until the process ends):
//   [ SSIM partial planes: kSsimPlanes * kMaxPixels floats ]
//   [ per-loss-tile L1 / SSIM sums: 2 * kMaxLossTiles floats ]
//   [ the view descriptor ][ the view's two sums ]
// Every value the kernels read is written earlier in the same call (the
// partial planes for every pixel of the image, a sum for every loss tile of
// the view, the descriptor and the two sums), so nothing needs clearing.
// Calls on one device share the array: calls that overlap on one device --
// two host threads, or two streams -- race.
#include <mutex>
#include <string>

#include "../../include/splat_hip.h"
#include "splat_common.h"
#include "splat_context.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr long long kMaxPixels = 1LL << 21;  // 2,097,152: covers 1920x1080
constexpr int kMaxSide = 8192;
constexpr int kMaxLossTiles = 4096;  // >= loss_tiles(W, H) for W, H <= kMaxSide, W*H <= kMaxPixels
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kTermsOffset = 0;
constexpr size_t kSumsOffset = align256(kTermsOffset + sizeof(float) * kSsimPlanes * kMaxPixels);
constexpr size_t kViewOffset = align256(kSumsOffset + sizeof(float) * 2 * kMaxLossTiles);
constexpr size_t kTotalsOffset = align256(kViewOffset + sizeof(ViewDesc));
constexpr size_t kScratchBytes = align256(kTotalsOffset + sizeof(float) * 2);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// scratch's device address on each device, from hipGetSymbolAddress with that
// device current (the array's host-side name is not a device address)
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

__global__ void k_put_view(ViewDesc d, ViewDesc* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = d;
}

__global__ void k_loss_value(const float* __restrict__ totals, float w, float inverse, float* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float l1 = totals[0] * inverse;
  *loss = w <= 0.f ? l1 : (1.f - w) * l1 + w * (1.f - totals[1] * inverse);
}

splat_status fail(splat_context* ctx, splat_status st, const std::string& msg) {
  ctx->error = msg;
  return st;
}

}  // namespace

// Looks up scratch's address on `device` (which must be current) once.
hipError_t noalloc_attach(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[device] = static_cast<unsigned char*>(ptr);
  return hipSuccess;
}

}  // namespace splat

using namespace splat;

extern "C" splat_status splat_loss_noalloc(splat_context* ctx, const float* rgb, const float* target, int32_t width,
                                          int32_t height, float ssim_weight, float* dL_drgb, float* loss) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!rgb || !target || !dL_drgb || !loss)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "loss_noalloc: null image, target, gradient or loss pointer");
  if (width <= 0 || height <= 0)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "loss_noalloc: width and height must be positive");
  const long long px = (long long)width * height;
  if (width > kMaxSide || height > kMaxSide || px > kMaxPixels || loss_tiles(width, height) > kMaxLossTiles)
    return fail(ctx, SPLAT_ERROR_CAPACITY,
                "loss_noalloc: " + std::to_string(width) + "x" + std::to_string(height) +
                    " exceeds the fixed scratch (at most 2097152 pixels, each side at most 8192)");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
  }
  if (!base) return fail(ctx, SPLAT_ERROR_STATE, "loss_noalloc: scratch not attached to this device");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return fail(ctx, SPLAT_ERROR_DEVICE, std::string("set device: ") + hipGetErrorString(e));

  ViewDesc d{};
  d.width = width;
  d.height = height;
  d.cols = (width + kTile - 1) / kTile;
  d.rows = (height + kTile - 1) / kTile;
  d.lt_base = 0;
  d.pixel_base = 0;
  d.target_base = 0;
  ViewDesc* d_view = reinterpret_cast<ViewDesc*>(base + kViewOffset);
  float* terms = reinterpret_cast<float*>(base + kTermsOffset);
  float* block_sums = reinterpret_cast<float*>(base + kSumsOffset);
  float* totals = reinterpret_cast<float*>(base + kTotalsOffset);

  LossParams lp{};
  for (int i = 0; i < 11; ++i) lp.kernel[i] = ctx->ssim_kernel[i];
  lp.ssim_weight = ssim_weight;
  lp.n_views = 1;
  lp.total_tiles = loss_tiles(width, height);
  hipLaunchKernelGGL(k_put_view, dim3(1), dim3(64), 0, ctx->stream, d, d_view);
  e = hipGetLastError();
  if (e == hipSuccess) e = launch_loss(lp, d_view, rgb, target, dL_drgb, terms, px, block_sums, ctx->stream);
  if (e == hipSuccess) e = launch_loss_reduce(d_view, 1, block_sums, totals, ctx->stream);
  if (e == hipSuccess) {
    const float inverse = 1.f / static_cast<float>(px * 3);
    hipLaunchKernelGGL(k_loss_value, dim3(1), dim3(64), 0, ctx->stream, totals, ssim_weight, inverse, loss);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(ctx, SPLAT_ERROR_DEVICE, std::string("loss_noalloc: ") + hipGetErrorString(e));
  return SPLAT_OK;
}
