// Host-side launchers of the splat kernels (internal to libparallax_splat).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "splat_common.h"

namespace splat {

// ---- projection ----
hipError_t launch_project(const float* params, long long cap, int n, int rest_stride, int sh_use,
                          const ViewDesc* views, int n_views, float dilation, float min_alpha, SplatRec* rec,
                          uint32_t* keys, uint32_t* vals, uint32_t* counts, hipStream_t stream);

// ---- scan / sort ----
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096 elements per block

// Scratch of a multi-level exclusive scan of up to `max_n` elements.
struct ScanWorkspace {
  uint32_t* buf = nullptr;  // per-level block sums, concatenated
  long long capacity = 0;   // elements in buf
  long long max_n = 0;
};
long long scan_workspace_elems(long long max_n);

// out[i] = sum_{j<i} in[idx ? idx[j] : j]; *total (device, nullable) = sum of all.
hipError_t launch_exclusive_scan(const uint32_t* in, const uint32_t* gather_idx, uint32_t* out, long long n,
                                 uint32_t* total, const ScanWorkspace& ws, hipStream_t stream);

struct SortWorkspace {
  uint32_t* hist = nullptr;  // 256 * blocks
  long long hist_capacity = 0;
  ScanWorkspace scan;
};
long long sort_hist_elems(long long max_n);

// Stable LSD radix sort of (key, value) pairs over key bits [begin_bit,
// end_bit), 8-bit digits. Ping-pongs between (keys, vals) and (keys_alt,
// vals_alt); *result_in_alt tells which pair holds the sorted output.
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, long long n,
                             int begin_bit, int end_bit, const SortWorkspace& ws, bool* result_in_alt,
                             hipStream_t stream);

// ---- binning ----
// For each depth-sorted slot s: writes the pair keys (global tile id) and
// values (slot) of its tile rect at offsets[s], and pair_base[slot].
hipError_t launch_emit_pairs(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* counts,
                             const SplatRec* rec, const ViewDesc* views, int n_views, long long per_view,
                             long long m, uint32_t* pair_keys, uint32_t* pair_vals, uint32_t* pair_base,
                             hipStream_t stream);
// ranges[2*t] / ranges[2*t+1] = [begin, end) of global tile t in the sorted pairs.
hipError_t launch_tile_ranges(const uint32_t* sorted_keys, long long p, uint32_t* ranges, int total_tiles,
                              hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* ptr, uint32_t value, long long n, hipStream_t stream);
hipError_t launch_fill_f32(float* ptr, float value, long long n, hipStream_t stream);

// ---- compositing ----
struct RasterParams {
  float bg[3];
  float min_alpha;
  float cutoff;
  int n_views;
  int total_tiles;
  long long per_view;  // slots per view (= gaussian count)
};
hipError_t launch_forward(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                          const uint32_t* pair_vals, const SplatRec* rec, float* rgb, float* T, int* last,
                          hipStream_t stream);
// Atomic mode: screen accumulators (kScreenStride floats per slot) via
// wave-reduced atomics. Deterministic mode: per-pair partials written to
// partial[pair_base[slot] + local pair] (10 floats each).
hipError_t launch_backward(const RasterParams& p, const ViewDesc* views, const uint32_t* ranges,
                           const uint32_t* pair_vals, const SplatRec* rec, const float* dL_drgb, const float* T,
                           const int* last, float* screen, const uint32_t* pair_base, float* partial,
                           bool deterministic, hipStream_t stream);

// ---- loss ----
struct LossParams {
  float kernel[11];
  float ssim_weight;
  int n_views;
  int total_tiles;  // 16x16 tiles over all views
};
// First pass: SSIM statistics, L1 gradient and the five term planes; per-block
// partial sums of L1 and SSIM into block_sums (2 floats per block).
hipError_t launch_loss_forward(const LossParams& p, const ViewDesc* views, const float* rgb, const float* target,
                               float* dL, float* terms, long long total_px, float* block_sums, hipStream_t stream);
// Second pass: blurs the term planes and finishes dL/drgb.
hipError_t launch_loss_backward(const LossParams& p, const ViewDesc* views, const float* rgb, const float* target,
                                const float* terms, long long total_px, float* dL, hipStream_t stream);
// Fixed-order reduction of the block sums into one (l1, ssim) pair per view.
hipError_t launch_loss_reduce(const ViewDesc* views, int n_views, const float* block_sums, float* out,
                              hipStream_t stream);

// ---- per-gaussian adjoint ----
struct AdjointParams {
  long long cap;
  int n;
  int rest_stride;
  int sh_use;
  int n_views;
  int deterministic;
};
hipError_t launch_adjoint(const AdjointParams& p, const ViewDesc* views, const float* params, const SplatRec* rec,
                          const float* screen, const uint32_t* pair_base, const uint32_t* counts,
                          const float* partial, float* grads, float* sg, int* vc, hipStream_t stream);

// ---- optimiser ----
struct AdamParams {
  float lr[6];
  float beta1, beta2, eps, c1, c2;
  long long cap;
  int n;
  int rest_stride;
};
hipError_t launch_adam(const AdamParams& p, float* params, float* grads, float* m, float* v, float* sg, int* vc,
                       float* stat_sg, int* stat_vc, hipStream_t stream);

}  // namespace splat
