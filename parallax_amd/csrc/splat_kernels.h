// Host-side launchers of the splat kernels (internal to libparallax_splat).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "splat_common.h"

namespace splat {

// ---- projection ----
// rec[slot] and keys[slot] (depth bits; 0xFFFFFFFF for a culled slot). With
// prects (every view of the batch at most kPackedTiles tiles a side):
// prects[slot] = the packed tile rect (kNoRect for a culled slot), and counts,
// scounts, rects are not written. Otherwise counts[slot] = tiles of the
// slot's rect, scounts[slot] = super-tiles of it, rects[slot] = the rect
// ((x0 | x1 << 16, y0 | y1 << 16), 8 bytes) for the binning's gathers.
// With key_ranges (device, 2 * ceil(n / 64) words), wave w of gaussians
// [64 w, 64 w + 64) writes the minimum and maximum of its kept keys to
// key_ranges[2 w], key_ranges[2 w + 1] (0xFFFFFFFF, 0 when it kept none).
hipError_t launch_project(const float* params, long long cap, int n, int rest_stride, int sh_use,
                          const ViewDesc* views, int n_views, float dilation, float min_alpha, SplatRec* rec,
                          uint32_t* keys, uint32_t* counts, uint32_t* scounts, uint2* rects, uint32_t* prects,
                          hipStream_t stream, uint32_t* key_ranges = nullptr);

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

// What a scan takes of each input word: the word, or the tiles / super-tiles
// of a packed rect.
enum { kScanPlain = 0, kScanRectTiles = 1, kScanRectSupers = 2 };
// out[i] = sum_{j<i} f(in[idx ? idx[j] : j]); *total (device, nullable) = sum
// of all; f per xform. With n_dev the item count is min(*n_dev, n) (n sizes
// the launches); out past it is not written. With side_total (device),
// *side_total += the sum of side_xform of the same inputs.
hipError_t launch_exclusive_scan(const uint32_t* in, const uint32_t* gather_idx, uint32_t* out, long long n,
                                 uint32_t* total, const ScanWorkspace& ws, hipStream_t stream,
                                 int xform = kScanPlain, const uint32_t* n_dev = nullptr,
                                 uint32_t* side_total = nullptr, int side_xform = kScanPlain);

struct SortWorkspace {
  uint32_t* hist = nullptr;  // 256 * blocks
  long long hist_capacity = 0;
  ScanWorkspace scan;
};
long long sort_hist_elems(long long max_n);

// Stable LSD radix sort of (key, value) pairs over key bits [begin_bit,
// end_bit), 8-bit digits (a plan may take 9); key bits above end_bit ride along. Ping-pongs
// between (keys, vals) and (keys_alt, vals_alt); *result_in_alt tells which
// pair holds the sorted output. With n_dev, the item count is read on the
// device (clamped to n, which then only sizes the launches), so the sort can
// sit in a captured graph whose item count changes from replay to replay.
// With vals2 (and vals2_alt), a second value array is permuted alongside and
// ends in the same buffer of the pair as the first. With identity_vals, the
// values start as the item indices (vals is not read, only written). With
// n_kept (device), items whose key is 0xFFFFFFFF are dropped, *n_kept is set
// to the count of the rest, and the sorted output holds those only.
//
// With a plan (begin_bit 0), the passes to run are read on the device:
// each pass covers digit_bits bits (8 or 9), and a pass at or past *passes
// does nothing (each of its kernels returns at once). The output always ends
// in (keys, vals, vals2) and *result_in_alt is false. With third buffers
// (keys_c ...) the input is in (keys, vals, vals2); without them it must be
// in the first set for an even number of passes and in the second
// (keys_alt ...) for an odd one. With key_base, the first pass rewrites each
// kept key k as k - *key_base, with, when *view_shift < 32, the item's view
// (item index / view_size) placed at bit *view_shift.
struct SortPlan {
  const uint32_t* passes;      // device: the passes to run
  const uint32_t* active;      // device, per pass: 0 when it is skipped (caps its scans' item count)
  const uint32_t* key_base;    // device, or null: no rewrite
  const uint32_t* view_shift;  // device, or null
  uint32_t view_size;
  int digit_bits;
  uint32_t* keys_c;            // third buffers, each of n elements, or null
  uint32_t* vals_c;
  uint32_t* vals2_c;           // with vals2 only
};
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, long long n,
                             int begin_bit, int end_bit, const SortWorkspace& ws, bool* result_in_alt,
                             hipStream_t stream, const uint32_t* n_dev = nullptr, uint32_t* vals2 = nullptr,
                             uint32_t* vals2_alt = nullptr, bool identity_vals = false, uint32_t* n_kept = nullptr,
                             const SortPlan* plan = nullptr);
// The plans of the binning's two sorts, from the projection's per-wave key
// ranges (n_waves (min, max) pairs), into plan[kPlanWords]: the depth sort's
// passes, per-pass active words, key base and view shift (32: no view bits),
// and the super-tile sort's passes and active words. view_bits: bits of the
// batch's largest view index when super-tiles are numbered view * 256 +
// super-tile within the view (every view at most 256 of them), else -1.
constexpr int kDepthDigitBits = 9;
constexpr int kPlanPasses = 0, kPlanActive = 1, kPlanKeyBase = 5, kPlanViewShift = 6, kPlanSuperPasses = 7,
              kPlanSuperActive = 8, kPlanWords = 10;
hipError_t launch_sort_plan(const uint32_t* key_ranges, int n_waves, int view_bits, uint32_t* plan,
                            hipStream_t stream);
// *out (device) += the sum of in[0..n).
hipError_t launch_sum_u32(const uint32_t* in, long long n, uint32_t* out, hipStream_t stream);

// ---- binning (binning.hip) ----
constexpr int kBinSegment = 256;   // splats per segment of a super-tile list
constexpr int kStBits = 20;        // super-tile id bits of a super-tile key
constexpr uint32_t kStMask = (1u << kStBits) - 1u;

// For each depth-sorted slot s < min(m, *m_dev): writes one (super-tile key,
// slot) pair per super-tile of its rect at soffsets[s] (*s_total: all pairs);
// a key holds the global super-tile id in its low kStBits bits and the
// splat's rect inside that super-tile above them. With pair_base:
// pair_base[slot] = offsets[s]. Pairs at or past s_cap are dropped (the
// capacity check flags the step).
// The rects are read from sorted_prects (the packed rects in depth order)
// when given, else gathered from rects[slot]. The pairs go to (skeys,
// svals), or to (skeys_odd, svals_odd) when sort_passes (device, nullable:
// the next sort's pass count under its plan) is odd.
hipError_t launch_emit_super(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* soffsets,
                             const uint32_t* s_total, const uint2* rects, const uint32_t* sorted_prects,
                             const ViewDesc* views, long long per_view, long long m, const uint32_t* m_dev,
                             uint32_t s_cap, uint32_t* skeys, uint32_t* svals, uint32_t* skeys_odd,
                             uint32_t* svals_odd, const uint32_t* sort_passes, uint32_t* pair_base,
                             hipStream_t stream);
// sranges[2*st] / [2*st+1] = [begin, end) of super-tile st in the sorted
// pairs; the pair count is min(*n_dev, n) with n_dev, else n.
hipError_t launch_super_ranges(const uint32_t* sorted_keys, long long n, const uint32_t* n_dev, uint32_t* sranges,
                               int n_st, hipStream_t stream);
// totals[3] = 1 if totals[0] (tile pairs) > p_cap or totals[1] (super-tile
// pairs) > s_cap, else 0.
hipError_t launch_check_caps(uint32_t* totals, uint32_t p_cap, uint32_t s_cap, hipStream_t stream);
// nseg[st] = segments of kBinSegment splats in super-tile st's list.
hipError_t launch_segment_counts(const uint32_t* sranges, int n_st, uint32_t* nseg, hipStream_t stream);
// seg_st[seg_base[st] + s] = st for every segment s of st.
hipError_t launch_segment_list(const uint32_t* nseg, const uint32_t* seg_base, int n_st, uint32_t* seg_st,
                               hipStream_t stream);
// Per (super-tile, tile, segment) counts into table[0 .. 64*ns_max] (the last
// entry and those of missing segments zeroed); *ns_total = real segments
// (device), ns_max >= *ns_total.
hipError_t launch_tile_count(const uint32_t* skeys, const uint32_t* sranges,
                             const uint32_t* nseg, const uint32_t* seg_base, const uint32_t* seg_st,
                             const uint32_t* ns_total, uint32_t ns_max, uint32_t* table, hipStream_t stream);
// Every tile's list, in depth order, into pair_vals at the runs `scan` (the
// exclusive scan of the count table) assigns; entries at or past p_cap are
// dropped.
hipError_t launch_tile_write(const uint32_t* skeys, const uint32_t* svals,
                             const uint32_t* sranges, const uint32_t* nseg, const uint32_t* seg_base,
                             const uint32_t* seg_st, const uint32_t* ns_total, uint32_t ns_max, const uint32_t* scan,
                             uint32_t p_cap, uint32_t* pair_vals, hipStream_t stream);
// ranges[2*t] / ranges[2*t+1] = [begin, end) of global tile t in pair_vals,
// clamped to p_cap.
hipError_t launch_tile_ranges_table(const ViewDesc* views, int n_views, int total_tiles, const uint32_t* scan,
                                    const uint32_t* nseg, const uint32_t* seg_base, uint32_t p_cap, uint32_t* ranges,
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
// Launch shape of the compositing kernels (performance only: every choice
// gives the same image, and the same gradients in deterministic mode).
struct RasterConfig {
  int fwd_shape = 0;     // forward workgroup: 0 16x16 (4 waves), 1 8x8 (1 wave), 2 32x8 (4 waves)
  bool fwd_queue = false;  // persistent work queue, longest tiles first
  int bwd_batch = 128;   // adjoint: splats staged per batch (64 or 128)
  bool bwd_queue = false;
  int cus = 256;         // compute units: persistent grids are sized from it
};
const char* raster_shape_name(int shape);
// keys / vals for a sort of the tiles by list length, longest first.
hipError_t launch_tile_order_keys(const uint32_t* ranges, int total_tiles, uint32_t* keys, uint32_t* vals,
                                  hipStream_t stream);
// `order` (tiles, longest first) and `counter` (one u32, cleared here) are
// read only by the queue variants.
hipError_t launch_forward(const RasterParams& p, const RasterConfig& cfg, const ViewDesc* views,
                          const uint32_t* ranges, const uint32_t* pair_vals, const SplatRec* rec, float* rgb, float* T,
                          int* last, const uint32_t* order, uint32_t* counter, hipStream_t stream);
// Atomic mode: screen accumulators (kScreenStride floats per slot) via
// wave-reduced atomics. Deterministic mode: per-pair partials written to
// partial[pair_base[slot] + local pair] (kPartialStride floats each), for the
// pairs the adjoint visits only, each tagged in slot 11 with *epoch; a record
// whose tag is not the current epoch reads as zero. visit[pair] (p_cap
// bytes) takes the epoch's low byte beside each record written, so the
// reduction loads only records whose byte matches.
hipError_t launch_backward(const RasterParams& p, const RasterConfig& cfg, const ViewDesc* views,
                           const uint32_t* ranges, const uint32_t* pair_vals, const SplatRec* rec,
                           const float* dL_drgb, const float* T, const int* last, float* screen,
                           const uint32_t* pair_base, float* partial, uint8_t* visit, uint32_t p_cap,
                           bool deterministic,
                           const uint32_t* epoch, const uint32_t* order, uint32_t* counter, hipStream_t stream);

// ---- loss ----
struct LossParams {
  float kernel[11];
  float ssim_weight;
  int n_views;
  int total_tiles;  // loss blocks (64x78 pixels, half units) over all views
};
// L1 + SSIM loss terms and dL/drgb of every view of the batch (two passes
// over 128x78-pixel loss units; `terms` holds kSsimPlanes floats per pixel
// between them); per-block (64x78 pixels) partial sums of L1 and SSIM into
// block_sums (2 floats per block; loss_tiles counts a view's blocks). Each
// view's target starts at targets + views[v].target_base.
constexpr int kSsimPlanes = 9;  // three partial planes per channel
int loss_tiles(int width, int height);
hipError_t launch_loss(const LossParams& p, const ViewDesc* views, const float* rgb, const float* targets,
                       float* dL, float* terms, long long total_px, float* block_sums, hipStream_t stream);
// Fixed-order reduction of the block sums into one (l1, ssim) pair per view.
hipError_t launch_loss_reduce(const ViewDesc* views, int n_views, const float* block_sums, float* out,
                              hipStream_t stream);
// Looks up, once per device, the device address of the fixed scratch array
// behind splat_loss_noalloc; `device` must be current.
hipError_t noalloc_attach(int device);

// ---- per-gaussian adjoint ----
struct AdjointParams {
  long long cap;
  int n;
  int rest_stride;
  int sh_use;
  int n_views;
  int deterministic;
};
struct AdamStep;
// Adam fused into the per-gaussian adjoint (a training step): see k_adjoint.
struct AdamFuse {
  const AdamStep* step = nullptr;
  const uint32_t* skip = nullptr;  // device flag: when set, the plain adjoint's accumulation only
  float* params = nullptr;         // the same buffer as the adjoint's params
  float* m = nullptr;
  float* v = nullptr;
  float* stat_sg = nullptr;
  int* stat_vc = nullptr;
  bool stored_clear = false;  // the stored gradient, sg and vc are all zeros: not read, not cleared
};
hipError_t launch_adjoint(const AdjointParams& p, const ViewDesc* views, const float* params, const SplatRec* rec,
                          const float* screen, float* grads, float* sg, int* vc, hipStream_t stream,
                          const AdamFuse* fuse = nullptr);
// Deterministic mode: screen[slot] = the slot's per-tile partials (its run
// [offsets[i], offsets[i+1]) of depth-sorted position i < min(m, *m_dev),
// the last run ending at *p_total, all clamped to p_cap) summed in a fixed
// order. `carry` holds det_reduce_carry_floats(p_cap) floats.
// Records tagged with another epoch than *epoch count as zeros; a record
// whose visit byte is not the epoch's low byte is one and is not loaded.
long long det_reduce_carry_floats(unsigned long long p_cap);
hipError_t launch_det_reduce(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* p_total,
                             long long m, const uint32_t* m_dev, uint32_t p_cap, const float* partial,
                             const uint8_t* visit, const uint32_t* epoch, float* screen, float* carry,
                             hipStream_t stream);
// The next epoch of the deterministic partials: *epoch + 1, skipping 0 (the
// tag of a cleared buffer). A device-side counter, so each replay of a
// captured step takes a new one; a record can be mistaken for current only
// after 2^32 - 1 further passes.
hipError_t launch_bump_epoch(uint32_t* epoch, hipStream_t stream);

// ---- optimiser ----
// The per-step scalars, read from device memory so a captured step can take
// new ones each replay.
struct AdamStep {
  float lr[6];  // position, log_scale, rotation, opacity, dc, rest
  float beta1, beta2, eps, c1, c2;
};
struct AdamShape {
  long long cap;
  int n;
  int rest_stride;
};
// Skips the update (and leaves gradients and statistics in place) when
// skip_flag is non-null and *skip_flag != 0.
hipError_t launch_adam(const AdamShape& s, const AdamStep* step, const uint32_t* skip_flag, float* params,
                       float* grads, float* m, float* v, float* sg, int* vc, float* stat_sg, int* stat_vc,
                       hipStream_t stream);

}  // namespace splat
