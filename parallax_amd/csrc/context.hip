// C ABI of the splat trainer: context, device buffers, and the per-batch
// launch sequence (projection -> depth sort -> pair emit -> tile sort ->
// ranges -> forward -> loss -> adjoint -> per-gaussian adjoint -> Adam).
// Host code only; the kernels live in project/sort/raster/loss/adjoint/adam.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/splat_hip.h"
#include "splat_common.h"
#include "splat_context.h"
#include "splat_kernels.h"

using namespace splat;

namespace {

thread_local std::string g_no_context_error;

splat_status fail(splat_context* ctx, splat_status st, const std::string& msg) {
  if (ctx) ctx->error = msg;
  else g_no_context_error = msg;
  return st;
}

splat_status hip_fail(splat_context* ctx, hipError_t e, const char* what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? SPLAT_ERROR_OUT_OF_MEMORY : SPLAT_ERROR_DEVICE, msg);
}

#define CK(expr, what)                                   \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
  } while (0)

// Float view as the projection uses it (SplatRaster.cpp:206-212).
ViewDesc make_desc(const splat_view& v) {
  ViewDesc d{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d.W[3 * i + j] = static_cast<float>(v.R[3 * i + j]);
    d.t[i] = static_cast<float>(v.t[i]);
    const double c = (-v.R[0 * 3 + i]) * v.t[0] + (-v.R[1 * 3 + i]) * v.t[1] + (-v.R[2 * 3 + i]) * v.t[2];
    d.C[i] = static_cast<float>(c);
  }
  d.fx = static_cast<float>(v.fx);
  d.fy = static_cast<float>(v.fy);
  d.cx = static_cast<float>(v.cx);
  d.cy = static_cast<float>(v.cy);
  d.thx = static_cast<float>(0.5 * v.width / v.fx);
  d.thy = static_cast<float>(0.5 * v.height / v.fy);
  d.width = v.width;
  d.height = v.height;
  d.cols = (v.width + kTile - 1) / kTile;
  d.rows = (v.height + kTile - 1) / kTile;
  d.scols = (d.cols + kSuper - 1) / kSuper;
  d.srows = (d.rows + kSuper - 1) / kSuper;
  return d;
}

int bit_width(unsigned long long x) {
  int b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

}  // namespace

// ------------------------------------------------------------- buffers --
namespace {
std::atomic<unsigned long long> g_alloc_generation{1};
}
unsigned long long alloc_generation() { return g_alloc_generation.load(); }

hipError_t DeviceBuffer::reserve(size_t want) {
  if (want <= bytes) return hipSuccess;
  g_alloc_generation.fetch_add(1);
  if (ptr) {
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return e;
  }
  size_t grow = want + want / 4 + 256;
  hipError_t e = hipMalloc(&ptr, grow);
  if (e != hipSuccess) {
    ptr = nullptr;
    return e;
  }
  bytes = grow;
  return hipSuccess;
}

void DeviceBuffer::release() {
  if (ptr) g_alloc_generation.fetch_add(1);
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void splat_context::drop_graphs() {
  for (Graph& g : graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs.clear();
}

splat_context::~splat_context() {
  drop_graphs();
  if (readback) (void)hipHostFree(readback);
  for (DeviceBuffer* b : all_buffers()) b->release();
  for (auto& kv : timers) {
    for (auto& pr : kv.second.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  }
  if (stream) (void)hipStreamDestroy(stream);
}

std::vector<DeviceBuffer*> splat_context::all_buffers() {
  return {&params, &grads, &m, &v, &sg, &vc, &stat_sg, &stat_vc, &targets, &d_views, &rec, &keys, &vals, &keys_alt,
          &vals_alt, &counts, &offsets, &pair_base, &scounts, &scounts_alt, &soffsets, &keys_c, &vals_c, &scounts_c, &key_ranges, &rects, &skeys, &svals, &skeys_alt, &svals_alt,
          &sranges, &nseg, &seg_base, &seg_st, &bin_table, &pair_vals, &ranges,
          &rgb, &T, &last, &dL, &terms, &block_sums, &loss_sums, &screen, &partial, &visit, &carry, &scan_ws, &hist_ws, &scalars,
          &scratch_params, &scratch_m, &scratch_v, &dens_flags, &dens_a, &dens_b, &dens_c, &dens_src, &d_adam,
          &order_keys, &order_vals, &order_keys_alt, &order_vals_alt};
}

// ------------------------------------------------------------- timing --
hipError_t splat_context::timed_begin(const char* name, hipEvent_t* start) {
  if (!timing || capturing) return hipSuccess;
  SPLAT_HIP_CHECK(hipEventCreate(start));
  (void)name;
  return hipEventRecord(*start, stream);
}

hipError_t splat_context::timed_end(const char* name, hipEvent_t start) {
  if (!timing || capturing) return hipSuccess;
  hipEvent_t stop;
  SPLAT_HIP_CHECK(hipEventCreate(&stop));
  SPLAT_HIP_CHECK(hipEventRecord(stop, stream));
  timers[name].pending.push_back({start, stop});
  return hipSuccess;
}

hipError_t splat_context::collect_timing() {
  for (auto& kv : timers) {
    for (auto& pr : kv.second.pending) {
      SPLAT_HIP_CHECK(hipEventSynchronize(pr.second));
      float ms = 0.f;
      SPLAT_HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
      kv.second.total_ms += ms;
      kv.second.launches += 1;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    kv.second.pending.clear();
  }
  return hipSuccess;
}

#define TIMED(name, call)                                           \
  do {                                                              \
    hipEvent_t _t0 = nullptr;                                       \
    CK(ctx->timed_begin(name, &_t0), "event");                      \
    CK((call), name);                                               \
    CK(ctx->timed_end(name, _t0), "event");                         \
  } while (0)

// ------------------------------------------------------------- pipeline --
namespace {

// The u32 device scalars: [0] tile pairs, [1] super-tile pairs, [3] overflow
// flag, [4] work-queue counter, [5] projected slots (the depth-sorted ones),
// [8] the deterministic partials' epoch, [9..14] the depth sort's plan.
constexpr int kSortedScalar = 5;
constexpr int kEpochScalar = 8;
constexpr int kSortPlanScalar = 9;
constexpr int kScalarWords = kSortPlanScalar + kPlanWords;

splat_status check_ready(splat_context* ctx) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) return fail(ctx, SPLAT_ERROR_STATE, "no capacity reserved: call splat_set_capacity first");
  return SPLAT_OK;
}

// Host side of a batch: descriptors, totals, and every buffer whose size does
// not depend on the pair counts. No device work but allocation.
splat_status batch_setup(splat_context* ctx, int n_views, const int32_t* view_ids, const splat_raster_options* o) {
  if (n_views <= 0 || n_views > kMaxBatchViews || !view_ids || !o)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "render_forward: 1..64 views and options required");
  for (int k = 0; k < n_views; ++k)
    if (view_ids[k] < 0 || view_ids[k] >= (int)ctx->views.size())
      return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "render_forward: view id out of range");
  ctx->opts = *o;
  ctx->batch_ids.assign(view_ids, view_ids + n_views);
  ctx->batch_desc.clear();
  int tiles = 0, supers = 0, ltiles = 0;
  long long px = 0;
  // super-tiles numbered view * 256 + super-tile within the view when every
  // view has at most 256 of them: the super-tile sort's digit is then the
  // id within the view (the sort plan's view grouping does the rest)
  bool st_by_view = true;
  for (int k = 0; k < n_views; ++k) {
    const ViewDesc d = make_desc(ctx->views[view_ids[k]]);
    if (d.scols * d.srows > 256) st_by_view = false;
  }
  for (int k = 0; k < n_views; ++k) {
    ViewDesc d = make_desc(ctx->views[view_ids[k]]);
    d.tile_base = tiles;
    d.st_base = st_by_view ? 256 * k : supers;
    d.pixel_base = px;
    d.target_base = ctx->has_target[view_ids[k]] ? ctx->target_offset[view_ids[k]] : 0;
    d.lt_base = ltiles;
    ltiles += loss_tiles(d.width, d.height);
    tiles += d.cols * d.rows;
    supers = st_by_view ? 256 * (k + 1) : supers + d.scols * d.srows;
    px += (long long)d.width * d.height;
    ctx->batch_desc.push_back(d);
  }
  ctx->st_view_bits = st_by_view ? bit_width((unsigned long long)(n_views - 1)) : -1;
  if (supers > (int)kStMask) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "render_forward: batch too large");
  ctx->packed = true;
  for (const ViewDesc& d : ctx->batch_desc)
    if (d.cols > kPackedTiles || d.rows > kPackedTiles) ctx->packed = false;
  ctx->shape_key.clear();
  for (const ViewDesc& d : ctx->batch_desc) {
    ctx->shape_key += std::to_string(d.width) + "x" + std::to_string(d.height) + ",";
  }
  {
    const auto it = ctx->tuned.find(ctx->shape_key);
    const int cus = ctx->rcfg.cus;
    ctx->rcfg = ctx->forced ? ctx->forced_cfg : (it != ctx->tuned.end() ? it->second : RasterConfig{});
    ctx->rcfg.cus = cus;
  }
  ctx->total_tiles = tiles;
  ctx->loss_tiles = ltiles;
  ctx->n_st = supers;
  ctx->total_px = px;
  ctx->sh_use = std::min(o->sh_degree, ctx->sh_degree);
  ctx->forward_done = false;
  ctx->loss_done = false;
  const long long n = ctx->n;
  const long long M = n * n_views;
  ctx->M = M;
  RasterParams rp{};
  for (int c = 0; c < 3; ++c) rp.bg[c] = o->background[c];
  rp.min_alpha = o->min_alpha;
  rp.cutoff = o->transmittance_cutoff;
  rp.n_views = n_views;
  rp.total_tiles = tiles;
  rp.per_view = n;
  ctx->raster = rp;
  CK(ctx->d_views.reserve(sizeof(ViewDesc) * n_views), "alloc views");
  CK(ctx->rgb.reserve(sizeof(float) * 3 * std::max(1LL, px)), "alloc image");
  CK(ctx->T.reserve(sizeof(float) * std::max(1LL, px)), "alloc image");
  CK(ctx->last.reserve(sizeof(int) * std::max(1LL, px)), "alloc image");
  CK(ctx->ranges.reserve(sizeof(uint32_t) * 2 * std::max(1, tiles)), "alloc ranges");
  CK(ctx->order_keys.reserve(sizeof(uint32_t) * std::max(1, tiles)), "alloc tile order");
  CK(ctx->order_vals.reserve(sizeof(uint32_t) * std::max(1, tiles)), "alloc tile order");
  CK(ctx->order_keys_alt.reserve(sizeof(uint32_t) * std::max(1, tiles)), "alloc tile order");
  CK(ctx->order_vals_alt.reserve(sizeof(uint32_t) * std::max(1, tiles)), "alloc tile order");
  CK(ctx->scalars.reserve(4 * 32), "alloc scalars");
  static_assert(kScalarWords <= 32, "device scalars");
  const long long Mc = std::max(1LL, M);
  CK(ctx->rec.reserve(sizeof(SplatRec) * Mc), "alloc records");
  CK(ctx->keys.reserve(4 * Mc), "alloc keys");
  CK(ctx->vals.reserve(4 * Mc), "alloc keys");
  CK(ctx->keys_alt.reserve(4 * Mc), "alloc keys");
  CK(ctx->vals_alt.reserve(4 * Mc), "alloc keys");
  CK(ctx->counts.reserve(4 * Mc), "alloc counts");
  CK(ctx->offsets.reserve(4 * Mc), "alloc offsets");
  CK(ctx->pair_base.reserve(4 * Mc), "alloc pair base");
  CK(ctx->scounts.reserve(4 * Mc), "alloc counts");
  CK(ctx->scounts_alt.reserve(4 * Mc), "alloc counts");
  CK(ctx->keys_c.reserve(4 * Mc), "alloc keys");
  CK(ctx->vals_c.reserve(4 * Mc), "alloc keys");
  CK(ctx->scounts_c.reserve(4 * Mc), "alloc counts");
  CK(ctx->key_ranges.reserve(8 * std::max(1LL, ((long long)ctx->n + 63) / 64)), "alloc key ranges");
  CK(ctx->rects.reserve(8 * Mc), "alloc rects");
  CK(ctx->soffsets.reserve(4 * Mc), "alloc offsets");
  CK(ctx->sranges.reserve(sizeof(uint32_t) * 2 * std::max(1, supers)), "alloc ranges");
  CK(ctx->nseg.reserve(sizeof(uint32_t) * std::max(1, supers)), "alloc segments");
  CK(ctx->seg_base.reserve(sizeof(uint32_t) * std::max(1, supers)), "alloc segments");
  CK(ctx->ensure_sort_workspace(std::max(M, 256LL * 4096)), "alloc sort workspace");
  return SPLAT_OK;
}

splat_status upload_batch(splat_context* ctx) {
  CK(hipMemcpyAsync(ctx->d_views.ptr, ctx->batch_desc.data(), sizeof(ViewDesc) * ctx->batch_desc.size(),
                    hipMemcpyHostToDevice, ctx->stream),
     "upload views");
  return SPLAT_OK;
}

// The deterministic adjoint's per-pair partials and their visit bytes. A new
// allocation is cleared: tag 0 is no epoch, so every record reads as zero
// until written (a visit byte of 0 only sends the reduction to the record's
// own tag). A reallocation always grows `bytes`; the allocator may return
// the freed address, so the pointer cannot tell.
splat_status reserve_partials(splat_context* ctx, unsigned long long p_cap) {
  p_cap = std::max<unsigned long long>(1, p_cap);
  const size_t before = ctx->partial.bytes;
  CK(ctx->partial.reserve(sizeof(float) * kPartialStride * p_cap), "alloc partials");
  if (ctx->partial.bytes != before)
    CK(launch_fill_f32(ctx->partial.f(), 0.f, (long long)(ctx->partial.bytes / sizeof(float)), ctx->stream),
       "clear partials");
  const size_t visit_before = ctx->visit.bytes;
  CK(ctx->visit.reserve((p_cap + 3) & ~3ULL), "alloc visit bytes");
  if (ctx->visit.bytes != visit_before)
    CK(launch_fill_u32(ctx->visit.u(), 0u, (long long)(ctx->visit.bytes / sizeof(uint32_t)), ctx->stream),
       "clear visit bytes");
  return SPLAT_OK;
}

// Buffers sized by the pair counts: tile pairs p_cap, super-tile pairs s_cap.
splat_status reserve_pairs(splat_context* ctx, unsigned long long p_cap, unsigned long long s_cap) {
  p_cap = std::max(1ULL, p_cap);
  s_cap = std::max(1ULL, s_cap);
  if (p_cap > 0xFFFFFFF0ULL || s_cap > 0xFFFFFFF0ULL)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "render_forward: more than 2^32 pairs in one batch");
  ctx->P_cap = p_cap;
  ctx->S_cap = s_cap;
  // segments: at most ceil(len / kBinSegment) per super-tile
  const unsigned long long ns_max = s_cap / kBinSegment + (unsigned long long)ctx->n_st;
  CK(ctx->pair_vals.reserve(4 * (p_cap + 1)), "alloc pairs");  // + the spare entry overflowing writes land on
  CK(ctx->skeys.reserve(4 * s_cap), "alloc super pairs");
  CK(ctx->svals.reserve(4 * s_cap), "alloc super pairs");
  CK(ctx->skeys_alt.reserve(4 * s_cap), "alloc super pairs");
  CK(ctx->svals_alt.reserve(4 * s_cap), "alloc super pairs");
  CK(ctx->seg_st.reserve(4 * ns_max), "alloc segments");
  CK(ctx->bin_table.reserve(4 * (64ull * ns_max + 1)), "alloc binning table");
  CK(ctx->ensure_sort_workspace(std::max<long long>(std::max<long long>((long long)s_cap, ctx->M),
                                                    (long long)(64ull * ns_max + 1))),
     "alloc sort workspace");
  if (ctx->opts.deterministic) {
    if (splat_status st = reserve_partials(ctx, p_cap)) return st;
    CK(ctx->carry.reserve(sizeof(float) * det_reduce_carry_floats(p_cap)), "alloc carries");
  }
  CK(ctx->screen.reserve(sizeof(float) * kScreenStride * std::max(1LL, ctx->M)), "alloc screen");
  return SPLAT_OK;
}

// Projection, depth sort and the pair counts: scalars[0] = tile pairs P,
// scalars[1] = super-tile pairs S; scalars[3] (overflow) cleared.
splat_status enqueue_front(splat_context* ctx) {
  const long long n = ctx->n, M = ctx->M;
  const int n_views = (int)ctx->batch_desc.size();
  const splat_raster_options* o = &ctx->opts;
  uint32_t* d_total = ctx->scalars.u();
  CK(launch_fill_u32(d_total, 0u, 4, ctx->stream), "clear counters");
  const bool packed = ctx->packed;
  TIMED("project", launch_project(ctx->params.f(), ctx->cap, (int)n, ctx->rs, ctx->sh_use, ctx->views_dev(), n_views,
                                  o->screen_dilation, o->min_alpha, ctx->recs(), ctx->keys.u(),
                                  packed ? nullptr : ctx->counts.u(), packed ? nullptr : ctx->scounts.u(),
                                  packed ? nullptr : reinterpret_cast<uint2*>(ctx->rects.u()),
                                  packed ? ctx->scounts.u() : nullptr, ctx->stream, ctx->key_ranges.u()));
  bool alt = false;
  // the packed rects (or the super-tile counts) ride along with the slots, so
  // what reads them in depth order needs no gather; the slots start as the
  // item indices, and culled slots (key 0xFFFFFFFF) leave in the first pass:
  // the sorted arrays hold the scalars[kSortedScalar] projected slots. The
  // keys are sorted less their minimum, in as many 8-bit passes as the
  // kept keys' range needs (the sort's plan, set on the device).
  uint32_t* n_sorted = d_total + kSortedScalar;
  uint32_t* plan_words = d_total + kSortPlanScalar;
  const SortPlan plan{plan_words + kPlanPasses, plan_words + kPlanActive, plan_words + kPlanKeyBase,
                      plan_words + kPlanViewShift, (uint32_t)n, kDepthDigitBits, ctx->keys_c.u(), ctx->vals_c.u(),
                      ctx->scounts_c.u()};
  TIMED("depth_sort", [&]() -> hipError_t {
    const hipError_t e = launch_sort_plan(ctx->key_ranges.u(), (int)((n + 63) / 64), ctx->st_view_bits, plan_words,
                                          ctx->stream);
    if (e != hipSuccess) return e;
    return launch_radix_sort(ctx->keys.u(), ctx->vals.u(), ctx->keys_alt.u(), ctx->vals_alt.u(), M, 0, 32,
                             ctx->sort_ws, &alt, ctx->stream, nullptr, ctx->scounts.u(), ctx->scounts_alt.u(), true,
                             n_sorted, &plan);
  }());
  ctx->sorted_slots = alt ? ctx->vals_alt.u() : ctx->vals.u();
  const uint32_t* sorted_v2 = alt ? ctx->scounts_alt.u() : ctx->scounts.u();
  ctx->sorted_prects = packed ? const_cast<uint32_t*>(sorted_v2) : nullptr;
  if (packed) {
    // the super-tile offsets in depth order are the emitter's places (their
    // total S, and on the side the tile-pair total P); the tile-pair offsets
    // are the deterministic adjoint's pair ids
    TIMED("pair_scan", launch_exclusive_scan(sorted_v2, nullptr, ctx->soffsets.u(), M, d_total + 1, ctx->sort_ws.scan,
                                             ctx->stream, kScanRectSupers, n_sorted, d_total, kScanRectTiles));
    if (o->deterministic)
      TIMED("pair_scan", launch_exclusive_scan(sorted_v2, nullptr, ctx->offsets.u(), M, nullptr, ctx->sort_ws.scan,
                                               ctx->stream, kScanRectTiles, n_sorted));
    return SPLAT_OK;
  }
  if (o->deterministic) {
    TIMED("pair_scan", launch_exclusive_scan(ctx->counts.u(), ctx->sorted_slots, ctx->offsets.u(), M, d_total,
                                             ctx->sort_ws.scan, ctx->stream, kScanPlain, n_sorted));
  } else {
    TIMED("pair_scan", launch_sum_u32(ctx->counts.u(), M, d_total, ctx->stream));
  }
  TIMED("pair_scan", launch_exclusive_scan(sorted_v2, nullptr, ctx->soffsets.u(), M, d_total + 1, ctx->sort_ws.scan,
                                           ctx->stream, kScanPlain, n_sorted));
  return SPLAT_OK;
}

splat_status enqueue_tile_order(splat_context* ctx);

// Binning within the reserved capacities (the pair counts are read on the
// device), then forward compositing.
splat_status enqueue_back(splat_context* ctx) {
  const long long n = ctx->n, M = ctx->M;
  const int n_views = (int)ctx->batch_desc.size();
  const int supers = ctx->n_st, tiles = ctx->total_tiles;
  const uint32_t p_cap = (uint32_t)ctx->P_cap, s_cap = (uint32_t)ctx->S_cap;
  const uint32_t ns_max = (uint32_t)(ctx->S_cap / kBinSegment + (unsigned long long)supers);
  uint32_t* d_total = ctx->scalars.u();
  CK(launch_check_caps(d_total, p_cap, s_cap, ctx->stream), "capacity check");
  uint32_t* pair_base = ctx->opts.deterministic ? ctx->pair_base.u() : nullptr;
  // with super-tiles numbered by view, the super-tile sort runs under the
  // plan the depth sort's plan kernel wrote: one pass over the id within a
  // view when the depth sort grouped the slots by view, else a second by view
  const bool st_by_view = ctx->st_view_bits >= 0;
  const uint32_t* plan_words = d_total + kSortPlanScalar;
  TIMED("emit_super", launch_emit_super(ctx->sorted_slots, ctx->offsets.u(), ctx->soffsets.u(), d_total + 1,
                                        reinterpret_cast<const uint2*>(ctx->rects.u()), ctx->sorted_prects,
                                        ctx->views_dev(), n, M, d_total + kSortedScalar, s_cap, ctx->skeys.u(),
                                        ctx->svals.u(), ctx->skeys_alt.u(), ctx->svals_alt.u(),
                                        st_by_view ? plan_words + kPlanSuperPasses : nullptr, pair_base,
                                        ctx->stream));
  bool salt = false;
  if (st_by_view) {
    const SortPlan splan{plan_words + kPlanSuperPasses, plan_words + kPlanSuperActive, nullptr, nullptr, 1u, 8,
                         nullptr, nullptr, nullptr};
    TIMED("super_sort", launch_radix_sort(ctx->skeys.u(), ctx->svals.u(), ctx->skeys_alt.u(), ctx->svals_alt.u(),
                                          s_cap, 0, 8 + ctx->st_view_bits, ctx->sort_ws, &salt, ctx->stream,
                                          d_total + 1, nullptr, nullptr, false, nullptr, &splan));
  } else {
    const int st_bits = std::max(1, bit_width((unsigned long long)std::max(0, supers - 1)));
    TIMED("super_sort", launch_radix_sort(ctx->skeys.u(), ctx->svals.u(), ctx->skeys_alt.u(), ctx->svals_alt.u(),
                                          s_cap, 0, st_bits, ctx->sort_ws, &salt, ctx->stream, d_total + 1));
  }
  const uint32_t* skeys = salt ? ctx->skeys_alt.u() : ctx->skeys.u();
  const uint32_t* svals = salt ? ctx->svals_alt.u() : ctx->svals.u();
  TIMED("segments", launch_super_ranges(skeys, s_cap, d_total + 1, ctx->sranges.u(), supers, ctx->stream));
  TIMED("segments", launch_segment_counts(ctx->sranges.u(), supers, ctx->nseg.u(), ctx->stream));
  TIMED("segments", launch_exclusive_scan(ctx->nseg.u(), nullptr, ctx->seg_base.u(), supers, d_total + 2,
                                          ctx->sort_ws.scan, ctx->stream));
  TIMED("segments", launch_segment_list(ctx->nseg.u(), ctx->seg_base.u(), supers, ctx->seg_st.u(), ctx->stream));
  TIMED("tile_count", launch_tile_count(skeys, ctx->sranges.u(), ctx->nseg.u(), ctx->seg_base.u(), ctx->seg_st.u(),
                                        d_total + 2, ns_max, ctx->bin_table.u(), ctx->stream));
  TIMED("tile_scan", launch_exclusive_scan(ctx->bin_table.u(), nullptr, ctx->bin_table.u(), 64LL * ns_max + 1,
                                           nullptr, ctx->sort_ws.scan, ctx->stream));
  TIMED("tile_write", launch_tile_write(skeys, svals, ctx->sranges.u(), ctx->nseg.u(), ctx->seg_base.u(),
                                        ctx->seg_st.u(), d_total + 2, ns_max, ctx->bin_table.u(), p_cap,
                                        ctx->pair_vals.u(), ctx->stream));
  TIMED("tile_ranges", launch_tile_ranges_table(ctx->views_dev(), n_views, tiles, ctx->bin_table.u(), ctx->nseg.u(),
                                                ctx->seg_base.u(), p_cap, ctx->ranges.u(), ctx->stream));
  ctx->sorted_pair_vals = ctx->pair_vals.u();
  if (ctx->rcfg.fwd_queue || ctx->rcfg.bwd_queue || (ctx->autotune && !ctx->capturing))
    if (splat_status s = enqueue_tile_order(ctx)) return s;
  TIMED("forward", launch_forward(ctx->raster, ctx->rcfg, ctx->views_dev(), ctx->ranges.u(), ctx->sorted_pair_vals,
                                  ctx->recs(), ctx->rgb.f(), ctx->T.f(), ctx->last.i(), ctx->tile_order,
                                  ctx->scalars.u() + 4, ctx->stream));
  return SPLAT_OK;
}

// Tiles by list length, longest first: the work order of the queue launches.
splat_status enqueue_tile_order(splat_context* ctx) {
  const int tiles = ctx->total_tiles;
  CK(launch_tile_order_keys(ctx->ranges.u(), tiles, ctx->order_keys.u(), ctx->order_vals.u(), ctx->stream),
     "tile order");
  bool alt = false;
  CK(launch_radix_sort(ctx->order_keys.u(), ctx->order_vals.u(), ctx->order_keys_alt.u(), ctx->order_vals_alt.u(),
                       tiles, 0, 16, ctx->sort_ws, &alt, ctx->stream),
     "tile order");
  ctx->tile_order = alt ? ctx->order_vals_alt.u() : ctx->order_vals.u();
  return SPLAT_OK;
}

// Mean time of `reps` launches of `launch` on the context's stream, in ms.
template <typename F>
splat_status time_launches(splat_context* ctx, int reps, F launch, float* ms) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a), "event");
  CK(hipEventCreate(&b), "event");
  hipError_t e = launch();  // warm-up
  if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch();
  if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  CK(e, "autotune");
  *ms = t / reps;
  return SPLAT_OK;
}

// The forward launch shape for this batch shape: every variant timed on the
// batch just binned, the fastest kept (all give the same image; the last one
// timed is re-run with the choice so the outputs are the chosen kernel's).
splat_status tune_forward(splat_context* ctx) {
  RasterConfig best = ctx->rcfg;
  float best_ms = 1e30f;
  for (int shape = 0; shape < 3; ++shape)
    for (int queue = 0; queue < 2; ++queue) {
      RasterConfig c = ctx->rcfg;
      c.fwd_shape = shape;
      c.fwd_queue = queue != 0;
      float ms = 0.f;
      if (splat_status s = time_launches(ctx, 4, [&] {
            return launch_forward(ctx->raster, c, ctx->views_dev(), ctx->ranges.u(), ctx->sorted_pair_vals,
                                  ctx->recs(), ctx->rgb.f(), ctx->T.f(), ctx->last.i(), ctx->tile_order,
                                  ctx->scalars.u() + 4, ctx->stream);
          }, &ms))
        return s;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  ctx->rcfg.fwd_shape = best.fwd_shape;
  ctx->rcfg.fwd_queue = best.fwd_queue;
  CK(launch_forward(ctx->raster, ctx->rcfg, ctx->views_dev(), ctx->ranges.u(), ctx->sorted_pair_vals, ctx->recs(),
                    ctx->rgb.f(), ctx->T.f(), ctx->last.i(), ctx->tile_order, ctx->scalars.u() + 4, ctx->stream),
     "forward");
  ctx->tuning_backward = true;
  return SPLAT_OK;
}

// The forward of one batch, its pair buffers sized exactly (one host read of
// the pair counts between projection and binning).
splat_status forward_impl(splat_context* ctx, int n_views, const int32_t* view_ids, const splat_raster_options* o) {
  if (splat_status s = batch_setup(ctx, n_views, view_ids, o)) return s;
  if (splat_status s = upload_batch(ctx)) return s;
  if (splat_status s = enqueue_front(ctx)) return s;
  uint32_t totals[2] = {0, 0};
  if (ctx->M > 0) {
    CK(hipMemcpyAsync(totals, ctx->scalars.u(), sizeof(totals), hipMemcpyDeviceToHost, ctx->stream),
       "read pair count");
    CK(hipStreamSynchronize(ctx->stream), "sync");
  }
  ctx->P = totals[0];
  ctx->S = totals[1];
  if (splat_status s = reserve_pairs(ctx, totals[0], totals[1])) return s;
  if (splat_status s = enqueue_back(ctx)) return s;
  ctx->tuning_backward = false;
  if (ctx->autotune && !ctx->forced && !ctx->tuned.count(ctx->shape_key))
    if (splat_status s = tune_forward(ctx)) return s;
  ctx->forward_done = true;
  return SPLAT_OK;
}

splat_status reserve_loss(splat_context* ctx, float w) {
  const long long px = ctx->total_px;
  const int V = (int)ctx->batch_ids.size();
  CK(ctx->dL.reserve(sizeof(float) * 3 * std::max(1LL, px)), "alloc dL");
  CK(ctx->block_sums.reserve(sizeof(float) * 2 * std::max(1, ctx->loss_tiles)), "alloc loss");
  CK(ctx->loss_sums.reserve(sizeof(float) * 2 * V), "alloc loss");
  if (w > 0.f) CK(ctx->terms.reserve(sizeof(float) * kSsimPlanes * std::max(1LL, px)), "alloc ssim terms");
  return SPLAT_OK;
}

// Targets uploaded after the batch was set up: refresh the views' target
// offsets. Returns whether the descriptors changed.
bool refresh_targets(splat_context* ctx) {
  bool stale = false;
  for (size_t k = 0; k < ctx->batch_desc.size(); ++k) {
    const long long want = ctx->target_offset[ctx->batch_ids[k]];
    if (ctx->batch_desc[k].target_base != want) {
      ctx->batch_desc[k].target_base = want;
      stale = true;
    }
  }
  return stale;
}

splat_status enqueue_loss(splat_context* ctx, float w) {
  LossParams lp{};
  for (int i = 0; i < 11; ++i) lp.kernel[i] = ctx->ssim_kernel[i];
  lp.ssim_weight = w;
  lp.n_views = (int)ctx->batch_ids.size();
  lp.total_tiles = ctx->loss_tiles;
  TIMED("loss", launch_loss(lp, ctx->views_dev(), ctx->rgb.f(), ctx->targets.f(), ctx->dL.f(), ctx->terms.f(),
                            ctx->total_px, ctx->block_sums.f(), ctx->stream));
  CK(launch_loss_reduce(ctx->views_dev(), lp.n_views, ctx->block_sums.f(), ctx->loss_sums.f(), ctx->stream),
     "loss reduce");
  return SPLAT_OK;
}

void losses_from_sums(const splat_context* ctx, const float* sums, float w, float* losses) {
  for (size_t k = 0; k < ctx->batch_desc.size(); ++k) {
    const ViewDesc& d = ctx->batch_desc[k];
    const float inverse = 1.f / static_cast<float>((long long)d.width * d.height * 3);
    const float l1 = sums[2 * k] * inverse;
    losses[k] = w <= 0.f ? l1 : (1.f - w) * l1 + w * (1.f - sums[2 * k + 1] * inverse);
  }
}

splat_status loss_impl(splat_context* ctx, float w, float* losses) {
  if (!ctx->forward_done) return fail(ctx, SPLAT_ERROR_STATE, "loss: no forward batch");
  for (int id : ctx->batch_ids)
    if (!ctx->has_target[id]) return fail(ctx, SPLAT_ERROR_STATE, "loss: a batch view has no target image");
  if (splat_status s = reserve_loss(ctx, w)) return s;
  if (refresh_targets(ctx))
    if (splat_status s = upload_batch(ctx)) return s;
  if (splat_status s = enqueue_loss(ctx, w)) return s;
  ctx->loss_done = true;
  ctx->last_ssim_weight = w;
  if (losses) {
    const int V = (int)ctx->batch_ids.size();
    std::vector<float> sums(2 * V);
    CK(hipMemcpyAsync(sums.data(), ctx->loss_sums.f(), sizeof(float) * 2 * V, hipMemcpyDeviceToHost, ctx->stream),
       "read loss");
    CK(hipStreamSynchronize(ctx->stream), "sync");
    losses_from_sums(ctx, sums.data(), w, losses);
  }
  return SPLAT_OK;
}

// With fuse_adam (a training step), the per-gaussian adjoint also applies
// the Adam step staged in d_adam, unless *adam_skip is set.
splat_status enqueue_backward(splat_context* ctx, bool fuse_adam = false, const uint32_t* adam_skip = nullptr) {
  const long long M = ctx->M;
  const bool det = ctx->opts.deterministic != 0;
  CK(ctx->screen.reserve(sizeof(float) * kScreenStride * std::max(1LL, M)), "alloc screen");
  uint32_t* epoch = ctx->scalars.u() + kEpochScalar;
  if (det) {
    if (splat_status st = reserve_partials(ctx, ctx->P_cap)) return st;
    CK(ctx->carry.reserve(sizeof(float) * det_reduce_carry_floats(std::max<unsigned long long>(1, ctx->P_cap))),
       "alloc carries");
    TIMED("epoch", launch_bump_epoch(epoch, ctx->stream));
  } else {
    TIMED("screen_clear", launch_fill_f32(ctx->screen.f(), 0.f, kScreenStride * M, ctx->stream));
  }
  const uint32_t p_cap = (uint32_t)ctx->P_cap;
  auto launch = [&](const RasterConfig& c) {
    return launch_backward(ctx->raster, c, ctx->views_dev(), ctx->ranges.u(), ctx->sorted_pair_vals, ctx->recs(),
                           ctx->dL.f(), ctx->T.f(), ctx->last.i(), ctx->screen.f(), ctx->pair_base.u(),
                           ctx->partial.f(), static_cast<uint8_t*>(ctx->visit.ptr), p_cap, det, epoch,
                           ctx->tile_order, ctx->scalars.u() + 4, ctx->stream);
  };
  if (ctx->tuning_backward && !ctx->capturing) {
    // the adjoint's launch shape for this batch shape: each variant timed on
    // this batch (the screen accumulators cleared again before the real run)
    RasterConfig best = ctx->rcfg;
    float best_ms = 1e30f;
    for (int batch = 64; batch <= 128; batch *= 2)
      for (int queue = 0; queue < 2; ++queue) {
        RasterConfig c = ctx->rcfg;
        c.bwd_batch = batch;
        c.bwd_queue = queue != 0;
        float ms = 0.f;
        if (splat_status s = time_launches(ctx, 4, [&] { return launch(c); }, &ms)) return s;
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    ctx->rcfg.bwd_batch = best.bwd_batch;
    ctx->rcfg.bwd_queue = best.bwd_queue;
    ctx->tuned[ctx->shape_key] = ctx->rcfg;
    ctx->tuning_backward = false;
    if (!det) CK(launch_fill_f32(ctx->screen.f(), 0.f, kScreenStride * M, ctx->stream), "clear");
  }
  TIMED("backward", launch(ctx->rcfg));
  AdjointParams ap{};
  ap.cap = ctx->cap;
  ap.n = ctx->n;
  ap.rest_stride = ctx->rs;
  ap.sh_use = ctx->sh_use;
  ap.n_views = (int)ctx->batch_ids.size();
  ap.deterministic = det ? 1 : 0;
  if (det)
    TIMED("det_reduce", launch_det_reduce(ctx->sorted_slots, ctx->offsets.u(), ctx->scalars.u(), M,
                                          ctx->scalars.u() + kSortedScalar, p_cap, ctx->partial.f(),
                                          static_cast<const uint8_t*>(ctx->visit.ptr), epoch, ctx->screen.f(),
                                          ctx->carry.f(), ctx->stream));
  if (fuse_adam) {
    AdamFuse fuse;
    fuse.step = static_cast<const AdamStep*>(ctx->d_adam.ptr);
    fuse.skip = adam_skip;
    fuse.params = ctx->params.f();
    fuse.m = ctx->m.f();
    fuse.v = ctx->v.f();
    fuse.stat_sg = ctx->stat_sg.f();
    fuse.stat_vc = ctx->stat_vc.i();
    fuse.stored_clear = ctx->grads_clear;
    TIMED("adjoint_adam", launch_adjoint(ap, ctx->views_dev(), ctx->params.f(), ctx->recs(), ctx->screen.f(),
                                         ctx->grads.f(), ctx->sg.f(), ctx->vc.i(), ctx->stream, &fuse));
    // the update leaves them at zero (a step the device skips on overflow is
    // cleared by the host before its redo)
    ctx->grads_clear = true;
  } else {
    TIMED("adjoint", launch_adjoint(ap, ctx->views_dev(), ctx->params.f(), ctx->recs(), ctx->screen.f(),
                                    ctx->grads.f(), ctx->sg.f(), ctx->vc.i(), ctx->stream));
    ctx->grads_clear = false;
  }
  return SPLAT_OK;
}

splat_status backward_impl(splat_context* ctx) {
  if (!ctx->forward_done) return fail(ctx, SPLAT_ERROR_STATE, "render_backward: no forward batch");
  if (!ctx->loss_done) return fail(ctx, SPLAT_ERROR_STATE, "render_backward: no image gradient (run splat_loss)");
  return enqueue_backward(ctx);
}

// The step's Adam scalars to device memory (ahead of the update on the stream).
splat_status stage_adam(splat_context* ctx, const splat_adam_params* a) {
  AdamStep& p = ctx->adam_host;
  p.lr[0] = a->lr_position;
  p.lr[1] = a->lr_log_scale;
  p.lr[2] = a->lr_rotation;
  p.lr[3] = a->lr_opacity;
  p.lr[4] = a->lr_dc;
  p.lr[5] = a->lr_rest;
  p.beta1 = a->beta1;
  p.beta2 = a->beta2;
  p.eps = a->epsilon;
  p.c1 = a->correction1;
  p.c2 = a->correction2;
  CK(ctx->d_adam.reserve(sizeof(AdamStep)), "alloc adam");
  CK(hipMemcpyAsync(ctx->d_adam.ptr, &ctx->adam_host, sizeof(AdamStep), hipMemcpyHostToDevice, ctx->stream),
     "upload adam");
  return SPLAT_OK;
}

splat_status enqueue_adam(splat_context* ctx, const uint32_t* skip) {
  AdamShape sh{};
  sh.cap = ctx->cap;
  sh.n = ctx->n;
  sh.rest_stride = ctx->rs;
  TIMED("adam", launch_adam(sh, static_cast<const AdamStep*>(ctx->d_adam.ptr), skip, ctx->params.f(), ctx->grads.f(),
                            ctx->m.f(), ctx->v.f(), ctx->sg.f(), ctx->vc.i(), ctx->stat_sg.f(), ctx->stat_vc.i(),
                            ctx->stream));
  return SPLAT_OK;
}

splat_status adam_impl(splat_context* ctx, const splat_adam_params* a) {
  if (splat_status s = stage_adam(ctx, a)) return s;
  if (splat_status s = enqueue_adam(ctx, nullptr)) return s;
  ctx->grads_clear = true;  // Adam clears the gradients and folds sg / vc
  return SPLAT_OK;
}

// One training step as a plain launch sequence (one host read of the pair
// counts inside the forward).
splat_status train_step_plain(splat_context* ctx, int n_views, const int32_t* view_ids,
                              const splat_raster_options* o, float w, const splat_adam_params* adam, float* losses) {
  if (splat_status s = forward_impl(ctx, n_views, view_ids, o)) return s;
  if (splat_status s = loss_impl(ctx, w, nullptr)) return s;
  if (splat_status s = stage_adam(ctx, adam)) return s;
  if (splat_status s = enqueue_backward(ctx, true)) return s;  // the adjoint applies Adam
  if (losses) {
    std::vector<float> sums(2 * n_views);
    CK(hipMemcpyAsync(sums.data(), ctx->loss_sums.f(), sizeof(float) * 2 * n_views, hipMemcpyDeviceToHost,
                      ctx->stream),
       "read loss");
    CK(hipStreamSynchronize(ctx->stream), "sync");
    losses_from_sums(ctx, sums.data(), w, losses);
  }
  return SPLAT_OK;
}

// Everything a captured step bakes in: shapes, sizes, options, capacities.
std::string graph_key(const splat_context* ctx, float w) {
  std::string k;
  auto put = [&k](const void* p, size_t n) { k.append(static_cast<const char*>(p), n); };
  const long long head[] = {ctx->cap, ctx->n, ctx->rs, ctx->sh_use, (long long)ctx->P_cap, (long long)ctx->S_cap};
  put(head, sizeof head);
  put(&ctx->opts, sizeof ctx->opts);
  put(&w, sizeof w);
  const char clear = ctx->grads_clear ? 1 : 0;  // which fused update the graph holds
  put(&clear, 1);
  const int cfg[4] = {ctx->rcfg.fwd_shape, ctx->rcfg.fwd_queue ? 1 : 0, ctx->rcfg.bwd_batch, ctx->rcfg.bwd_queue ? 1 : 0};
  put(cfg, sizeof cfg);
  for (const ViewDesc& d : ctx->batch_desc) {
    const int wh[2] = {d.width, d.height};
    put(wh, sizeof wh);
  }
  return k;
}

constexpr size_t kMaxGraphs = 16;

// One training step replayed from a captured graph: no host synchronisation
// inside the step, one read of the losses and counters after it. Pair buffers
// run at capacities learnt from earlier steps; a step whose pairs outgrow them
// skips its update on the device, and is redone as a plain step after the
// capacities grow.
splat_status train_step_graph(splat_context* ctx, int n_views, const int32_t* view_ids,
                              const splat_raster_options* o, float w, const splat_adam_params* adam, float* losses) {
  if (ctx->graph_p_cap == 0) {
    if (splat_status s = train_step_plain(ctx, n_views, view_ids, o, w, adam, losses)) return s;
    ctx->graph_p_cap = ctx->P + ctx->P / 4 + 65536;
    ctx->graph_s_cap = ctx->S + ctx->S / 4 + 4096;
    return SPLAT_OK;
  }
  if (splat_status s = batch_setup(ctx, n_views, view_ids, o)) return s;
  for (int id : ctx->batch_ids)
    if (!ctx->has_target[id]) return fail(ctx, SPLAT_ERROR_STATE, "train_step: a batch view has no target image");
  if (ctx->autotune && !ctx->forced && !ctx->tuned.count(ctx->shape_key))
    return train_step_plain(ctx, n_views, view_ids, o, w, adam, losses);  // tunes this batch shape
  if (splat_status s = reserve_pairs(ctx, ctx->graph_p_cap, ctx->graph_s_cap)) return s;
  if (splat_status s = reserve_loss(ctx, w)) return s;
  CK(ctx->d_adam.reserve(sizeof(AdamStep)), "alloc adam");
  const size_t rb_bytes = sizeof(float) * 2 * n_views + 16;
  if (ctx->readback_bytes < rb_bytes) {
    if (ctx->readback) (void)hipHostFree(ctx->readback);
    ctx->readback = nullptr;
    ctx->readback_bytes = 0;
    CK(hipHostMalloc(reinterpret_cast<void**>(&ctx->readback), sizeof(float) * 2 * kMaxBatchViews + 16),
       "alloc pinned");
    ctx->readback_bytes = sizeof(float) * 2 * kMaxBatchViews + 16;
  }
  if (splat_status s = upload_batch(ctx)) return s;
  if (splat_status s = stage_adam(ctx, adam)) return s;
  const std::string key = graph_key(ctx, w);
  const unsigned long long gen = alloc_generation();
  splat_context::Graph* g = nullptr;
  for (auto& e : ctx->graphs)
    if (e.key == key && e.generation == gen) g = &e;
  if (!g) {
    for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();)
      if (it->generation != gen) {
        (void)hipGraphExecDestroy(it->exec);
        it = ctx->graphs.erase(it);
      } else {
        ++it;
      }
    if (ctx->graphs.size() >= kMaxGraphs) ctx->drop_graphs();
    splat_context::Graph fresh;
    fresh.key = key;
    fresh.generation = gen;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "begin capture");
    ctx->capturing = true;
    splat_status s = enqueue_front(ctx);
    if (!s) s = enqueue_back(ctx);
    if (!s) s = enqueue_loss(ctx, w);
    if (!s) s = enqueue_backward(ctx, true, ctx->scalars.u() + 3);  // Adam fused; skipped on overflow
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (s) {
      if (graph) (void)hipGraphDestroy(graph);
      return s;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "end capture");
    if (alloc_generation() != gen) {
      (void)hipGraphDestroy(graph);
      return fail(ctx, SPLAT_ERROR_DEVICE, "train_step: a buffer moved during capture");
    }
    const hipError_t ie = hipGraphInstantiate(&fresh.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return hip_fail(ctx, ie, "instantiate graph");
    fresh.sorted_slots = ctx->sorted_slots;
    fresh.raster = ctx->raster;
    ctx->graphs.push_back(fresh);
    g = &ctx->graphs.back();
  }
  CK(hipGraphLaunch(g->exec, ctx->stream), "launch graph");
  ctx->sorted_slots = g->sorted_slots;
  ctx->sorted_pair_vals = ctx->pair_vals.u();
  ctx->raster = g->raster;
  float* sums = reinterpret_cast<float*>(ctx->readback);
  uint32_t* counters = ctx->readback + 2 * n_views;
  CK(hipMemcpyAsync(sums, ctx->loss_sums.f(), sizeof(float) * 2 * n_views, hipMemcpyDeviceToHost, ctx->stream),
     "read loss");
  CK(hipMemcpyAsync(counters, ctx->scalars.u(), 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
     "read counters");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->P = counters[0];
  ctx->S = counters[1];
  if (counters[3]) {
    // outgrew the capacities: the update was skipped on the device; clear this
    // step's partial gradients and statistics and redo it with room to spare
    ctx->graph_p_cap = ctx->P + ctx->P / 4 + 65536;
    ctx->graph_s_cap = ctx->S + ctx->S / 4 + 4096;
    CK(launch_fill_f32(ctx->grads.f(), 0.f, ctx->cap * (14 + ctx->rs), ctx->stream), "clear");
    CK(launch_fill_f32(ctx->sg.f(), 0.f, ctx->cap, ctx->stream), "clear");
    CK(launch_fill_u32(ctx->vc.u(), 0u, ctx->cap, ctx->stream), "clear");
    ctx->grads_clear = true;
    return train_step_plain(ctx, n_views, view_ids, o, w, adam, losses);
  }
  ctx->forward_done = true;
  ctx->loss_done = true;
  ctx->last_ssim_weight = w;
  if (losses) losses_from_sums(ctx, sums, w, losses);
  return SPLAT_OK;
}

}  // namespace

hipError_t splat_context::ensure_sort_workspace(long long max_n) {
  if (max_n <= sort_ws.scan.max_n && sort_hist_elems(max_n) <= sort_ws.hist_capacity) return hipSuccess;
  const long long hist_elems = sort_hist_elems(max_n);
  // the scan workspace must cover both the histogram table and max_n items
  const long long scan_n = std::max(max_n, hist_elems);
  SPLAT_HIP_CHECK(hist_ws.reserve(sizeof(uint32_t) * hist_elems));
  SPLAT_HIP_CHECK(scan_ws.reserve(sizeof(uint32_t) * scan_workspace_elems(scan_n)));
  sort_ws.hist = hist_ws.u();
  sort_ws.hist_capacity = (long long)(hist_ws.bytes / sizeof(uint32_t));
  sort_ws.scan.buf = scan_ws.u();
  sort_ws.scan.capacity = (long long)(scan_ws.bytes / sizeof(uint32_t));
  sort_ws.scan.max_n = scan_n;
  return hipSuccess;
}

// ================================================================ C ABI ==
extern "C" {

void splat_default_options(splat_raster_options* o) {
  if (!o) return;
  o->background[0] = o->background[1] = o->background[2] = 0.f;
  o->sh_degree = 0;
  o->min_alpha = 1.f / 255.f;
  o->transmittance_cutoff = 1e-4f;
  o->screen_dilation = 0.3f;
  o->deterministic = 0;
}

splat_status splat_availability(char* name, size_t name_len, char* reason, size_t reason_len) {
  auto put = [](char* dst, size_t len, const std::string& s) {
    if (dst && len) {
      std::snprintf(dst, len, "%s", s.c_str());
    }
  };
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    put(reason, reason_len, e != hipSuccess ? std::string("no HIP device: ") + hipGetErrorString(e)
                                             : std::string("no HIP device visible"));
    return SPLAT_ERROR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) {
    put(reason, reason_len, std::string("device query failed: ") + hipGetErrorString(e));
    return SPLAT_ERROR_NO_DEVICE;
  }
  put(name, name_len, std::string(prop.name) + " (" + prop.gcnArchName + ")");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    put(reason, reason_len, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
    return SPLAT_ERROR_NO_DEVICE;
  }
  put(reason, reason_len, "");
  return SPLAT_OK;
}

splat_status splat_create(int32_t device, splat_context** out) {
  if (!out) return fail(nullptr, SPLAT_ERROR_INVALID_ARGUMENT, "splat_create: null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(nullptr, SPLAT_ERROR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count) return fail(nullptr, SPLAT_ERROR_INVALID_ARGUMENT, "device index out of range");
  splat_context* ctx = new splat_context;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    ctx->rcfg.cus = cus > 0 ? cus : 256;
  }
  // the fixed scratch of splat_loss_noalloc, looked up with this device current
  if (e == hipSuccess) e = noalloc_attach(device);
  if (e != hipSuccess) {
    std::string msg = std::string("context set-up failed: ") + hipGetErrorString(e);
    delete ctx;
    return fail(nullptr, SPLAT_ERROR_DEVICE, msg);
  }
  // ssimKernel (SplatRaster.cpp:587-600), host libm as the CPU trainer
  float total = 0.f;
  for (int i = -5; i <= 5; ++i) {
    const float v = std::exp(-0.5f * static_cast<float>(i * i) / (1.5f * 1.5f));
    ctx->ssim_kernel[i + 5] = v;
    total += v;
  }
  for (float& v : ctx->ssim_kernel) v /= total;
  *out = ctx;
  return SPLAT_OK;
}

void splat_destroy(splat_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

const char* splat_last_error(const splat_context* ctx) {
  return ctx ? ctx->error.c_str() : g_no_context_error.c_str();
}

splat_status splat_set_capacity(splat_context* ctx, int32_t max_gaussians, int32_t sh_degree) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (max_gaussians <= 0 || sh_degree < 0 || sh_degree > 3)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "set_capacity: max_gaussians > 0 and sh_degree in [0, 3]");
  CK(hipSetDevice(ctx->device), "set device");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  const int per = sh_degree == 0 ? 0 : (sh_degree == 1 ? 3 : (sh_degree == 2 ? 8 : 15));
  ctx->rs = 3 * per;
  ctx->sh_degree = sh_degree;
  ctx->cap = max_gaussians;
  ctx->n = 0;
  const size_t floats = (size_t)max_gaussians * (14 + ctx->rs);
  for (DeviceBuffer* b : {&ctx->params, &ctx->grads, &ctx->m, &ctx->v}) CK(b->reserve(sizeof(float) * floats), "alloc parameters");
  CK(ctx->sg.reserve(sizeof(float) * max_gaussians), "alloc stats");
  CK(ctx->vc.reserve(sizeof(int) * max_gaussians), "alloc stats");
  CK(ctx->stat_sg.reserve(sizeof(float) * max_gaussians), "alloc stats");
  CK(ctx->stat_vc.reserve(sizeof(int) * max_gaussians), "alloc stats");
  CK(launch_fill_f32(ctx->params.f(), 0.f, (long long)floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->grads.f(), 0.f, (long long)floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->m.f(), 0.f, (long long)floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->v.f(), 0.f, (long long)floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->sg.f(), 0.f, max_gaussians, ctx->stream), "clear");
  CK(launch_fill_u32(ctx->vc.u(), 0u, max_gaussians, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->stat_sg.f(), 0.f, max_gaussians, ctx->stream), "clear");
  CK(launch_fill_u32(ctx->stat_vc.u(), 0u, max_gaussians, ctx->stream), "clear");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->forward_done = ctx->loss_done = false;
  ctx->grads_clear = true;
  return SPLAT_OK;
}

int32_t splat_count(const splat_context* ctx) { return ctx ? ctx->n : 0; }

splat_status splat_upload_cloud(splat_context* ctx, const splat_cloud* c) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!c || c->n < 0 || c->n > ctx->cap)
    return fail(ctx, SPLAT_ERROR_CAPACITY, "upload_cloud: count exceeds the reserved capacity");
  if (c->rest_stride != ctx->rs)
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "upload_cloud: rest_stride does not match the capacity's SH degree");
  if (c->n > 0 && (!c->pos || !c->log_scale || !c->rot || !c->opacity || !c->dc || (ctx->rs > 0 && !c->rest)))
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "upload_cloud: null parameter array");
  CK(hipSetDevice(ctx->device), "set device");
  const long long cap = ctx->cap, n = c->n;
  float* P = ctx->params.f();
  auto up = [&](long long off, const float* src, long long count) -> hipError_t {
    if (count <= 0) return hipSuccess;
    return hipMemcpyAsync(P + off, src, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream);
  };
  CK(up(0, c->pos, 3 * n), "upload");
  CK(up(3 * cap, c->log_scale, 3 * n), "upload");
  CK(up(6 * cap, c->rot, 4 * n), "upload");
  CK(up(10 * cap, c->opacity, n), "upload");
  CK(up(11 * cap, c->dc, 3 * n), "upload");
  CK(up(14 * cap, c->rest, (long long)ctx->rs * n), "upload");
  const long long floats = cap * (14 + ctx->rs);
  CK(launch_fill_f32(ctx->grads.f(), 0.f, floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->m.f(), 0.f, floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->v.f(), 0.f, floats, ctx->stream), "clear");
  CK(launch_fill_f32(ctx->sg.f(), 0.f, cap, ctx->stream), "clear");
  CK(launch_fill_u32(ctx->vc.u(), 0u, cap, ctx->stream), "clear");
  ctx->grads_clear = true;
  CK(launch_fill_f32(ctx->stat_sg.f(), 0.f, cap, ctx->stream), "clear");
  CK(launch_fill_u32(ctx->stat_vc.u(), 0u, cap, ctx->stream), "clear");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->n = (int)n;
  ctx->cloud_degree = c->sh_degree;
  ctx->forward_done = ctx->loss_done = false;
  return SPLAT_OK;
}

splat_status splat_download_cloud(splat_context* ctx, splat_cloud* out) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!out) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "download_cloud: null out");
  CK(hipSetDevice(ctx->device), "set device");
  const long long cap = ctx->cap, n = ctx->n;
  const float* P = ctx->params.f();
  auto down = [&](float* dst, long long off, long long count) -> hipError_t {
    if (!dst || count <= 0) return hipSuccess;
    return hipMemcpyAsync(dst, P + off, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream);
  };
  CK(down(out->pos, 0, 3 * n), "download");
  CK(down(out->log_scale, 3 * cap, 3 * n), "download");
  CK(down(out->rot, 6 * cap, 4 * n), "download");
  CK(down(out->opacity, 10 * cap, n), "download");
  CK(down(out->dc, 11 * cap, 3 * n), "download");
  CK(down(out->rest, 14 * cap, (long long)ctx->rs * n), "download");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  out->n = (int32_t)n;
  out->rest_stride = ctx->rs;
  out->sh_degree = ctx->cloud_degree;
  return SPLAT_OK;
}

splat_status splat_upload_views(splat_context* ctx, int32_t n, const splat_view* views, const float* const* images) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (n < 0 || (n > 0 && !views)) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "upload_views: bad arguments");
  for (int k = 0; k < n; ++k) {
    const splat_view& v = views[k];
    if (v.width <= 0 || v.height <= 0 || !(v.fx > 0.0) || !(v.fy > 0.0))
      return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "upload_views: a view has invalid intrinsics");
  }
  CK(hipSetDevice(ctx->device), "set device");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->views.assign(views, views + n);
  ctx->target_offset.assign(n, 0);
  ctx->has_target.assign(n, false);
  long long total = 0;
  for (int k = 0; k < n; ++k) {
    ctx->target_offset[k] = total;
    total += 3LL * views[k].width * views[k].height;
  }
  CK(ctx->targets.reserve(sizeof(float) * std::max(1LL, total)), "alloc targets");
  if (images) {
    for (int k = 0; k < n; ++k) {
      if (!images[k]) continue;
      CK(hipMemcpyAsync(ctx->targets.f() + ctx->target_offset[k], images[k],
                        sizeof(float) * 3 * (size_t)views[k].width * views[k].height, hipMemcpyHostToDevice,
                        ctx->stream),
         "upload targets");
      ctx->has_target[k] = true;
    }
  }
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->forward_done = ctx->loss_done = false;
  return SPLAT_OK;
}

splat_status splat_render_forward(splat_context* ctx, int32_t n_views, const int32_t* view_ids,
                                  const splat_raster_options* options) {
  if (splat_status s = check_ready(ctx)) return s;
  CK(hipSetDevice(ctx->device), "set device");
  return forward_impl(ctx, n_views, view_ids, options);
}

splat_status splat_download_image(splat_context* ctx, int32_t slot, float* rgb, float* transmittance,
                                  int32_t* last_contributor) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!ctx->forward_done || slot < 0 || slot >= (int)ctx->batch_desc.size())
    return fail(ctx, SPLAT_ERROR_STATE, "download_image: no such batch slot");
  const ViewDesc& d = ctx->batch_desc[slot];
  const size_t px = (size_t)d.width * d.height;
  if (rgb) CK(hipMemcpyAsync(rgb, ctx->rgb.f() + 3 * d.pixel_base, 12 * px, hipMemcpyDeviceToHost, ctx->stream), "download");
  if (transmittance)
    CK(hipMemcpyAsync(transmittance, ctx->T.f() + d.pixel_base, 4 * px, hipMemcpyDeviceToHost, ctx->stream), "download");
  if (last_contributor)
    CK(hipMemcpyAsync(last_contributor, ctx->last.i() + d.pixel_base, 4 * px, hipMemcpyDeviceToHost, ctx->stream),
       "download");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_download_projection(splat_context* ctx, int32_t slot, int32_t* count, int32_t* index, float* mean2,
                                       float* conic3, float* depth, float* radius, float* color3, float* opacity,
                                       int32_t* clamped3, int32_t* rect4) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!ctx->forward_done || slot < 0 || slot >= (int)ctx->batch_desc.size())
    return fail(ctx, SPLAT_ERROR_STATE, "download_projection: no such batch slot");
  const long long n = ctx->n, M = ctx->M;
  std::vector<SplatRec> recs((size_t)n);
  std::vector<uint32_t> order((size_t)M);
  uint32_t sorted = 0;  // the depth-sorted (projected) slots of the batch
  if (n > 0) {
    CK(hipMemcpyAsync(recs.data(), ctx->recs() + (long long)slot * n, sizeof(SplatRec) * n, hipMemcpyDeviceToHost,
                      ctx->stream),
       "download");
    CK(hipMemcpyAsync(&sorted, ctx->scalars.u() + kSortedScalar, 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    CK(hipMemcpyAsync(order.data(), ctx->sorted_slots, 4 * M, hipMemcpyDeviceToHost, ctx->stream), "download");
  }
  CK(hipStreamSynchronize(ctx->stream), "sync");
  int k = 0;
  for (long long s = 0; s < std::min<long long>(M, sorted); ++s) {
    const long long sl = order[(size_t)s];
    if (sl / n != slot) continue;
    const int g = (int)(sl % n);
    const SplatRec& r = recs[(size_t)g];
    if (!(r.flags & kProjected)) continue;
    if (index) index[k] = g;
    if (mean2) { mean2[2 * k] = r.mx; mean2[2 * k + 1] = r.my; }
    if (conic3) { conic3[3 * k] = r.ca; conic3[3 * k + 1] = r.cb; conic3[3 * k + 2] = r.cc; }
    if (depth) depth[k] = r.depth;
    if (radius) radius[k] = r.radius;
    if (color3) { color3[3 * k] = r.cr; color3[3 * k + 1] = r.cg; color3[3 * k + 2] = r.cbl; }
    if (opacity) opacity[k] = r.opacity;
    if (clamped3) for (int c = 0; c < 3; ++c) clamped3[3 * k + c] = (r.flags >> c) & 1u;
    if (rect4) { rect4[4 * k] = r.x0; rect4[4 * k + 1] = r.x1; rect4[4 * k + 2] = r.y0; rect4[4 * k + 3] = r.y1; }
    ++k;
  }
  if (count) *count = k;
  return SPLAT_OK;
}

splat_status splat_download_tiles(splat_context* ctx, int32_t slot, int32_t* offsets, int32_t* entries,
                                  int32_t* entry_count) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!ctx->forward_done || slot < 0 || slot >= (int)ctx->batch_desc.size())
    return fail(ctx, SPLAT_ERROR_STATE, "download_tiles: no such batch slot");
  const ViewDesc& d = ctx->batch_desc[slot];
  const int tiles = d.cols * d.rows;
  std::vector<uint32_t> ranges(2 * (size_t)tiles);
  CK(hipMemcpyAsync(ranges.data(), ctx->ranges.u() + 2 * (long long)d.tile_base, 8 * (size_t)tiles,
                    hipMemcpyDeviceToHost, ctx->stream),
     "download");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  long long total = 0;
  for (int t = 0; t < tiles; ++t) total += ranges[2 * t + 1] - ranges[2 * t];
  if (entry_count) *entry_count = (int32_t)total;
  if (offsets) {
    long long run = 0;
    for (int t = 0; t < tiles; ++t) {
      offsets[t] = (int32_t)run;
      run += ranges[2 * t + 1] - ranges[2 * t];
    }
    offsets[tiles] = (int32_t)run;
  }
  if (entries && total > 0) {
    // the view's pairs are contiguous in the sorted order
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for (int t = 0; t < tiles; ++t)
      if (ranges[2 * t + 1] > ranges[2 * t]) {
        lo = std::min(lo, ranges[2 * t]);
        hi = std::max(hi, ranges[2 * t + 1]);
      }
    std::vector<uint32_t> vals(hi - lo);
    CK(hipMemcpyAsync(vals.data(), ctx->sorted_pair_vals + lo, 4 * (size_t)(hi - lo), hipMemcpyDeviceToHost,
                      ctx->stream),
       "download");
    CK(hipStreamSynchronize(ctx->stream), "sync");
    long long k = 0;
    for (int t = 0; t < tiles; ++t)
      for (uint32_t e = ranges[2 * t]; e < ranges[2 * t + 1]; ++e) entries[k++] = (int32_t)(vals[e - lo] % ctx->n);
  }
  return SPLAT_OK;
}

splat_status splat_loss(splat_context* ctx, float ssim_weight, float* losses) {
  if (splat_status s = check_ready(ctx)) return s;
  CK(hipSetDevice(ctx->device), "set device");
  return loss_impl(ctx, ssim_weight, losses);
}

splat_status splat_set_image_gradient(splat_context* ctx, int32_t slot, const float* dL_drgb) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!ctx->forward_done || slot < 0 || slot >= (int)ctx->batch_desc.size() || !dL_drgb)
    return fail(ctx, SPLAT_ERROR_STATE, "set_image_gradient: no such batch slot");
  const ViewDesc& d = ctx->batch_desc[slot];
  CK(ctx->dL.reserve(sizeof(float) * 3 * std::max(1LL, ctx->total_px)), "alloc dL");
  CK(hipMemcpyAsync(ctx->dL.f() + 3 * d.pixel_base, dL_drgb, 12 * (size_t)d.width * d.height, hipMemcpyHostToDevice,
                    ctx->stream),
     "upload dL");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->loss_done = true;
  return SPLAT_OK;
}

splat_status splat_download_image_gradient(splat_context* ctx, int32_t slot, float* dL_drgb) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!ctx->loss_done || slot < 0 || slot >= (int)ctx->batch_desc.size() || !dL_drgb)
    return fail(ctx, SPLAT_ERROR_STATE, "download_image_gradient: no image gradient for that slot");
  const ViewDesc& d = ctx->batch_desc[slot];
  CK(hipMemcpyAsync(dL_drgb, ctx->dL.f() + 3 * d.pixel_base, 12 * (size_t)d.width * d.height, hipMemcpyDeviceToHost,
                    ctx->stream),
     "download dL");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_render_backward(splat_context* ctx) {
  if (splat_status s = check_ready(ctx)) return s;
  CK(hipSetDevice(ctx->device), "set device");
  return backward_impl(ctx);
}

splat_status splat_download_gradients(splat_context* ctx, splat_gradients* out) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!out) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "download_gradients: null out");
  const long long cap = ctx->cap, n = ctx->n;
  const float* G = ctx->grads.f();
  auto down = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  };
  CK(down(out->pos, G, 12 * n), "download");
  CK(down(out->log_scale, G + 3 * cap, 12 * n), "download");
  CK(down(out->rot, G + 6 * cap, 16 * n), "download");
  CK(down(out->opacity, G + 10 * cap, 4 * n), "download");
  CK(down(out->dc, G + 11 * cap, 12 * n), "download");
  CK(down(out->rest, G + 14 * cap, 4 * (size_t)ctx->rs * n), "download");
  CK(down(out->screen_gradient, ctx->sg.f(), 4 * n), "download");
  CK(down(out->visible_count, ctx->vc.i(), 4 * n), "download");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_zero_gradients(splat_context* ctx) {
  if (splat_status s = check_ready(ctx)) return s;
  CK(launch_fill_f32(ctx->grads.f(), 0.f, ctx->cap * (14 + ctx->rs), ctx->stream), "clear");
  CK(launch_fill_f32(ctx->sg.f(), 0.f, ctx->cap, ctx->stream), "clear");
  CK(launch_fill_u32(ctx->vc.u(), 0u, ctx->cap, ctx->stream), "clear");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->grads_clear = true;
  return SPLAT_OK;
}

splat_status splat_adam_step(splat_context* ctx, const splat_adam_params* params) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!params) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "adam_step: null params");
  CK(hipSetDevice(ctx->device), "set device");
  if (splat_status s = adam_impl(ctx, params)) return s;
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_train_step(splat_context* ctx, int32_t n_views, const int32_t* view_ids,
                              const splat_raster_options* options, float ssim_weight, const splat_adam_params* adam,
                              float* losses) {
  if (splat_status s = check_ready(ctx)) return s;
  if (!adam) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "train_step: null adam params");
  if (!options) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "train_step: null options");
  CK(hipSetDevice(ctx->device), "set device");
  if (ctx->graph_mode) return train_step_graph(ctx, n_views, view_ids, options, ssim_weight, adam, losses);
  return train_step_plain(ctx, n_views, view_ids, options, ssim_weight, adam, losses);
}

splat_status splat_set_graph_mode(splat_context* ctx, int32_t enabled) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  ctx->graph_mode = enabled != 0;
  if (!ctx->graph_mode) ctx->drop_graphs();
  return SPLAT_OK;
}

splat_status splat_set_autotune(splat_context* ctx, int32_t enabled) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  ctx->autotune = enabled != 0;
  return SPLAT_OK;
}

splat_status splat_set_raster_config(splat_context* ctx, const splat_raster_config* config) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!config) {
    ctx->forced = false;
    return SPLAT_OK;
  }
  if (config->forward_shape < 0 || config->forward_shape > 2 ||
      (config->backward_batch != 64 && config->backward_batch != 128))
    return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "set_raster_config: forward_shape 0..2, backward_batch 64 or 128");
  ctx->forced = true;
  ctx->forced_cfg.fwd_shape = config->forward_shape;
  ctx->forced_cfg.fwd_queue = config->forward_queue != 0;
  ctx->forced_cfg.bwd_batch = config->backward_batch;
  ctx->forced_cfg.bwd_queue = config->backward_queue != 0;
  return SPLAT_OK;
}

splat_status splat_get_raster_config(const splat_context* ctx, splat_raster_config* out) {
  if (!ctx || !out) return SPLAT_ERROR_INVALID_ARGUMENT;
  out->forward_shape = ctx->rcfg.fwd_shape;
  out->forward_queue = ctx->rcfg.fwd_queue ? 1 : 0;
  out->backward_batch = ctx->rcfg.bwd_batch;
  out->backward_queue = ctx->rcfg.bwd_queue ? 1 : 0;
  return SPLAT_OK;
}

splat_status splat_download_stats(splat_context* ctx, float* screen_gradient, int32_t* visible_count) {
  if (splat_status s = check_ready(ctx)) return s;
  if (screen_gradient)
    CK(hipMemcpyAsync(screen_gradient, ctx->stat_sg.f(), 4 * (size_t)ctx->n, hipMemcpyDeviceToHost, ctx->stream),
       "download");
  if (visible_count)
    CK(hipMemcpyAsync(visible_count, ctx->stat_vc.i(), 4 * (size_t)ctx->n, hipMemcpyDeviceToHost, ctx->stream),
       "download");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

splat_status splat_upload_stats(splat_context* ctx, const float* screen_gradient, const int32_t* visible_count) {
  if (splat_status s = check_ready(ctx)) return s;
  if (screen_gradient)
    CK(hipMemcpyAsync(ctx->stat_sg.f(), screen_gradient, 4 * (size_t)ctx->n, hipMemcpyHostToDevice, ctx->stream),
       "upload");
  if (visible_count)
    CK(hipMemcpyAsync(ctx->stat_vc.i(), visible_count, 4 * (size_t)ctx->n, hipMemcpyHostToDevice, ctx->stream),
       "upload");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

namespace {
// Copies between the capacity layout and the count-concatenated layout.
splat_status moments_copy(splat_context* ctx, float* host_m, float* host_v, bool to_host) {
  const long long cap = ctx->cap, n = ctx->n;
  const long long offs[6] = {0, 3 * cap, 6 * cap, 10 * cap, 11 * cap, 14 * cap};
  const long long widths[6] = {3, 3, 4, 1, 3, ctx->rs};
  long long packed = 0;
  for (int q = 0; q < 6; ++q) {
    const size_t bytes = 4 * (size_t)(widths[q] * n);
    if (bytes > 0) {
      for (int which = 0; which < 2; ++which) {
        float* host = which == 0 ? host_m : host_v;
        float* dev = (which == 0 ? ctx->m.f() : ctx->v.f()) + offs[q];
        if (!host) continue;
        if (to_host)
          CK(hipMemcpyAsync(host + packed, dev, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
        else
          CK(hipMemcpyAsync(dev, host + packed, bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
      }
    }
    packed += widths[q] * n;
  }
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}
}  // namespace

splat_status splat_download_moments(splat_context* ctx, float* m, float* v) {
  if (splat_status s = check_ready(ctx)) return s;
  return moments_copy(ctx, m, v, true);
}

splat_status splat_upload_moments(splat_context* ctx, const float* m, const float* v) {
  if (splat_status s = check_ready(ctx)) return s;
  return moments_copy(ctx, const_cast<float*>(m), const_cast<float*>(v), false);
}

splat_status splat_set_timing(splat_context* ctx, int32_t enabled) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  ctx->timing = enabled != 0;
  return SPLAT_OK;
}

int32_t splat_timing_count(const splat_context* ctx) { return ctx ? (int32_t)ctx->timers.size() : 0; }

splat_status splat_timing_entry(splat_context* ctx, int32_t i, const char** name, double* total_ms, int64_t* launches) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  CK(ctx->collect_timing(), "timing");
  if (i < 0 || i >= (int)ctx->timers.size()) return fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "timing index");
  auto it = ctx->timers.begin();
  std::advance(it, i);
  if (name) *name = it->first.c_str();
  if (total_ms) *total_ms = it->second.total_ms;
  if (launches) *launches = it->second.launches;
  return SPLAT_OK;
}

splat_status splat_reset_timing(splat_context* ctx) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  CK(ctx->collect_timing(), "timing");
  ctx->timers.clear();
  return SPLAT_OK;
}

int64_t splat_last_pair_count(const splat_context* ctx) { return ctx ? (int64_t)ctx->P : 0; }

splat_status splat_memory(splat_context* ctx, uint64_t* context_bytes, uint64_t* device_used_bytes) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  CK(hipSetDevice(ctx->device), "set device");
  uint64_t mine = 0;
  for (DeviceBuffer* b : ctx->all_buffers()) mine += b->bytes;
  if (context_bytes) *context_bytes = mine;
  if (device_used_bytes) {
    size_t free_b = 0, total_b = 0;
    CK(hipMemGetInfo(&free_b, &total_b), "memory info");
    *device_used_bytes = (uint64_t)(total_b - free_b);
  }
  return SPLAT_OK;
}

splat_status splat_synchronize(splat_context* ctx) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return SPLAT_OK;
}

}  // extern "C"
