// The context behind the opaque splat_context handle (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/splat_hip.h"
#include "splat_common.h"
#include "splat_kernels.h"

// A device allocation that only grows.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t want);
  void release();
  float* f() const { return static_cast<float*>(ptr); }
  uint32_t* u() const { return static_cast<uint32_t*>(ptr); }
  int* i() const { return static_cast<int*>(ptr); }
};

// Bumped whenever any DeviceBuffer (re)allocates or frees: a captured graph
// recorded under an older value may hold stale pointers.
unsigned long long alloc_generation();

struct KernelTimer {
  double total_ms = 0.0;
  long long launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct splat_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;

  // scene: parameters, gradients and moments laid out by capacity
  long long cap = 0;
  int sh_degree = 0;
  int rs = 0;
  int n = 0;
  int cloud_degree = 0;
  DeviceBuffer params, grads, m, v;
  DeviceBuffer sg, vc;            // this step's screen gradient / visible count
  DeviceBuffer stat_sg, stat_vc;  // accumulated since the last densify

  // cameras and targets
  std::vector<splat_view> views;
  std::vector<long long> target_offset;
  std::vector<bool> has_target;
  DeviceBuffer targets;

  // the current batch
  splat_raster_options opts{};
  std::vector<int> batch_ids;
  std::vector<splat::ViewDesc> batch_desc;
  DeviceBuffer d_views;
  int total_tiles = 0;
  int loss_tiles = 0;
  long long total_px = 0;
  int sh_use = 0;
  long long M = 0;
  unsigned long long P = 0;
  splat::RasterParams raster{};
  bool forward_done = false;
  bool loss_done = false;
  // the stored gradient (grads) and the step's sg / vc are all zeros: true
  // after allocation, upload, densify, a clear, an Adam step and every
  // training step; false after a standalone adjoint, which adds into them
  bool grads_clear = true;
  float last_ssim_weight = 0.f;
  float ssim_kernel[11] = {};

  DeviceBuffer rec, keys, vals, keys_alt, vals_alt, counts, offsets, pair_base;
  DeviceBuffer scounts, scounts_alt, soffsets;      // super-tile pairs per slot (sorted along), their scan
  DeviceBuffer keys_c, vals_c, scounts_c;           // the depth sort's third buffers (three passes)
  DeviceBuffer key_ranges;                          // per projection wave: min, max of its kept keys
  DeviceBuffer rects;                                // each slot's tile rect, 8 bytes
  DeviceBuffer skeys, svals, skeys_alt, svals_alt;  // (super-tile, slot) pairs
  DeviceBuffer sranges, nseg, seg_base, seg_st, bin_table;
  DeviceBuffer pair_vals, ranges;                   // per-tile lists (slots), [begin, end) per tile
  unsigned long long S = 0;                         // super-tile pairs of the batch
  int n_st = 0;                                     // super-tiles of the batch
  // every view at most kPackedTiles tiles a side: packed rects ride along with
  // the depth sort (in scounts), and the emitter reads them in depth order
  bool packed = false;
  int st_view_bits = -1;  // super-tiles numbered view * 256 + id within the view: bits of the last view index; else -1
  uint32_t* sorted_prects = nullptr;
  uint32_t* sorted_slots = nullptr;
  uint32_t* sorted_pair_vals = nullptr;
  // capacities the pair-count-sized buffers were reserved for (kernel bounds)
  unsigned long long P_cap = 0, S_cap = 0;

  // compositing launch shape: the current one, and the tuned one per batch shape
  splat::RasterConfig rcfg{};
  bool autotune = false;
  bool forced = false;  // forced_cfg for every batch (splat_set_raster_config)
  splat::RasterConfig forced_cfg{};
  bool tuning_backward = false;  // the forward of this batch was just tuned
  std::map<std::string, splat::RasterConfig> tuned;
  std::string shape_key;  // resolutions of the current batch
  DeviceBuffer order_keys, order_vals, order_keys_alt, order_vals_alt;  // tiles, longest list first
  uint32_t* tile_order = nullptr;

  // captured training steps
  struct Graph {
    std::string key;
    hipGraphExec_t exec = nullptr;
    unsigned long long generation = 0;
    uint32_t* sorted_slots = nullptr;
    splat::RasterParams raster{};
  };
  bool graph_mode = false;
  bool capturing = false;
  unsigned long long graph_p_cap = 0, graph_s_cap = 0;  // capacities graph steps run with
  std::vector<Graph> graphs;
  splat::AdamStep adam_host{};
  DeviceBuffer d_adam;
  uint32_t* readback = nullptr;  // pinned: loss sums then the step counters
  size_t readback_bytes = 0;
  void drop_graphs();
  DeviceBuffer rgb, T, last, dL, terms, block_sums, loss_sums;
  DeviceBuffer screen, partial, visit, carry;
  DeviceBuffer scan_ws, hist_ws, scalars;  // scalars: [0] tile pairs, [1] super-tile pairs, [2] segments, [3] overflow
  splat::SortWorkspace sort_ws;

  // densify scratch
  DeviceBuffer scratch_params, scratch_m, scratch_v;
  DeviceBuffer dens_flags, dens_a, dens_b, dens_c, dens_src;

  // timing
  bool timing = false;
  std::map<std::string, KernelTimer> timers;

  ~splat_context();
  std::vector<DeviceBuffer*> all_buffers();
  hipError_t ensure_sort_workspace(long long max_n);
  hipError_t timed_begin(const char* name, hipEvent_t* start);
  hipError_t timed_end(const char* name, hipEvent_t start);
  hipError_t collect_timing();
  splat::SplatRec* recs() const { return static_cast<splat::SplatRec*>(rec.ptr); }
  const splat::ViewDesc* views_dev() const { return static_cast<const splat::ViewDesc*>(d_views.ptr); }
};
