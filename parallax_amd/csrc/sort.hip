// Device scan, stable LSD radix sort, and tile binning.
//
// Binning follows the engine's buildTiles (engine/native/cpp/src/splat/
// SplatRaster.cpp:114-151): every projected splat is entered into each
// 16x16 tile of its rect, and each tile's list runs near to far. The engine
// sorts all projected splats by depth (:260-261) and then scatters them
// stably into tiles. Here the same order comes from two stable sorts: the
// (view, gaussian) slots by float depth bits, then the emitted (tile, slot)
// pairs by tile id. Exact depth ties keep slot order, which is gaussian index
// order within a view -- the checker's stable_sort gives the same order.
//
// The radix sort is a classic three-kernel LSD sort with 8-bit digits over
// 4096-element blocks: per-block digit histograms (per-wave LDS copies to
// limit atomic contention), an exclusive scan of the digit-major histogram
// table, and a stable scatter whose in-block ranks come from wave64 ballots
// (one ballot per digit bit selects the lanes sharing a digit, and a popcount
// below the lane gives the rank). All global loads and stores of keys and
// values are coalesced; the scatter's stores are grouped by digit.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    const uint32_t s = lds[w];
    if (w < wave) before += s;
    tot += s;
  }
  *total = tot;
  __syncthreads();
  return x - v + before;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t* in, const uint32_t* idx, uint32_t* out,
                                                           long long n, uint32_t* block_sums, uint32_t* total) {
  __shared__ uint32_t lds[kScanThreads / 64];
  const long long base = (long long)blockIdx.x * kScanTile;
  uint32_t running = 0;
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * kScanThreads + threadIdx.x;
    uint32_t v = 0;
    if (j < n) v = in[idx ? idx[j] : j];
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, &tot, lds);
    if (j < n) out[j] = running + ex;
    running += tot;
  }
  if (threadIdx.x == 0) {
    if (block_sums) block_sums[blockIdx.x] = running;
    if (total && gridDim.x == 1) *total = running;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* out, long long n, const uint32_t* offsets) {
  const uint32_t add = offsets[blockIdx.x];
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * kScanThreads + threadIdx.x;
    if (j < n) out[j] += add;
  }
}

hipError_t scan_level(const uint32_t* in, const uint32_t* idx, uint32_t* out, long long n, uint32_t* total,
                      uint32_t* ws, hipStream_t stream) {
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_local, dim3(1), dim3(kScanThreads), 0, stream, in, idx, out, n, nullptr, total);
    return hipGetLastError();
  }
  uint32_t* sums = ws;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, in, idx, out, n, sums,
                     nullptr);
  SPLAT_HIP_CHECK(hipGetLastError());
  SPLAT_HIP_CHECK(scan_level(sums, nullptr, sums, nb, total, ws + nb, stream));
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, out, n, sums);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ keys, long long n, int shift,
                                                    uint32_t mask, uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t h[4][256];
  const int tid = threadIdx.x, wave = tid >> 6;
  for (int w = 0; w < 4; ++w) h[w][tid] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * 256 + tid;
    if (j < n) atomicAdd(&h[wave][(keys[j] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(long long)tid * nblocks + blockIdx.x] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals,
                                                       uint32_t* __restrict__ keys_out,
                                                       uint32_t* __restrict__ vals_out, long long n, int shift,
                                                       int bits, const uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_cnt[4][256];
  const int tid = threadIdx.x, wave = tid >> 6;
  s_base[tid] = hist[(long long)tid * nblocks + blockIdx.x];
  const uint32_t mask = (1u << bits) - 1u;
  const long long base = (long long)blockIdx.x * kScanTile;
  const uint64_t below = __lanemask_lt();
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * 256 + tid;
    const bool valid = j < n;
    const uint32_t k = valid ? keys[j] : 0u;
    const uint32_t vv = valid ? vals[j] : 0u;
    const uint32_t d = (k >> shift) & mask;
#pragma unroll
    for (int w = 0; w < 4; ++w) s_cnt[w][tid] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0) s_cnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t run = s_base[tid];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint32_t c = s_cnt[w][tid];
        s_cnt[w][tid] = run;
        run += c;
      }
      s_base[tid] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = s_cnt[wave][d] + rank;
      keys_out[pos] = k;
      vals_out[pos] = vv;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_emit_pairs(const uint32_t* __restrict__ sorted_slots,
                                                    const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ counts,
                                                    const SplatRec* __restrict__ rec,
                                                    const ViewDesc* __restrict__ views, long long per_view,
                                                    long long m, uint32_t* __restrict__ pair_keys,
                                                    uint32_t* __restrict__ pair_vals,
                                                    uint32_t* __restrict__ pair_base) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const uint32_t slot = sorted_slots[s];
  const uint32_t c = counts[slot];
  if (c == 0) return;
  const uint32_t off = offsets[s];
  const int v = (int)(slot / per_view);
  const SplatRec& r = rec[slot];
  const int cols = views[v].cols, tile_base = views[v].tile_base;
  pair_base[slot] = off;
  uint32_t k = off;
  for (int ty = r.y0; ty <= r.y1; ++ty)
    for (int tx = r.x0; tx <= r.x1; ++tx) {
      pair_keys[k] = (uint32_t)(tile_base + ty * cols + tx);
      pair_vals[k] = slot;
      ++k;
    }
}

__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* __restrict__ keys, long long p,
                                                     uint32_t* __restrict__ ranges) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) ranges[2 * (long long)k] = (uint32_t)i;
  if (i == p - 1 || keys[i + 1] != k) ranges[2 * (long long)k + 1] = (uint32_t)(i + 1);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

unsigned grid_for(long long n, int threads, long long max_blocks = 65535LL * 16) {
  long long b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

}  // namespace

long long scan_workspace_elems(long long max_n) {
  long long total = 0, n = max_n;
  while (n > kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    total += n;
  }
  return total + 16;
}

long long sort_hist_elems(long long max_n) { return 256LL * ((max_n + kScanTile - 1) / kScanTile + 1); }

hipError_t launch_exclusive_scan(const uint32_t* in, const uint32_t* gather_idx, uint32_t* out, long long n,
                                 uint32_t* total, const ScanWorkspace& ws, hipStream_t stream) {
  if (n <= 0) return total ? launch_fill_u32(total, 0u, 1, stream) : hipSuccess;
  if (n > ws.max_n) return hipErrorInvalidValue;
  return scan_level(in, gather_idx, out, n, total, ws.buf, stream);
}

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, long long n,
                             int begin_bit, int end_bit, const SortWorkspace& ws, bool* result_in_alt,
                             hipStream_t stream) {
  *result_in_alt = false;
  if (n <= 1 || end_bit <= begin_bit) return hipSuccess;
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (256 * nb > ws.hist_capacity) return hipErrorInvalidValue;
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int bits = (end_bit - shift) < 8 ? (end_bit - shift) : 8;
    const uint32_t mask = (1u << bits) - 1u;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(256), 0, stream, ki, n, shift, mask, ws.hist,
                       (int)nb);
    SPLAT_HIP_CHECK(hipGetLastError());
    SPLAT_HIP_CHECK(launch_exclusive_scan(ws.hist, nullptr, ws.hist, 256 * nb, nullptr, ws.scan, stream));
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(256), 0, stream, ki, vi, ko, vo, n, shift, bits,
                       ws.hist, (int)nb);
    SPLAT_HIP_CHECK(hipGetLastError());
    uint32_t* t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
    *result_in_alt = !*result_in_alt;
  }
  return hipSuccess;
}

hipError_t launch_emit_pairs(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* counts,
                             const SplatRec* rec, const ViewDesc* views, int n_views, long long per_view,
                             long long m, uint32_t* pair_keys, uint32_t* pair_vals, uint32_t* pair_base,
                             hipStream_t stream) {
  (void)n_views;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, sorted_slots, offsets,
                     counts, rec, views, per_view, m, pair_keys, pair_vals, pair_base);
  return hipGetLastError();
}

hipError_t launch_tile_ranges(const uint32_t* sorted_keys, long long p, uint32_t* ranges, int total_tiles,
                              hipStream_t stream) {
  SPLAT_HIP_CHECK(launch_fill_u32(ranges, 0u, 2LL * total_tiles, stream));
  if (p <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, stream, sorted_keys, p,
                     ranges);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* ptr, uint32_t value, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, stream, ptr, value, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* ptr, float value, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256)), dim3(256), 0, stream, ptr, value, n);
  return hipGetLastError();
}

}  // namespace splat
