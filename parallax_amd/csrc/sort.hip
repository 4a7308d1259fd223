// Device scan and stable LSD radix sort.
//
// The binning uses two stable sorts: the (view, gaussian) slots by float depth
// bits (the engine sorts its projected splats by depth, engine/native/cpp/src/
// splat/SplatRaster.cpp:260-261; exact ties keep slot order, which is gaussian
// index order within a view, as the checker's stable_sort), then the
// (super-tile, slot) pairs by super-tile id.
//
// The radix sort is a three-kernel LSD sort with 8-bit digits over
// 4096-element blocks: per-block digit histograms (per-wave LDS copies to
// limit atomic contention), an exclusive scan of the digit-major histogram
// table, and a stable scatter whose in-block ranks come from wave64 ballots
// (one ballot per digit bit selects the lanes sharing a digit, and a popcount
// below the lane gives the rank). The scatter sorts each block by digit in
// LDS first, so its global stores go out as contiguous per-digit runs. An
// optional second value array moves with the first (the depth sort carries
// each slot's super-tile count, so its scan reads them in order).
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

// One 4096-item tile per workgroup. Items are loaded coalesced (item
// r * 256 + thread) into a padded LDS tile, each thread then scans its 16
// consecutive items serially, one workgroup scan combines the thread totals,
// and the prefixes go back out coalesced. (Every read precedes the first
// barrier, so the scan may run in place.)
__device__ __forceinline__ int scan_pad(int i) { return i + (i >> 5); }

// n_dev (nullable): the item count, read on the device and clamped to n
// (the capacity the grid was sized for).
__device__ __forceinline__ long long item_count(long long n, const uint32_t* n_dev) {
  if (!n_dev) return n;
  const long long d = (long long)*n_dev;
  return d < n ? d : n;
}

// The value a scan or sum takes of an input word.
__device__ __forceinline__ uint32_t scan_value(uint32_t x, int xform) {
  return xform == kScanRectTiles ? rect_tiles(x) : (xform == kScanRectSupers ? rect_supers(x) : x);
}

// With side_total, each block also sums side_xform of its inputs: into
// side_parts[block] when given (a later sum adds them to *side_total: one
// word's atomics from thousands of blocks would queue behind each other),
// else with one atomic add to *side_total.
__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t* in, const uint32_t* idx, uint32_t* out,
                                                           long long n_cap, const uint32_t* n_dev, int xform,
                                                           uint32_t* block_sums, uint32_t* total,
                                                           uint32_t* side_total, int side_xform,
                                                           uint32_t* side_parts) {
  __shared__ uint32_t lds[kScanThreads / 64];
  __shared__ uint32_t tile[kScanTile + kScanTile / 32];
  const long long n = item_count(n_cap, n_dev);
  const long long base = (long long)blockIdx.x * kScanTile;
  const int tid = threadIdx.x;
  uint32_t side = 0;
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * kScanThreads + tid;
    const uint32_t x = j < n ? in[idx ? idx[j] : j] : 0u;
    tile[scan_pad(r * kScanThreads + tid)] = j < n ? scan_value(x, xform) : 0u;
    if (side_total && j < n) side += scan_value(x, side_xform);
  }
  if (side_total) {  // block-uniform
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) side += (uint32_t)__shfl_xor((int)side, d, 64);
  }
  __syncthreads();
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = tile[scan_pad(tid * kScanItems + k)];
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(sum, &tot, lds);  // has barriers: the tile reads above are done
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    tile[scan_pad(tid * kScanItems + k)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * kScanThreads + tid;
    if (j < n) out[j] = tile[scan_pad(r * kScanThreads + tid)];
  }
  if (side_total) {  // after every store
    if ((tid & 63) == 0) lds[tid >> 6] = side;
    __syncthreads();
    if (tid == 0) {
      uint32_t sum = 0;
#pragma unroll
      for (int w = 0; w < kScanThreads / 64; ++w) sum += lds[w];
      if (side_parts) side_parts[blockIdx.x] = sum;
      else if (sum) atomicAdd(side_total, sum);
    }
  }
  if (tid == 0) {
    if (block_sums) block_sums[blockIdx.x] = tot;
    if (total && gridDim.x == 1) *total = tot;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* out, long long n_cap, const uint32_t* n_dev,
                                                         const uint32_t* offsets) {
  const long long n = item_count(n_cap, n_dev);
  const uint32_t add = offsets[blockIdx.x];
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = base + (long long)r * kScanThreads + threadIdx.x;
    if (j < n) out[j] += add;
  }
}

// The first level reads n_dev and the transform; the levels above scan plain
// block sums (zero for blocks past the items).
hipError_t scan_level(const uint32_t* in, const uint32_t* idx, uint32_t* out, long long n, const uint32_t* n_dev,
                      int xform, uint32_t* total, uint32_t* side_total, int side_xform, uint32_t* ws,
                      hipStream_t stream) {
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_local, dim3(1), dim3(kScanThreads), 0, stream, in, idx, out, n, n_dev, xform, nullptr,
                       total, side_total, side_xform, nullptr);
    return hipGetLastError();
  }
  // the level's block sums, then (with side_total) the blocks' side sums
  uint32_t* sums = ws;
  uint32_t* side_parts = side_total ? ws + nb : nullptr;
  uint32_t* above = ws + (side_total ? 2 * nb : nb);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, in, idx, out, n, n_dev, xform,
                     sums, nullptr, side_total, side_xform, side_parts);
  SPLAT_HIP_CHECK(hipGetLastError());
  SPLAT_HIP_CHECK(scan_level(sums, nullptr, sums, nb, nullptr, kScanPlain, total, nullptr, kScanPlain, above,
                             stream));
  if (side_total) SPLAT_HIP_CHECK(launch_sum_u32(side_parts, nb, side_total, stream));
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, out, n, n_dev, sums);
  return hipGetLastError();
}

// With drop, keys equal to kDropKey are not counted (the scatter drops them).
constexpr uint32_t kDropKey = 0xFFFFFFFFu;

// One radix pass's buffers: set k of keys, values and second values is
// (k[i], v[i], w[i]). Without a plan the pass reads set 0 and writes set 1.
// With a plan (*passes on the device, 2 to 4) a pass at or past *passes does
// nothing, and the sets are chosen so the last pass run writes set 0: two or
// four passes go 0 -> 1 -> 0 (-> 1 -> 0), three go 0 -> 1 -> 2 -> 0, and the
// first pass takes every kept key less *key_base.
struct SortRoute {
  uint32_t* k[3];
  uint32_t* v[3];
  uint32_t* w[3];
  const uint32_t* passes;
  const uint32_t* key_base;
  int pass;
};

// kPlan: the route is read from the plan (the depth sort); otherwise set 0
// to set 1 with keys as they are, known at compile time.
template <bool kPlan>
__device__ __forceinline__ bool route_sets(const SortRoute& R, int& src, int& dst) {
  if (!kPlan) {
    src = 0;
    dst = 1;
    return true;
  }
  const int p = (int)*R.passes, i = R.pass;
  if (i >= p) return false;
  src = i == 0 ? 0 : (p == 3 ? (i == 1 ? 1 : 2) : (i & 1));
  dst = i == p - 1 ? 0 : (p == 3 ? (i == 0 ? 1 : 2) : ((i & 1) ? 0 : 1));
  return true;
}

// Set k's buffer of one array, chosen by selects (no indexed access).
__device__ __forceinline__ uint32_t* route_pick(uint32_t* const (&a)[3], int k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
}

// Per-block digit counts, into per-wave LDS copies. With kAggregate, each
// round of 64 items finds the lanes sharing a digit (one ballot per digit
// bit) and only the group's lowest lane adds the group's size: a skewed
// digit (a narrow last digit, the few exponents in the top byte of positive
// float keys) then costs one LDS atomic per round, not up to 64; for spread
// digits the plain atomics are cheaper.
template <bool kAggregate, bool kPlan>
__global__ __launch_bounds__(256) void k_radix_hist(SortRoute R, long long n_cap,
                                                    const uint32_t* __restrict__ n_dev, int shift, int bits,
                                                    int drop, uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t h[4][256];
  int src, dst;
  if (!route_sets<kPlan>(R, src, dst)) return;  // a pass the plan skips (uniform)
  const uint32_t* __restrict__ keys = route_pick(R.k, src);
  // the first pass of a plan takes the keys less the base
  const uint32_t key_base = (kPlan && R.pass == 0) ? *R.key_base : 0u;
  const long long n = item_count(n_cap, n_dev);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t mask = (1u << bits) - 1u;
  for (int w = 0; w < 4; ++w) h[w][tid] = 0;
  __syncthreads();
  // the wave's own 1024 consecutive items
  const long long wbase = (long long)blockIdx.x * kScanTile + (long long)wave * (kScanTile / 4);
  uint32_t k[kScanItems];
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = wbase + (long long)r * 64 + lane;
    k[r] = j < n ? keys[j] : kDropKey;
    if (kPlan && k[r] != kDropKey) k[r] -= key_base;
  }
  const uint64_t below = __lanemask_lt();
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = wbase + (long long)r * 64 + lane;
    const bool valid = j < n && !(drop && k[r] == kDropKey);
    const uint32_t d = (k[r] >> shift) & mask;
    if constexpr (kAggregate) {
      uint64_t peers = __ballot(valid);
      for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      if (valid && (peers & below) == 0ull) atomicAdd(&h[wave][d], (uint32_t)__popcll(peers));
    } else {
      if (valid) atomicAdd(&h[wave][d], 1u);
    }
  }
  __syncthreads();
  hist[(long long)tid * nblocks + blockIdx.x] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// Stable scatter of one 4096-item block for one digit. Each wave ranks its
// own contiguous 1024 items with no block barriers: per round of 64, the
// lanes sharing a digit are found with one ballot per digit bit, and the
// group's lowest lane adds the group's size to a wave-private LDS counter per
// digit with a returning atomic, whose old value is the group's base. Then
// per-digit wave offsets and block digit starts place every item at its
// stable position in an LDS copy of the block, sorted by digit; the block is
// written out from LDS so each digit's run goes to global memory contiguously.
constexpr int kRsWaveItems = kScanTile / 4;  // 1024 items per wave
constexpr int kRsRounds = kRsWaveItems / 64;  // 16

// With kV2, a second value array rides along with the first (staged through
// LDS after the keys and values, so its stores are runs too). Without vals
// (null), item j's value is j.
template <bool kV2, bool kPlan>
__global__ __launch_bounds__(256) void k_radix_scatter(SortRoute R, int identity, long long n_cap,
                                                       const uint32_t* __restrict__ n_dev, int shift, int bits,
                                                       int drop, const uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ hist_prefix, int nblocks) {
  int src, dst;
  if (!route_sets<kPlan>(R, src, dst)) return;  // a pass the plan skips (uniform)
  const uint32_t* __restrict__ keys = route_pick(R.k, src);
  const uint32_t* __restrict__ vals = identity ? nullptr : route_pick(R.v, src);
  const uint32_t* __restrict__ vals2 = route_pick(R.w, src);
  uint32_t* __restrict__ keys_out = route_pick(R.k, dst);
  uint32_t* __restrict__ vals_out = route_pick(R.v, dst);
  uint32_t* __restrict__ vals2_out = route_pick(R.w, dst);
  const uint32_t key_base = (kPlan && R.pass == 0) ? *R.key_base : 0u;
  const long long n = item_count(n_cap, n_dev);
  if ((long long)blockIdx.x * kScanTile >= n) return;  // whole block past the items (uniform)
  __shared__ uint32_t s_cnt[4][256];   // per-wave digit counts, then per-wave digit offsets
  __shared__ uint32_t s_start[256];    // block-local start of each digit
  __shared__ uint32_t s_global[256];   // global start of each digit for this block
  __shared__ uint32_t s_count;         // items of the block kept
  __shared__ uint32_t s_keys[kScanTile];
  __shared__ uint32_t s_vals[kScanTile];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t mask = (1u << bits) - 1u;
  const long long base = (long long)blockIdx.x * kScanTile;
  const long long wbase = base + (long long)wave * kRsWaveItems;
#pragma unroll
  for (int w = 0; w < 4; ++w) s_cnt[w][tid] = 0;
  {
    // the block-level scan of the digit table, plus its second level's
    // prefix when the table spans several scan blocks
    const long long e = (long long)tid * nblocks + blockIdx.x;
    s_global[tid] = hist[e] + (hist_prefix ? hist_prefix[e / kScanTile] : 0u);
  }
  __syncthreads();
  const uint64_t below = __lanemask_lt();
  uint32_t k[kRsRounds], v[kRsRounds], v2[kV2 ? kRsRounds : 1], rank[kRsRounds];
  int leader[kRsRounds];
  uint32_t old[kRsRounds];
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const long long j = wbase + (long long)r * 64 + lane;
    k[r] = j < n ? keys[j] : kDropKey;
    if (kPlan && k[r] != kDropKey) k[r] -= key_base;
    const bool valid = j < n && !(drop && k[r] == kDropKey);
    v[r] = valid ? (vals ? vals[j] : (uint32_t)j) : 0u;
    if constexpr (kV2) v2[r] = valid ? vals2[j] : 0u;
    const uint32_t d = (k[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : ~bb;
    }
    // The group's lowest lane adds the group's size to the wave's digit
    // counter; the counter's old value is the group's base. The rounds'
    // atomics go out back to back (LDS applies them in order) and the bases
    // are fetched after the last one.
    rank[r] = valid ? (uint32_t)__popcll(peers & below) : 0xFFFFFFFFu;
    leader[r] = valid ? __builtin_ctzll(peers) : lane;
    old[r] = 0u;
    if (valid && (peers & below) == 0ull) old[r] = atomicAdd(&s_cnt[wave][d], (uint32_t)__popcll(peers));
  }
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const uint32_t base = (uint32_t)__shfl((int)old[r], leader[r], 64);
    if (rank[r] != 0xFFFFFFFFu) rank[r] += base;
  }
  __syncthreads();
  // per digit: offsets of each wave inside the digit, and the digit's total
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t c = s_cnt[w][tid];
    s_cnt[w][tid] = total;
    total += c;
  }
  // block-wide exclusive scan of the digit totals (thread = digit)
  {
    uint32_t x = total;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    __shared__ uint32_t s_wsum[4];
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += s_wsum[w];
    s_start[tid] = x - total + add;
    if (tid == 255) s_count = x + add;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    if (rank[r] == 0xFFFFFFFFu) continue;
    const uint32_t d = (k[r] >> shift) & mask;
    const uint32_t pos = s_start[d] + s_cnt[wave][d] + rank[r];
    s_keys[pos] = k[r];
    s_vals[pos] = v[r];
  }
  __syncthreads();
  const int count = (int)s_count;
  for (int p = tid; p < count; p += 256) {
    const uint32_t kk = s_keys[p];
    const uint32_t d = (kk >> shift) & mask;
    const uint32_t dst = s_global[d] + (uint32_t)p - s_start[d];
    keys_out[dst] = kk;
    vals_out[dst] = s_vals[p];
  }
  if constexpr (kV2) {
    __syncthreads();  // every value is out of s_vals: it takes the second values
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
      if (rank[r] == 0xFFFFFFFFu) continue;
      const uint32_t d = (k[r] >> shift) & mask;
      s_vals[s_start[d] + s_cnt[wave][d] + rank[r]] = v2[r];
    }
    __syncthreads();
    for (int p = tid; p < count; p += 256) {
      const uint32_t d = (s_keys[p] >> shift) & mask;
      vals2_out[s_global[d] + (uint32_t)p - s_start[d]] = s_vals[p];
    }
  }
}

// *out += the sum of in[0..n): per workgroup a register and LDS reduction,
// then one integer atomic (exact whatever the order).
__global__ __launch_bounds__(256) void k_sum_u32(const uint32_t* __restrict__ in, long long n, uint32_t* out) {
  __shared__ uint32_t s_part[4];
  uint32_t sum = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    sum += in[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) sum += (uint32_t)__shfl_xor((int)sum, d, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t total = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (total) atomicAdd(out, total);
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// 16-byte stores over a 16-byte-aligned buffer of n4 float4s.
__global__ void k_fill_f32x4(float4* p, float v, long long n4) {
  const float4 q = make_float4(v, v, v, v);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    p[i] = q;
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
  // + the first level's side sums
  return total + (max_n + kScanTile - 1) / kScanTile + 16;
}

long long sort_hist_elems(long long max_n) { return 256LL * ((max_n + kScanTile - 1) / kScanTile + 1); }

hipError_t launch_exclusive_scan(const uint32_t* in, const uint32_t* gather_idx, uint32_t* out, long long n,
                                 uint32_t* total, const ScanWorkspace& ws, hipStream_t stream, int xform,
                                 const uint32_t* n_dev, uint32_t* side_total, int side_xform) {
  if (n <= 0) return total ? launch_fill_u32(total, 0u, 1, stream) : hipSuccess;
  if (n > ws.max_n) return hipErrorInvalidValue;
  return scan_level(in, gather_idx, out, n, n_dev, xform, total, side_total, side_xform, ws.buf, stream);
}

// An in-place exclusive scan of data[0..n) whose readers add the second
// level themselves: afterwards the scan of element i is data[i] +
// (*prefix ? (*prefix)[i / kScanTile] : 0). Two levels when n fits them (the
// last level's add is skipped); otherwise the full scan, and *prefix = null.
// With n_dev (device), the scan covers min(n, *n_dev) items (a pass a sort
// plan skips gives 0: nothing is read or written).
static hipError_t scan_no_add(uint32_t* data, long long n, uint32_t* total, const ScanWorkspace& ws,
                              const uint32_t** prefix, hipStream_t stream, const uint32_t* n_dev = nullptr) {
  *prefix = nullptr;
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= 1 || nb > kScanTile || n > ws.max_n)
    return launch_exclusive_scan(data, nullptr, data, n, total, ws, stream, kScanPlain, n_dev);
  uint32_t* sums = ws.buf;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, data, nullptr, data, n, n_dev,
                     (int)kScanPlain, sums, nullptr, nullptr, (int)kScanPlain, nullptr);
  SPLAT_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_local, dim3(1), dim3(kScanThreads), 0, stream, sums, nullptr, sums, nb, n_dev,
                     (int)kScanPlain, nullptr, total, nullptr, (int)kScanPlain, nullptr);
  *prefix = sums;
  return hipGetLastError();
}

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, long long n,
                             int begin_bit, int end_bit, const SortWorkspace& ws, bool* result_in_alt,
                             hipStream_t stream, const uint32_t* n_dev, uint32_t* vals2, uint32_t* vals2_alt,
                             bool identity_vals, uint32_t* n_kept, const SortPlan* plan) {
  *result_in_alt = false;
  if (end_bit <= begin_bit || n <= 0) return hipSuccess;
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (256 * nb > ws.hist_capacity) return hipErrorInvalidValue;
  if (plan && (begin_bit != 0 || end_bit != 32 || (plan->vals2_c == nullptr) != (vals2 == nullptr)))
    return hipErrorInvalidValue;  // a plan covers a full 32-bit key
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt, *wi = vals2, *wo = vals2_alt;
  int pass = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++pass) {
    const int bits = (end_bit - shift) < 8 ? (end_bit - shift) : 8;
    const bool first = shift == begin_bit;
    // with n_kept, the first pass drops kDropKey and the later ones sort what it kept
    const int drop = n_kept && first ? 1 : 0;
    const uint32_t* count = n_kept && !first ? n_kept : n_dev;
    SortRoute R{};
    if (plan) {
      R = SortRoute{{keys, keys_alt, plan->keys_c}, {vals, vals_alt, plan->vals_c}, {vals2, vals2_alt, plan->vals2_c},
                    plan->passes, plan->key_base, pass};
    } else {
      R = SortRoute{{ki, ko, nullptr}, {vi, vo, nullptr}, {wi, wo, nullptr}, nullptr, nullptr, pass};
    }
    const bool aggregate = bits <= 4 || shift + bits == 32;
    auto hist_kernel = plan ? (aggregate ? k_radix_hist<true, true> : k_radix_hist<false, true>)
                            : (aggregate ? k_radix_hist<true, false> : k_radix_hist<false, false>);
    hipLaunchKernelGGL(hist_kernel, dim3((unsigned)nb), dim3(256), 0, stream, R, n, count, shift, bits, drop, ws.hist,
                       (int)nb);
    SPLAT_HIP_CHECK(hipGetLastError());
    const uint32_t* prefix = nullptr;
    SPLAT_HIP_CHECK(scan_no_add(ws.hist, 256 * nb, drop ? n_kept : nullptr, ws.scan, &prefix, stream,
                                plan ? plan->active + pass : nullptr));
    const int identity = (identity_vals && first) ? 1 : 0;
    auto scatter_kernel = plan ? (vals2 ? k_radix_scatter<true, true> : k_radix_scatter<false, true>)
                               : (vals2 ? k_radix_scatter<true, false> : k_radix_scatter<false, false>);
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)nb), dim3(256), 0, stream, R, identity, n, count, shift, bits,
                       drop, ws.hist, prefix, (int)nb);
    SPLAT_HIP_CHECK(hipGetLastError());
    if (!plan) {
      uint32_t* t = ki; ki = ko; ko = t;
      t = vi; vi = vo; vo = t;
      t = wi; wi = wo; wo = t;
      *result_in_alt = !*result_in_alt;
    }
  }
  return hipSuccess;
}

namespace {

// The depth sort's plan from the projection's per-wave ranges of kept keys:
// every kept key less the minimum fits the low 8 * passes bits, so the
// passes above them would only copy.
__global__ __launch_bounds__(1024) void k_sort_plan(const uint32_t* __restrict__ key_ranges, int n_waves,
                                                    uint32_t* __restrict__ plan) {
  __shared__ uint32_t s_lo[16], s_hi[16];
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (int w = threadIdx.x; w < n_waves; w += 1024) {
    lo = min(lo, key_ranges[2 * w]);
    hi = max(hi, key_ranges[2 * w + 1]);
  }
  lo = __ockl_wfred_min_u32(lo);
  hi = __ockl_wfred_max_u32(hi);
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 16; ++k) {
    lo = min(lo, s_lo[k]);
    hi = max(hi, s_hi[k]);
  }
  uint32_t passes = 2, base = 0;
  if (hi >= lo) {  // some key was kept
    base = lo;
    const uint32_t span = hi - lo;
    const int bits = span ? 32 - __clz((int)span) : 0;
    passes = (uint32_t)max(2, (bits + 7) / 8);
  }
  plan[0] = passes;
  for (int i = 0; i < 4; ++i) plan[1 + i] = (uint32_t)i < passes ? 0xFFFFFFFFu : 0u;
  plan[5] = base;
}

}  // namespace

hipError_t launch_sort_plan(const uint32_t* key_ranges, int n_waves, uint32_t* plan, hipStream_t stream) {
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(1024), 0, stream, key_ranges, n_waves, plan);
  return hipGetLastError();
}

hipError_t launch_sum_u32(const uint32_t* in, long long n, uint32_t* out, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sum_u32, dim3(grid_for(n, 256 * 8, 1024)), dim3(256), 0, stream, in, n, out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* ptr, uint32_t value, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, stream, ptr, value, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* ptr, float value, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if ((reinterpret_cast<uintptr_t>(ptr) & 15) == 0 && (n & 3) == 0) {
    hipLaunchKernelGGL(k_fill_f32x4, dim3(grid_for(n / 4, 256)), dim3(256), 0, stream, reinterpret_cast<float4*>(ptr),
                       value, n / 4);
  } else {
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256)), dim3(256), 0, stream, ptr, value, n);
  }
  return hipGetLastError();
}

}  // namespace splat
