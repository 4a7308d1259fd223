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
// With a plan (*passes on the device) a pass at or past *passes does nothing,
// and the last pass run always writes set 0: with a third set (k[2]) the
// input is in set 0 and two or four passes go 0 -> 1 -> 0 (-> 1 -> 0), three
// go 0 -> 1 -> 2 -> 0; without one, the input is in set 1 for an odd number
// of passes and in set 0 for an even one, and the passes alternate. A plan's
// first pass may also rewrite every kept key: less *key_base, and, when
// *view_shift < 32, with its item's view (index / view_size) at that bit.
struct SortRoute {
  uint32_t* k[3];
  uint32_t* v[3];
  uint32_t* w[3];
  const uint32_t* passes;
  const uint32_t* key_base;
  const uint32_t* view_shift;
  uint32_t view_size;
  float view_inv;  // 1 / view_size, rounded
  int pass;
};

// kPlan: the route is read from the plan; otherwise set 0 to set 1 with keys
// as they are, known at compile time.
template <bool kPlan>
__device__ __forceinline__ bool route_sets(const SortRoute& R, int& src, int& dst) {
  if (!kPlan) {
    src = 0;
    dst = 1;
    return true;
  }
  const int p = (int)*R.passes, i = R.pass;
  if (i >= p) return false;
  if (R.k[2]) {
    src = i == 0 ? 0 : (p == 3 ? (i == 1 ? 1 : 2) : (i & 1));
    dst = i == p - 1 ? 0 : (p == 3 ? (i == 0 ? 1 : 2) : ((i & 1) ? 0 : 1));
  } else {
    src = (p - i) & 1;
    dst = (p - i - 1) & 1;
  }
  return true;
}

// Set k's buffer of one array, chosen by selects (no indexed access).
__device__ __forceinline__ uint32_t* route_pick(uint32_t* const (&a)[3], int k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
}

// The first pass of a plan: a kept key's sort key (see SortRoute).
struct KeyRewrite {
  uint32_t base = 0u, shift = 32u, size = 1u;
  float inv = 1.f;
  bool on = false;
  __device__ __forceinline__ uint32_t operator()(uint32_t k, long long j) const {
    if (!on || k == kDropKey) return k;
    k -= base;
    if (shift < 32u) {
      // j / size: the float quotient is within one of it (j < 2^32 and at
      // most 64 views), and one step corrects it
      uint32_t v = (uint32_t)((float)j * inv);
      if ((unsigned long long)v * size > (unsigned long long)j) --v;
      else if ((unsigned long long)(v + 1) * size <= (unsigned long long)j) ++v;
      k |= v << shift;
    }
    return k;
  }
};
template <bool kPlan>
__device__ __forceinline__ KeyRewrite key_rewrite(const SortRoute& R) {
  KeyRewrite w;
  if (kPlan && R.pass == 0 && R.key_base) {
    w.on = true;
    w.base = *R.key_base;
    w.shift = R.view_shift ? *R.view_shift : 32u;
    w.size = R.view_size;
    w.inv = R.view_inv;
  }
  return w;
}

// Per-block digit counts (1 << kDigitBits digits), into per-wave LDS copies.
// With kAggregate, each round of 64 items finds the lanes sharing a digit
// (one ballot per digit bit) and only the group's lowest lane adds the
// group's size: a skewed digit (a narrow last digit, the few exponents in the
// top byte of positive float keys) then costs one LDS atomic per round, not
// up to 64; for spread digits the plain atomics are cheaper.
// A pass's workgroup: four waves over 4096 items for 8-bit digits, eight
// over 8192 for 9-bit ones (each wave 1024 items), so a digit's run in a
// workgroup averages 16 items either way.
template <int kDigitBits>
struct RadixShape {
  static constexpr int kWaves = kDigitBits == 9 ? 8 : 4;
  static constexpr int kThreads = 64 * kWaves;
  static constexpr int kItems = 1024 * kWaves;
};

template <bool kAggregate, bool kPlan, int kDigitBits>
__global__ __launch_bounds__(RadixShape<kDigitBits>::kThreads) void k_radix_hist(SortRoute R, long long n_cap,
                                                    const uint32_t* __restrict__ n_dev, int shift, int bits,
                                                    int drop, uint32_t* __restrict__ hist, int nblocks) {
  constexpr int kRadix = 1 << kDigitBits;
  constexpr int kWaves = RadixShape<kDigitBits>::kWaves, kThreads = RadixShape<kDigitBits>::kThreads;
  __shared__ uint32_t h[kWaves][kRadix];
  int src, dst;
  if (!route_sets<kPlan>(R, src, dst)) return;  // a pass the plan skips (uniform)
  const uint32_t* __restrict__ keys = route_pick(R.k, src);
  const KeyRewrite rewrite = key_rewrite<kPlan>(R);
  const long long n = item_count(n_cap, n_dev);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t mask = (1u << bits) - 1u;
  for (int d = tid; d < kRadix; d += kThreads)
    for (int w = 0; w < kWaves; ++w) h[w][d] = 0;
  __syncthreads();
  // the wave's own 1024 consecutive items
  const long long wbase = (long long)blockIdx.x * RadixShape<kDigitBits>::kItems + (long long)wave * 1024;
  uint32_t k[kScanItems];
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const long long j = wbase + (long long)r * 64 + lane;
    k[r] = j < n ? rewrite(keys[j], j) : kDropKey;
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
  for (int d = tid; d < kRadix; d += kThreads) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) c += h[w][d];
    hist[(long long)d * nblocks + blockIdx.x] = c;
  }
}

// Stable scatter of one 4096-item block for one digit. Each wave ranks its
// own contiguous 1024 items with no block barriers: per round of 64, the
// lanes sharing a digit are found with one ballot per digit bit, and the
// group's lowest lane adds the group's size to a wave-private LDS counter per
// digit with a returning atomic, whose old value is the group's base. The
// counters are 16-bit halves of 32-bit words (a wave counts at most 1024
// items), digits 2c and 2c + 1 in word c, so a 9-bit digit's block still
// fits four workgroups per CU. Then per-digit wave offsets and block digit
// starts place every item at its stable position in an LDS copy of the
// block, sorted by digit; the block is written out from LDS so each digit's
// run goes to global memory contiguously.
constexpr int kRsWaveItems = kScanTile / 4;  // 1024 items per wave
constexpr int kRsRounds = kRsWaveItems / 64;  // 16

__device__ __forceinline__ uint32_t half_of(uint32_t word, uint32_t d) { return (word >> (16u * (d & 1u))) & 0xFFFFu; }

// With kV2, a second value array rides along with the first (staged through
// LDS after the keys and values, so its stores are runs too). Without vals
// (identity), item j's value is j. With chunk_sums, the digit table holds
// the scan within each 4096-entry chunk and chunk_sums the chunks' totals,
// which every block scans for itself; then, with kept_total, block 0 also
// writes their sum (the items the dropping pass keeps).
template <bool kV2, bool kPlan, int kDigitBits>
__global__ __launch_bounds__(RadixShape<kDigitBits>::kThreads) void k_radix_scatter(SortRoute R, int identity,
                                                       long long n_cap,
                                                       const uint32_t* __restrict__ n_dev, int shift, int bits,
                                                       int drop, const uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ chunk_sums,
                                                       uint32_t* __restrict__ kept_total, int nblocks) {
  constexpr int kRadix = 1 << kDigitBits;
  constexpr int kWaves = RadixShape<kDigitBits>::kWaves, kThreads = RadixShape<kDigitBits>::kThreads;
  constexpr int kItems = RadixShape<kDigitBits>::kItems;
  constexpr int kWords = kRadix / 2;           // counter words per wave
  static_assert(kRadix == kThreads || kRadix == 2 * kThreads, "one or two digits per thread");
  constexpr int kPer = kRadix / kThreads;      // digits per thread in the digit-wise steps (1 or 2)
  int src, dst;
  if (!route_sets<kPlan>(R, src, dst)) return;  // a pass the plan skips (uniform)
  const uint32_t* __restrict__ keys = route_pick(R.k, src);
  const uint32_t* __restrict__ vals = identity ? nullptr : route_pick(R.v, src);
  const uint32_t* __restrict__ vals2 = route_pick(R.w, src);
  uint32_t* __restrict__ keys_out = route_pick(R.k, dst);
  uint32_t* __restrict__ vals_out = route_pick(R.v, dst);
  uint32_t* __restrict__ vals2_out = route_pick(R.w, dst);
  const KeyRewrite rewrite = key_rewrite<kPlan>(R);
  const long long n = item_count(n_cap, n_dev);
  // a whole block past the items leaves (uniform), but for block 0 when it
  // writes the kept total
  const bool past = (long long)blockIdx.x * kItems >= n;
  if (past && !(kept_total != nullptr && chunk_sums != nullptr && blockIdx.x == 0)) return;
  __shared__ uint32_t s_cnt[kWaves][kWords];  // per-wave digit counts, then per-wave digit offsets (16-bit halves)
  __shared__ uint32_t s_cw[kWaves];           // wave totals of the chunk-sum scan
  __shared__ uint32_t s_start[kRadix];        // block-local start of each digit, then global start less it
  __shared__ uint32_t s_count;                // items of the block kept
  __shared__ uint32_t s_keys[kItems];
  __shared__ uint32_t s_vals[kItems];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t mask = (1u << bits) - 1u;
  const long long base = (long long)blockIdx.x * kItems;
  const long long wbase = base + (long long)wave * kRsWaveItems;
  for (int c = tid; c < kWords; c += kThreads)
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][c] = 0;
  // the global start of this thread's digits: the digit table's scan, plus,
  // when it spans several 4096-entry chunks, the sum of the chunks before
  // (an exclusive scan of the at most kScanTile <= kItems chunk totals, in
  // s_keys, which is free until the block's items are placed)
  uint32_t gstart[kPer];
  if (chunk_sums) {  // uniform
    constexpr int kPerT = kScanTile / kThreads;  // chunk totals per thread
    static_assert(kScanTile % kThreads == 0 && kScanTile <= kItems, "chunk totals fit s_keys");
    const int nchunks = (int)(((long long)kRadix * nblocks + kScanTile - 1) / kScanTile);
    uint32_t v[kPerT], sum = 0;
#pragma unroll
    for (int k = 0; k < kPerT; ++k) {
      const int c = tid * kPerT + k;
      v[k] = c < nchunks ? chunk_sums[c] : 0u;
      sum += v[k];
    }
    uint32_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) s_cw[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t t = s_cw[w];
      if (w < wave) before += t;
      all += t;
    }
    uint32_t run = x - sum + before;
#pragma unroll
    for (int k = 0; k < kPerT; ++k) {
      s_keys[tid * kPerT + k] = run;
      run += v[k];
    }
    if (kept_total && blockIdx.x == 0 && tid == 0) *kept_total = all;
    __syncthreads();
    if (past) return;  // block 0 with no items of its own: the total is written (uniform)
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const long long e = (long long)(tid * kPer + i) * nblocks + blockIdx.x;
      gstart[i] = hist[e] + s_keys[e / kScanTile];
    }
  } else {
#pragma unroll
    for (int i = 0; i < kPer; ++i) gstart[i] = hist[(long long)(tid * kPer + i) * nblocks + blockIdx.x];
  }
  __syncthreads();
  const uint64_t below = __lanemask_lt();
  uint32_t k[kRsRounds], v[kRsRounds], v2[kV2 ? kRsRounds : 1], rank[kRsRounds];
  int leader[kRsRounds];
  uint32_t old[kRsRounds];
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const long long j = wbase + (long long)r * 64 + lane;
    k[r] = j < n ? rewrite(keys[j], j) : kDropKey;
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
    if (valid && (peers & below) == 0ull)
      old[r] = half_of(atomicAdd(&s_cnt[wave][d >> 1], (uint32_t)__popcll(peers) << (16u * (d & 1u))), d);
  }
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const uint32_t base = (uint32_t)__shfl((int)old[r], leader[r], 64);
    if (rank[r] != 0xFFFFFFFFu) rank[r] += base;
  }
  __syncthreads();
  // per digit: offsets of each wave inside the digit, and the digit's total;
  // thread t takes digits kPer * t .. kPer * t + kPer - 1
  uint32_t tot[kPer], sum = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) tot[i] = 0;
  if (kPer == 2) {  // word t holds both of this thread's digits
    uint32_t run = 0;  // both halves at once: each stays below 65536
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_cnt[w][tid];
      s_cnt[w][tid] = run;
      run += c;
    }
    tot[0] = run & 0xFFFFu;
    tot[kPer - 1] = run >> 16;
  } else if ((tid & 1) == 0) {  // the even thread of a pair updates the shared word
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_cnt[w][tid >> 1];
      s_cnt[w][tid >> 1] = run;
      run += c;
    }
    tot[0] = run & 0xFFFFu;
    s_start[tid + 1] = run >> 16;  // the odd digit's total, handed over below
  }
  if (kPer == 1) {
    __syncthreads();
    if (tid & 1) tot[0] = s_start[tid];
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) sum += tot[i];
  // block-wide exclusive scan of the digit totals in digit order
  {
    uint32_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    __shared__ uint32_t s_wsum[kWaves];
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += s_wsum[w];
    uint32_t run = x - sum + add;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      s_start[tid * kPer + i] = run;
      run += tot[i];
    }
    if (tid == kThreads - 1) s_count = x + add;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    if (rank[r] == 0xFFFFFFFFu) continue;
    const uint32_t d = (k[r] >> shift) & mask;
    rank[r] += s_start[d] + half_of(s_cnt[wave][d >> 1], d);  // from here: the item's place in the block
    s_keys[rank[r]] = k[r];
    s_vals[rank[r]] = v[r];
  }
  __syncthreads();
  // from here s_start[d] is the digit's global start less its block start
#pragma unroll
  for (int i = 0; i < kPer; ++i) s_start[tid * kPer + i] = gstart[i] - s_start[tid * kPer + i];
  __syncthreads();
  const int count = (int)s_count;
  for (int p = tid; p < count; p += kThreads) {
    const uint32_t kk = s_keys[p];
    const uint32_t dst = s_start[(kk >> shift) & mask] + (uint32_t)p;
    keys_out[dst] = kk;
    vals_out[dst] = s_vals[p];
  }
  if constexpr (kV2) {
    __syncthreads();  // every value is out of s_vals: it takes the second values
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r)
      if (rank[r] != 0xFFFFFFFFu) s_vals[rank[r]] = v2[r];
    __syncthreads();
    for (int p = tid; p < count; p += kThreads)
      vals2_out[s_start[(s_keys[p] >> shift) & mask] + (uint32_t)p] = s_vals[p];
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

long long sort_hist_elems(long long max_n) { return 512LL * ((max_n + kScanTile - 1) / kScanTile + 1); }

hipError_t launch_exclusive_scan(const uint32_t* in, const uint32_t* gather_idx, uint32_t* out, long long n,
                                 uint32_t* total, const ScanWorkspace& ws, hipStream_t stream, int xform,
                                 const uint32_t* n_dev, uint32_t* side_total, int side_xform) {
  if (n <= 0) return total ? launch_fill_u32(total, 0u, 1, stream) : hipSuccess;
  if (n > ws.max_n) return hipErrorInvalidValue;
  return scan_level(in, gather_idx, out, n, n_dev, xform, total, side_total, side_xform, ws.buf, stream);
}

// An in-place exclusive scan of data[0..n) whose readers finish it
// themselves: when n spans 2 .. kScanTile chunks of kScanTile, only each
// chunk is scanned and *chunk_sums gets the chunks' totals (unscanned; the
// scan of element i is data[i] + the sum of the totals before chunk
// i / kScanTile, and their sum is the total, which the readers then write).
// Otherwise the full scan, *total written and *chunk_sums = null. With n_dev
// (device), the scan covers min(n, *n_dev) items (a pass a sort plan skips
// gives 0: nothing is read or written).
static hipError_t scan_no_add(uint32_t* data, long long n, uint32_t* total, const ScanWorkspace& ws,
                              const uint32_t** chunk_sums, hipStream_t stream, const uint32_t* n_dev = nullptr) {
  *chunk_sums = nullptr;
  const long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb <= 1 || nb > kScanTile || n > ws.max_n)
    return launch_exclusive_scan(data, nullptr, data, n, total, ws, stream, kScanPlain, n_dev);
  uint32_t* sums = ws.buf;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, data, nullptr, data, n, n_dev,
                     (int)kScanPlain, sums, nullptr, nullptr, (int)kScanPlain, nullptr);
  *chunk_sums = sums;
  return hipGetLastError();
}

namespace {

template <bool kPlan, int kDigitBits>
auto hist_kernel(bool aggregate) {
  return aggregate ? k_radix_hist<true, kPlan, kDigitBits> : k_radix_hist<false, kPlan, kDigitBits>;
}
template <bool kPlan, int kDigitBits>
auto scatter_kernel(bool v2) {
  return v2 ? k_radix_scatter<true, kPlan, kDigitBits> : k_radix_scatter<false, kPlan, kDigitBits>;
}

}  // namespace

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, long long n,
                             int begin_bit, int end_bit, const SortWorkspace& ws, bool* result_in_alt,
                             hipStream_t stream, const uint32_t* n_dev, uint32_t* vals2, uint32_t* vals2_alt,
                             bool identity_vals, uint32_t* n_kept, const SortPlan* plan) {
  *result_in_alt = false;
  if (end_bit <= begin_bit || n <= 0) return hipSuccess;
  const int digit = plan ? plan->digit_bits : 8;
  if (digit != 8 && digit != 9) return hipErrorInvalidValue;
  const int threads = digit == 9 ? RadixShape<9>::kThreads : RadixShape<8>::kThreads;
  const long long items = digit == 9 ? RadixShape<9>::kItems : RadixShape<8>::kItems;
  const long long nb = (n + items - 1) / items;
  if ((1LL << digit) * nb > ws.hist_capacity) return hipErrorInvalidValue;
  if (plan && (begin_bit != 0 || (plan->keys_c && (plan->vals2_c == nullptr) != (vals2 == nullptr))))
    return hipErrorInvalidValue;
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt, *wi = vals2, *wo = vals2_alt;
  int pass = 0;
  for (int shift = begin_bit; shift < end_bit; shift += digit, ++pass) {
    const int bits = (end_bit - shift) < digit ? (end_bit - shift) : digit;
    const bool first = shift == begin_bit;
    // with n_kept, the first pass drops kDropKey and the later ones sort what it kept
    const int drop = n_kept && first ? 1 : 0;
    const uint32_t* count = n_kept && !first ? n_kept : n_dev;
    SortRoute R{};
    if (plan) {
      const uint32_t size = plan->view_size ? plan->view_size : 1u;
      R = SortRoute{{keys, keys_alt, plan->keys_c}, {vals, vals_alt, plan->vals_c},
                    {vals2, vals2_alt, plan->vals2_c}, plan->passes, plan->key_base, plan->view_shift,
                    size, 1.f / (float)size, pass};
    } else {
      R = SortRoute{{ki, ko, nullptr}, {vi, vo, nullptr}, {wi, wo, nullptr}, nullptr, nullptr, nullptr, 1u, 1.f, pass};
    }
    const bool aggregate = bits <= 4 || shift + bits == 32;
    auto hk = plan ? (digit == 9 ? hist_kernel<true, 9>(aggregate) : hist_kernel<true, 8>(aggregate))
                   : hist_kernel<false, 8>(aggregate);
    hipLaunchKernelGGL(hk, dim3((unsigned)nb), dim3(threads), 0, stream, R, n, count, shift, bits, drop, ws.hist,
                       (int)nb);
    SPLAT_HIP_CHECK(hipGetLastError());
    const uint32_t* prefix = nullptr;
    SPLAT_HIP_CHECK(scan_no_add(ws.hist, (1LL << digit) * nb, drop ? n_kept : nullptr, ws.scan, &prefix, stream,
                                plan ? plan->active + pass : nullptr));
    const int identity = (identity_vals && first) ? 1 : 0;
    auto sk = plan ? (digit == 9 ? scatter_kernel<true, 9>(vals2 != nullptr) : scatter_kernel<true, 8>(vals2 != nullptr))
                   : scatter_kernel<false, 8>(vals2 != nullptr);
    // the dropping pass's kept total: written by the full scan, or with
    // chunk totals by the scatter's block 0
    uint32_t* kept_total = drop && prefix ? n_kept : nullptr;
    hipLaunchKernelGGL(sk, dim3((unsigned)nb), dim3(threads), 0, stream, R, identity, n, count, shift, bits, drop,
                       ws.hist, prefix, kept_total, (int)nb);
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

// The plans of the depth and super-tile sorts from the projection's per-wave
// ranges of kept keys (layout: the header's kPlan* words). Every kept key
// less the minimum fits span_bits bits. When view_bits >= 0 and the view id
// fits above them (span_bits + view_bits <= 31, so no key becomes the drop
// key), the depth sort also orders by view, above depth: its output, and the
// super-tile pairs emitted in that order, are grouped by view, and one pass
// over the 8-bit super-tile id within a view finishes the super-tile sort;
// otherwise the super-tile sort takes a second pass, by view. The depth sort
// runs 9-bit digits, as many passes as the key's bits need (at least 2).
__global__ __launch_bounds__(1024) void k_sort_plan(const uint32_t* __restrict__ key_ranges, int n_waves,
                                                    int view_bits, uint32_t* __restrict__ plan) {
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
  uint32_t base = 0;
  int span_bits = 0;
  if (hi >= lo) {  // some key was kept
    base = lo;
    const uint32_t span = hi - lo;
    span_bits = span ? 32 - __clz((int)span) : 0;
  }
  const bool grouped = view_bits >= 0 && span_bits + view_bits <= 31;
  const int key_bits = span_bits + (grouped ? view_bits : 0);
  const uint32_t passes = (uint32_t)max(2, (key_bits + kDepthDigitBits - 1) / kDepthDigitBits);
  plan[kPlanPasses] = passes;
  for (int i = 0; i < 4; ++i) plan[kPlanActive + i] = (uint32_t)i < passes ? 0xFFFFFFFFu : 0u;
  plan[kPlanKeyBase] = base;
  plan[kPlanViewShift] = grouped ? (uint32_t)span_bits : 32u;
  const uint32_t super_passes = grouped || view_bits <= 0 ? 1u : 2u;
  plan[kPlanSuperPasses] = super_passes;
  for (int i = 0; i < 2; ++i) plan[kPlanSuperActive + i] = (uint32_t)i < super_passes ? 0xFFFFFFFFu : 0u;
}

}  // namespace

hipError_t launch_sort_plan(const uint32_t* key_ranges, int n_waves, int view_bits, uint32_t* plan,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(1024), 0, stream, key_ranges, n_waves, view_bits, plan);
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
