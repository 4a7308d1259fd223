// Two-level tile binning: per-tile lists without a sort over (tile, splat)
// pairs.
//
// The lists are the engine's buildTiles (engine/native/cpp/src/splat/
// SplatRaster.cpp:114-151): every projected splat enters each 16x16 tile of its
// rect, and each tile's list runs near to far (the depth order of :260-261,
// exact ties in gaussian index order).
//
// At 1080p a splat covers ~14 tiles, so sorting the (tile, splat) pairs moves
// ~14 times more data than there are splats. Here only (super-tile, splat)
// pairs are sorted -- a super-tile is 8x8 tiles, and most splats touch one or
// two -- and the per-tile lists are then cut out of each super-tile's
// depth-ordered list directly:
//   1. emit:    for every depth-sorted slot, one (super-tile id, slot) pair per
//               super-tile its rect touches (stable radix sort by id next);
//   2. segment: each super-tile's list is cut into segments of kBinSegment (256) splats;
//   3. count:   one wave per segment; per chunk of 64 splats, eight row and
//               eight column ballots (from the rect bits the emitter packed
//               into the key) give each tile's covering lanes as one AND, and
//               lane t counts tile t's with a popcount;
//   4. scan:    an exclusive scan over the (super-tile, tile, segment) table
//               gives every (tile, segment) run its place, and the tile ranges;
//   5. write:   the same wave pass again, writing each splat into each covered
//               tile's run at its rank among the covering lanes (mbcnt), in
//               depth order.
// Every tile's list thus lands contiguously and in the order of the super-tile
// list, which is depth order; pair memory is 4 bytes per (tile, splat).
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

__device__ __forceinline__ int find_view_st(const ViewDesc* views, int n_views, int st) {
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].st_base <= st) v = k;
  return v;
}

// Super-tile pair emission, one wave per 64 consecutive depth-sorted slots.
// The wave's pairs are one contiguous range of the output (soffsets is an
// exclusive scan in the same order, *s_total its sum), so the lanes stride
// over that range together and each lane finds the slot owning its pair by a
// binary search over the wave's 64 offsets. A key is the global super-tile id
// in its low kStBits bits and the splat's rect inside that super-tile (three
// bits per bound) above them, where the sort by id carries it along. With
// pair_base, also records each slot's first tile-pair index, the
// deterministic adjoint's pair id base. The rects are read from the compact
// 8-byte-a-slot copy the projection writes: in depth order these reads are
// random, and the compact array stays in the last-level cache where the
// 64-byte records would not.
__global__ __launch_bounds__(256) void k_emit_super(const uint32_t* __restrict__ sorted_slots,
                                                    const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ soffsets,
                                                    const uint32_t* __restrict__ s_total,
                                                    const uint2* __restrict__ rects,
                                                    const uint32_t* __restrict__ sorted_prects,
                                                    const ViewDesc* __restrict__ views, long long per_view,
                                                    long long m_cap, const uint32_t* __restrict__ m_dev,
                                                    uint32_t s_cap, uint32_t* __restrict__ skeys_even,
                                                    uint32_t* __restrict__ svals_even,
                                                    uint32_t* __restrict__ skeys_odd,
                                                    uint32_t* __restrict__ svals_odd,
                                                    const uint32_t* __restrict__ sort_passes,
                                                    uint32_t* __restrict__ pair_base) {
  // the sort after this one ends in the even set whatever its pass count:
  // its input goes to the odd set when that count is odd
  const bool odd = sort_passes && (*sort_passes & 1u);
  uint32_t* __restrict__ skeys = odd ? skeys_odd : skeys_even;
  uint32_t* __restrict__ svals = odd ? svals_odd : svals_even;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long m = m_dev ? min(m_cap, (long long)*m_dev) : m_cap;
  uint32_t slot = 0, c = 0, off = 0xFFFFFFFFu;
  int x0 = 0, x1 = 0, y0 = 0, y1 = 0, w = 1, scols = 0, st_base = 0;
  if (s < m) {
    slot = sorted_slots[s];
    off = soffsets[s];
    c = (s + 1 < m ? soffsets[s + 1] : *s_total) - off;
    if (c && sorted_prects) {  // the packed rect rode along with the depth sort
      const uint32_t r = sorted_prects[s];
      x0 = rect_x0(r);
      x1 = rect_x1(r);
      y0 = rect_y0(r);
      y1 = rect_y1(r);
    } else if (c) {
      const uint2 packed = rects[slot];  // x0, x1, y0, y1 as int16
      x0 = (int)(int16_t)(packed.x & 0xFFFF);
      x1 = (int)(int16_t)(packed.x >> 16);
      y0 = (int)(int16_t)(packed.y & 0xFFFF);
      y1 = (int)(int16_t)(packed.y >> 16);
    }
    if (c) {
      const int v = (int)(slot / per_view);
      w = x1 / kSuper - x0 / kSuper + 1;
      scols = views[v].scols;
      st_base = views[v].st_base;
      if (pair_base) pair_base[slot] = offsets[s];
    }
  }
  const uint32_t first = __shfl(off, 0, 64);
  if (first == 0xFFFFFFFFu) return;  // the whole wave is past the end
  const uint64_t valid = __ballot(s < m);
  const int last_lane = 63 - __builtin_clzll(valid);
  const uint32_t stop = __shfl(off + c, last_lane, 64);
  // The trip count is wave-uniform, so all lanes take part in every shuffle;
  // only the stores are predicated.
  for (uint32_t chunk = first; chunk < stop; chunk += 64) {
    const uint32_t i = chunk + (uint32_t)lane;
    int owner = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      const int cand = owner + step;
      const uint32_t oc = __shfl(off, cand & 63, 64);
      if (cand <= last_lane && oc <= i) owner = cand;
    }
    const uint32_t o_off = __shfl(off, owner, 64);
    const int ow = __shfl(w, owner, 64);
    const int ox0 = __shfl(x0, owner, 64), ox1 = __shfl(x1, owner, 64);
    const int oy0 = __shfl(y0, owner, 64), oy1 = __shfl(y1, owner, 64);
    const int obase = __shfl(st_base, owner, 64);
    const int ocols = __shfl(scols, owner, 64);
    const uint32_t oslot = __shfl(slot, owner, 64);
    if (i < stop && i < s_cap) {  // past s_cap: the capacity check has flagged the step
      const uint32_t local = i - o_off;
      const int sy = oy0 / kSuper + (int)(local / (uint32_t)ow);
      const int sx = ox0 / kSuper + (int)(local % (uint32_t)ow);
      const int lx0 = max(ox0 - sx * kSuper, 0), lx1 = min(ox1 - sx * kSuper, kSuper - 1);
      const int ly0 = max(oy0 - sy * kSuper, 0), ly1 = min(oy1 - sy * kSuper, kSuper - 1);
      skeys[i] = (uint32_t)(obase + sy * ocols + sx) | ((uint32_t)lx0 << kStBits) | ((uint32_t)lx1 << (kStBits + 3)) |
                 ((uint32_t)ly0 << (kStBits + 6)) | ((uint32_t)ly1 << (kStBits + 9));
      svals[i] = oslot;
    }
  }
}

// Four sorted keys a thread (one 16-byte load, the neighbours from the lanes
// beside it): a super-tile's list starts where its id first appears and
// ends after its last.
__global__ __launch_bounds__(256) void k_super_ranges(const uint32_t* __restrict__ keys, long long n_cap,
                                                      const uint32_t* __restrict__ n_dev,
                                                      uint32_t* __restrict__ ranges, int n_st) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // keys 4q .. 4q + 3
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long i0 = 4 * q;
  if (i0 >= n) return;
  uint32_t k[6];  // keys i0 - 1 .. i0 + 4 (ids; past the ends: none)
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  if (i0 + 4 <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
    k[1] = v.x & kStMask; k[2] = v.y & kStMask; k[3] = v.z & kStMask; k[4] = v.w & kStMask;
  } else {
    for (int a = 0; a < 4; ++a) k[1 + a] = i0 + a < n ? keys[i0 + a] & kStMask : kNone;
  }
  k[0] = i0 > 0 ? keys[i0 - 1] & kStMask : kNone;
  k[5] = i0 + 4 < n ? keys[i0 + 4] & kStMask : kNone;
#pragma unroll
  for (int a = 1; a <= 4; ++a) {
    const uint32_t id = k[a];
    if (id >= (uint32_t)n_st) continue;  // past the end (or an id the emitter cannot write)
    if (k[a - 1] != id) ranges[2 * (long long)id] = (uint32_t)(i0 + a - 1);
    if (k[a + 1] != id) ranges[2 * (long long)id + 1] = (uint32_t)(i0 + a);
  }
}

__global__ __launch_bounds__(256) void k_segment_counts(const uint32_t* __restrict__ sranges, int n_st,
                                                        uint32_t* __restrict__ nseg) {
  const int st = blockIdx.x * blockDim.x + threadIdx.x;
  if (st >= n_st) return;
  const uint32_t len = sranges[2 * st + 1] - sranges[2 * st];
  nseg[st] = (len + kBinSegment - 1) / kBinSegment;
}

__global__ __launch_bounds__(256) void k_segment_list(const uint32_t* __restrict__ nseg,
                                                      const uint32_t* __restrict__ seg_base, int n_st,
                                                      uint32_t* __restrict__ seg_st) {
  const int st = blockIdx.x * blockDim.x + threadIdx.x;
  if (st >= n_st) return;
  const uint32_t b = seg_base[st], c = nseg[st];
  for (uint32_t s = 0; s < c; ++s) seg_st[b + s] = (uint32_t)st;
}

// One segment as seen by one wave: which super-tile, which slice of its list,
// and where its table entries start.
struct Segment {
  uint32_t st, s, nsg, b0, b1;
  uint32_t table0;  // 64 * seg_base[st]
};

__device__ __forceinline__ Segment segment_of(uint32_t g, const uint32_t* sranges, const uint32_t* nseg,
                                              const uint32_t* seg_base, const uint32_t* seg_st) {
  Segment S;
  S.st = seg_st[g];
  S.s = g - seg_base[S.st];
  S.nsg = nseg[S.st];
  S.table0 = 64u * seg_base[S.st];
  const uint32_t b = sranges[2 * S.st], e = sranges[2 * S.st + 1];
  S.b0 = b + S.s * kBinSegment;
  S.b1 = min(S.b0 + (uint32_t)kBinSegment, e);
  return S;
}

// The coverage ballots of one chunk of 64 list entries: bit j of rows[k] is
// set when lane j's rect (from the rect bits of its super-tile key) spans
// tile row k of the super-tile, bit j of cols[k] when it spans tile column k.
// Tile (lx, ly) is covered by exactly the lanes of rows[ly] & cols[lx].
struct ChunkCover {
  uint64_t rows[kSuper], cols[kSuper];
};

__device__ __forceinline__ ChunkCover chunk_cover(bool valid, uint32_t key) {
  const int lx0 = (key >> kStBits) & 7, lx1 = (key >> (kStBits + 3)) & 7;
  const int ly0 = (key >> (kStBits + 6)) & 7, ly1 = (key >> (kStBits + 9)) & 7;
  ChunkCover c;
#pragma unroll
  for (int k = 0; k < kSuper; ++k) {
    c.rows[k] = __ballot(valid && ly0 <= k && k <= ly1);
    c.cols[k] = __ballot(valid && lx0 <= k && k <= lx1);
  }
  return c;
}

// Lane t: how many lanes of the chunk cover tile t = ly * 8 + lx.
__device__ __forceinline__ uint32_t lane_tile_count(const ChunkCover& c, int lane) {
  const int ly = lane >> 3, lx = lane & 7;
  uint64_t r = 0ull, q = 0ull;
#pragma unroll
  for (int k = 0; k < kSuper; ++k) {
    r = ly == k ? c.rows[k] : r;
    q = lx == k ? c.cols[k] : q;
  }
  return (uint32_t)__popcll(r & q);
}

// Per (tile, segment) splat counts into the table at
// 64 * seg_base[st] + tile * nseg[st] + s; entries of segments past the last
// real one (up to ns_max) are zeroed. One wave per segment, grid-stride.
__global__ __launch_bounds__(256) void k_tile_count(const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ sranges,
                                                    const uint32_t* __restrict__ nseg,
                                                    const uint32_t* __restrict__ seg_base,
                                                    const uint32_t* __restrict__ seg_st,
                                                    const uint32_t* __restrict__ ns_total, uint32_t ns_max,
                                                    uint32_t* __restrict__ table) {
  const int lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t ns = *ns_total;
  for (uint32_t g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); g < ns_max; g += waves) {
    if (g >= ns) {
      table[64ull * g + lane] = 0u;
      continue;
    }
    const Segment S = segment_of(g, sranges, nseg, seg_base, seg_st);
    uint32_t cnt = 0;  // lane t: splats of this segment covering tile t
    for (uint32_t c = S.b0; c < S.b1; c += 64) {
      const uint32_t i = c + (uint32_t)lane;
      const bool valid = i < S.b1;
      cnt += lane_tile_count(chunk_cover(valid, valid ? skeys[i] : 0u), lane);
    }
    table[S.table0 + (uint64_t)lane * S.nsg + S.s] = cnt;
  }
}

// Writes every splat of every segment into each covered tile's run. Same
// segments as k_tile_count; `scan` is the exclusive scan of its table. The
// wave holds the segment's (up to four) chunks of 64 list entries in
// registers, each as its slot and a 64-bit mask of the super-tile's tiles its
// rect covers, and walks the tiles in the outer loop: per tile and chunk, the
// covering lanes store at the tile's next slot plus their rank among those
// lanes (mbcnt), so each run is in list order and each tile's whole run for
// the segment goes out as one burst of adjacent stores. Both loops are
// unrolled: a lane's coverage of a tile is one constant-bit test, and its
// ballot is the store's execution mask. Entries at or past p_cap (a step the
// capacity check flags) all land on the spare entry pair_vals[p_cap]. With
// kNarrow (4 (p_cap + 1) < 2^32), the stores take 32-bit byte offsets from
// the list's base.
constexpr int kSegChunks = kBinSegment / 64;

// One segment's runs, tile by tile: the covering lanes of each chunk place
// their slots at their rank in the wave's LDS slice (the chunk's ballot is
// the store's execution mask), then the tile's whole run for the segment goes
// to its place in one coalesced store per 64 entries. Tiles alternate between
// two LDS buffers, so a tile's placement needs no wait for the previous
// tile's read; the LDS keeps one wave's accesses in order, and the wave
// barriers keep the compiler from moving them across each other.
template <bool kNarrow, bool kBound>
__device__ __forceinline__ void write_segment_runs(uint32_t base, const uint32_t (&slot)[kSegChunks],
                                                   const uint32_t (&cov)[kSegChunks][2], uint32_t p_cap,
                                                   uint32_t* __restrict__ pair_vals, uint32_t* __restrict__ s_runs) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 64; ++t) {
    uint32_t* s_run = s_runs + (t & 1) * kBinSegment;
    const uint32_t at = (uint32_t)__builtin_amdgcn_readlane((int)base, t);
    uint32_t n = 0;  // entries of the tile's run so far
#pragma unroll
    for (int c = 0; c < kSegChunks; ++c) {
      const uint64_t b = __ballot((cov[c][t >> 5] >> (t & 31)) & 1u);
      const uint32_t k = n + (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
      if (__builtin_amdgcn_inverse_ballot_w64(b)) s_run[k] = slot[c];
      n += (uint32_t)__popcll(b);
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = lane; k < n; k += 64) {
      uint32_t to = at + k;
      if constexpr (kBound) to = min(to, p_cap);
      if constexpr (kNarrow)
        *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(pair_vals) + (uint32_t)(to << 2)) = s_run[k];
      else
        pair_vals[to] = s_run[k];
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <bool kNarrow>
__global__ __launch_bounds__(256) void k_tile_write(const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals,
                                                    const uint32_t* __restrict__ sranges,
                                                    const uint32_t* __restrict__ nseg,
                                                    const uint32_t* __restrict__ seg_base,
                                                    const uint32_t* __restrict__ seg_st,
                                                    const uint32_t* __restrict__ ns_total,
                                                    const uint32_t* __restrict__ scan, uint32_t p_cap,
                                                    uint32_t* __restrict__ pair_vals) {
  static_assert(kBinSegment % 64 == 0 && kSuper == 8, "segments of whole chunks, 8x8 super-tiles");
  __shared__ uint32_t s_runs[4][2 * kBinSegment];  // per wave: two tiles' runs of the segment
  const int lane = threadIdx.x & 63;
  uint32_t* s_run = s_runs[threadIdx.x >> 6];
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t ns = *ns_total;
  for (uint32_t g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); g < ns; g += waves) {
    const Segment S = segment_of(g, sranges, nseg, seg_base, seg_st);
    const uint32_t base = scan[S.table0 + (uint64_t)lane * S.nsg + S.s];  // lane t: tile t's first slot
    uint32_t slot[kSegChunks], cov[kSegChunks][2];  // cov: tiles (ty * 8 + tx) the rect covers
#pragma unroll
    for (int c = 0; c < kSegChunks; ++c) {
      const uint32_t i = S.b0 + 64u * c + (uint32_t)lane;
      const bool valid = i < S.b1;
      const uint32_t key = valid ? skeys[i] : 0u;
      slot[c] = valid ? svals[i] : 0u;
      const uint32_t lx0 = (key >> kStBits) & 7, lx1 = (key >> (kStBits + 3)) & 7;
      const uint32_t ly0 = (key >> (kStBits + 6)) & 7, ly1 = (key >> (kStBits + 9)) & 7;
      const uint64_t cols = ((2ull << lx1) - 1ull) & ~((1ull << lx0) - 1ull);                    // bits lx0..lx1
      const uint64_t rows = ((2ull << (8 * ly1 + 7)) - 1ull) & ~((1ull << (8 * ly0)) - 1ull);  // bytes ly0..ly1
      const uint64_t m = valid ? (cols * 0x0101010101010101ull) & rows : 0ull;
      cov[c][0] = (uint32_t)m;
      cov[c][1] = (uint32_t)(m >> 32);
    }
    // the segment's runs all end by the end of tile 63's run: when that is
    // within p_cap (every step the capacity check passes), no store needs a
    // bound
    const uint32_t seg_end = scan[S.table0 + 63ull * S.nsg + S.s + 1];
    if (seg_end <= p_cap)
      write_segment_runs<kNarrow, false>(base, slot, cov, p_cap, pair_vals, s_run);
    else
      write_segment_runs<kNarrow, true>(base, slot, cov, p_cap, pair_vals, s_run);
  }
}

// [begin, end) of every global tile from the scanned table.
__global__ __launch_bounds__(256) void k_tile_ranges_table(const ViewDesc* __restrict__ views, int n_views,
                                                           int total_tiles, const uint32_t* __restrict__ scan,
                                                           const uint32_t* __restrict__ nseg,
                                                           const uint32_t* __restrict__ seg_base, uint32_t p_cap,
                                                           uint32_t* __restrict__ ranges) {
  const int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= total_tiles) return;
  int v = 0;
  for (int k = 1; k < n_views; ++k)
    if (views[k].tile_base <= tile) v = k;
  const ViewDesc& V = views[v];
  const int local = tile - V.tile_base;
  const int tx = local % V.cols, ty = local / V.cols;
  const int st = V.st_base + (ty / kSuper) * V.scols + tx / kSuper;
  const int t = (ty % kSuper) * kSuper + tx % kSuper;
  const uint64_t i0 = 64ull * seg_base[st] + (uint64_t)t * nseg[st];
  ranges[2 * (long long)tile] = min(scan[i0], p_cap);
  ranges[2 * (long long)tile + 1] = min(scan[i0 + nseg[st]], p_cap);
}

// totals[0] = tile pairs, totals[1] = super-tile pairs; totals[3] = 1 when
// either exceeds the buffers the step was sized for.
__global__ void k_check_caps(uint32_t* totals, uint32_t p_cap, uint32_t s_cap) {
  if (threadIdx.x == 0 && blockIdx.x == 0) totals[3] = (totals[0] > p_cap || totals[1] > s_cap) ? 1u : 0u;
}

}  // namespace

hipError_t launch_check_caps(uint32_t* totals, uint32_t p_cap, uint32_t s_cap, hipStream_t stream) {
  hipLaunchKernelGGL(k_check_caps, dim3(1), dim3(64), 0, stream, totals, p_cap, s_cap);
  return hipGetLastError();
}

hipError_t launch_emit_super(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* soffsets,
                             const uint32_t* s_total, const uint2* rects, const uint32_t* sorted_prects,
                             const ViewDesc* views, long long per_view, long long m, const uint32_t* m_dev,
                             uint32_t s_cap, uint32_t* skeys, uint32_t* svals, uint32_t* skeys_odd,
                             uint32_t* svals_odd, const uint32_t* sort_passes, uint32_t* pair_base,
                             hipStream_t stream) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit_super, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, sorted_slots, offsets,
                     soffsets, s_total, rects, sorted_prects, views, per_view, m, m_dev, s_cap, skeys, svals,
                     skeys_odd, svals_odd, sort_passes, pair_base);
  return hipGetLastError();
}

hipError_t launch_super_ranges(const uint32_t* sorted_keys, long long n, const uint32_t* n_dev, uint32_t* sranges,
                               int n_st, hipStream_t stream) {
  SPLAT_HIP_CHECK(launch_fill_u32(sranges, 0u, 2LL * n_st, stream));
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_super_ranges, dim3((unsigned)(((n + 3) / 4 + 255) / 256)), dim3(256), 0, stream, sorted_keys, n,
                     n_dev, sranges, n_st);
  return hipGetLastError();
}

hipError_t launch_segment_counts(const uint32_t* sranges, int n_st, uint32_t* nseg, hipStream_t stream) {
  if (n_st <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_segment_counts, dim3((unsigned)((n_st + 255) / 256)), dim3(256), 0, stream, sranges, n_st,
                     nseg);
  return hipGetLastError();
}

hipError_t launch_segment_list(const uint32_t* nseg, const uint32_t* seg_base, int n_st, uint32_t* seg_st,
                               hipStream_t stream) {
  if (n_st <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_segment_list, dim3((unsigned)((n_st + 255) / 256)), dim3(256), 0, stream, nseg, seg_base,
                     n_st, seg_st);
  return hipGetLastError();
}

namespace {
unsigned segment_grid(uint32_t ns_max) {
  // four waves (segments) per workgroup; enough workgroups to fill the device
  // several times over, grid-stride beyond that
  const uint32_t want = (ns_max + 3) / 4;
  return want < 1 ? 1u : (want > 8192u ? 8192u : want);
}
}  // namespace

hipError_t launch_tile_count(const uint32_t* skeys, const uint32_t* sranges,
                             const uint32_t* nseg, const uint32_t* seg_base, const uint32_t* seg_st,
                             const uint32_t* ns_total, uint32_t ns_max, uint32_t* table, hipStream_t stream) {
  // the entry past the last segment's table block holds the grand total after the scan
  SPLAT_HIP_CHECK(launch_fill_u32(table + 64ull * ns_max, 0u, 1, stream));
  if (ns_max == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_count, dim3(segment_grid(ns_max)), dim3(256), 0, stream, skeys, sranges,
                     nseg, seg_base, seg_st, ns_total, ns_max, table);
  return hipGetLastError();
}

hipError_t launch_tile_write(const uint32_t* skeys, const uint32_t* svals,
                             const uint32_t* sranges, const uint32_t* nseg, const uint32_t* seg_base,
                             const uint32_t* seg_st, const uint32_t* ns_total, uint32_t ns_max, const uint32_t* scan,
                             uint32_t p_cap, uint32_t* pair_vals, hipStream_t stream) {
  if (ns_max == 0) return hipSuccess;
  if (4ull * ((unsigned long long)p_cap + 1ull) <= 0xFFFFFFFFull)
    hipLaunchKernelGGL(k_tile_write<true>, dim3(segment_grid(ns_max)), dim3(256), 0, stream, skeys, svals, sranges,
                       nseg, seg_base, seg_st, ns_total, scan, p_cap, pair_vals);
  else
    hipLaunchKernelGGL(k_tile_write<false>, dim3(segment_grid(ns_max)), dim3(256), 0, stream, skeys, svals, sranges,
                       nseg, seg_base, seg_st, ns_total, scan, p_cap, pair_vals);
  return hipGetLastError();
}

hipError_t launch_tile_ranges_table(const ViewDesc* views, int n_views, int total_tiles, const uint32_t* scan,
                                    const uint32_t* nseg, const uint32_t* seg_base, uint32_t p_cap, uint32_t* ranges,
                                    hipStream_t stream) {
  if (total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_ranges_table, dim3((unsigned)((total_tiles + 255) / 256)), dim3(256), 0, stream, views,
                     n_views, total_tiles, scan, nseg, seg_base, p_cap, ranges);
  return hipGetLastError();
}

}  // namespace splat
