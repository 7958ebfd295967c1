// Stable LSD radix sort of (key, uint32 value) pairs, 8-bit digits.
//
// Per digit pass: a tile histogram kernel, an exclusive scan over the
// digit-major count table, and a scatter kernel that ranks each element
// inside its tile with wave ballots (peers with the same digit) and per-wave
// digit counts in LDS, so the relative order of equal digits is preserved.
// Everything is plain kernels (no memset nodes), so a caller can capture the
// sort in a HIP graph.
//
// Segmented form: given a table of tiles that never straddle a segment
// boundary, with the counts laid out segment-major ((segment, digit, tile of
// the segment)), the same exclusive scan yields each tile's output position
// inside its own segment's range, so one pass sorts every segment at once and
// the segments keep their places.

#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "pm_internal.hpp"

namespace pm {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per block
constexpr int kScanChunk = kSortThreads * kSortItems; // counts per scan block
static_assert(kSortTile == kSortTileElems, "tile size shared with the tile table");

// The block's tile: its elements [base, base + n) and where its digit counts
// go (cbase + digit * stride), from the table when there is one, else the
// uniform tiling (digit-major over all tiles).
struct TileView {
  int64_t base;
  int n;
  uint32_t cbase;
  uint32_t stride;
};
__device__ __forceinline__ TileView tile_view(const SortTile* tiles, int64_t n,
                                              int64_t n_tiles) {
  if (tiles) {
    const SortTile tl = tiles[blockIdx.x];
    return {tl.base, static_cast<int>(tl.n), tl.cbase, tl.stride};
  }
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
  return {base, static_cast<int>(std::min<int64_t>(kSortTile, n - base)), blockIdx.x,
          static_cast<uint32_t>(n_tiles)};
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void digit_hist_kernel(
    const K* __restrict__ keys, int64_t n, int shift, uint32_t* __restrict__ counts,
    int64_t n_tiles, const SortTile* __restrict__ tiles) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const TileView tv = tile_view(tiles, n, n_tiles);
  K k[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = r * kSortThreads + t;
    k[r] = li < tv.n ? keys[tv.base + li] : K(0);
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = r * kSortThreads + t;
    if (li < tv.n) atomicAdd(&h[static_cast<uint32_t>(k[r] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[static_cast<int64_t>(tv.cbase) + static_cast<int64_t>(t) * tv.stride] = h[t];
}

// Exclusive scan helpers over a uint32 array (length up to 2^31 chunks).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum,
                                                         uint32_t* total) {
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t off = 0;
  uint32_t tot = 0;
  for (int i = 0; i < kSortThreads / 64; ++i) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  *total = tot;
  __syncthreads();
  return off + x - v;
}

__global__ __launch_bounds__(kSortThreads) void scan_reduce_kernel(
    const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ partial) {
  __shared__ uint32_t wsum[kSortThreads / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  uint32_t s = 0;
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kSortItems + r;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, wsum, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block: exclusive scan of the partial sums, in place.
__global__ __launch_bounds__(kSortThreads) void scan_top_kernel(uint32_t* partial,
                                                                int64_t n) {
  __shared__ uint32_t wsum[kSortThreads / 64];
  uint32_t carry = 0;
  for (int64_t base = 0; base < n; base += kSortThreads) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < n ? partial[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, wsum, &tot);
    if (i < n) partial[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSortThreads) void scan_down_kernel(
    uint32_t* __restrict__ data, int64_t n, const uint32_t* __restrict__ partial) {
  __shared__ uint32_t wsum[kSortThreads / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk +
                       static_cast<int64_t>(threadIdx.x) * kSortItems;
  uint32_t v[kSortItems];
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + r;
    v[r] = i < n ? data[i] : 0u;
    s += v[r];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, wsum, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + r;
    if (i < n) data[i] = run;
    run += v[r];
  }
}

// Exclusive scan of one digit's row of a uniform (digit-major) count table
// in place, one block per digit: the row's entries become each tile's count
// of the digit in the tiles before it, and totals[digit] the digit's count
// over every tile.  The scatter adds the digits' starts, the exclusive scan of
// the totals, itself.
__global__ __launch_bounds__(kSortThreads) void digit_scan_kernel(uint32_t* __restrict__ counts,
                                                                  int64_t n_tiles,
                                                                  uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[kSortThreads / 64];
  uint32_t* row = counts + static_cast<int64_t>(blockIdx.x) * n_tiles;
  uint32_t carry = 0;
  for (int64_t b = 0; b < n_tiles; b += kScanChunk) {
    const int64_t i0 = b + static_cast<int64_t>(threadIdx.x) * kSortItems;
    uint32_t v[kSortItems];
    uint32_t sum = 0;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      v[r] = i0 + r < n_tiles ? row[i0 + r] : 0u;
      sum += v[r];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(sum, wsum, &tot) + carry;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      if (i0 + r < n_tiles) row[i0 + r] = run;
      run += v[r];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Scatter of one digit pass.  Wave w ranks the consecutive elements
// [w * 1024, (w + 1) * 1024) of the tile, 64 per round: wave ballots find
// each lane's equal-digit peers, and the wave's own running count of every
// digit (LDS, touched by no other wave) places them after the wave's earlier
// rounds.  One barrier later, each digit's counts over the waves give the
// waves' offsets within the digit and the digits' starts within the tile,
// and the tile is staged in LDS in digit order (stable: element order is
// wave, round, lane).  The staged tile is then written out with consecutive
// lanes on consecutive elements, so each digit's run of the tile lands in
// consecutive addresses instead of one scattered element per lane.
template <typename K>
__global__ __launch_bounds__(kSortThreads) void digit_scatter_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift,
    const uint32_t* __restrict__ offsets, int64_t n_tiles, const SortTile* __restrict__ tiles,
    const uint32_t* __restrict__ totals) {
  constexpr int kWaves = kSortThreads / 64;
  constexpr int kWaveElems = kSortItems * 64;
  __shared__ uint32_t glob[256];            // first output position of each digit (this tile)
  __shared__ uint32_t loc[256];             // first staged position of each digit
  __shared__ uint32_t wcnt[kWaves][256];    // a wave's running digit counts, then its offsets
  __shared__ uint32_t wsum[kWaves];
  __shared__ K skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const TileView tv = tile_view(tiles, n, n_tiles);
  const int64_t base = tv.base;
  const int tile_n = tv.n;
  // the tile's output start of digit t, from the exclusive scan of the
  // count table (per digit row, plus the digit's start, when totals are given)
  uint32_t g = offsets[static_cast<int64_t>(tv.cbase) + static_cast<int64_t>(t) * tv.stride];
  if (totals) {
    uint32_t all;
    g += block_exclusive_scan(totals[t], wsum, &all);
  }
  glob[t] = g;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) wcnt[i][t] = 0;
  // every element of the thread is loaded up front, so the tile's loads are
  // in flight together rather than one round at a time
  K keys_r[kSortItems];
  uint32_t vals_r[kSortItems];
  uint32_t pos_r[kSortItems];  // place among the wave's elements of its digit
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = w * kWaveElems + r * 64 + lane;
    keys_r[r] = 0;
    vals_r[r] = 0;
    pos_r[r] = 0;
    if (li < tile_n) {
      keys_r[r] = kin[base + li];
      vals_r[r] = vin[base + li];
    }
  }
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t* mine = wcnt[w];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = w * kWaveElems + r * 64 + lane;
    const bool valid = li < tile_n;
    const uint32_t d = valid ? static_cast<uint32_t>(keys_r[r] >> shift) & 255u : 0u;
    const uint64_t vmask = __ballot(valid);
    uint64_t peers = valid ? vmask : ~vmask;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt_mask));
    // every lane reads its digit's count before the group's first lane
    // (one instruction later) advances it
    const uint32_t before = valid ? mine[d] : 0u;
    pos_r[r] = before + rank;
    if (valid && rank == 0) mine[d] = before + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();
  // digit t: the waves' exclusive offsets (in place) and the tile's count
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    const uint32_t c = wcnt[i][t];
    wcnt[i][t] = cnt;
    cnt += c;
  }
  uint32_t tot;
  const uint32_t l0 = block_exclusive_scan(cnt, wsum, &tot);
  loc[t] = l0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = w * kWaveElems + r * 64 + lane;
    if (li < tile_n) {
      const uint32_t d = static_cast<uint32_t>(keys_r[r] >> shift) & 255u;
      const uint32_t p = loc[d] + wcnt[w][d] + pos_r[r];
      skey[p] = keys_r[r];
      sval[p] = vals_r[r];
    }
  }
  __syncthreads();
  for (int j = t; j < tile_n; j += kSortThreads) {
    const K key = skey[j];
    const uint32_t d = static_cast<uint32_t>(key >> shift) & 255u;
    const uint32_t pos = glob[d] + static_cast<uint32_t>(j) - loc[d];
    kout[pos] = key;
    vout[pos] = sval[j];
  }
}

int exclusive_scan_impl(uint32_t* partial, uint32_t* data, int64_t n, hipStream_t s) {
  const int64_t n_chunks = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(n_chunks), dim3(kSortThreads), 0, s, data,
                     n, partial);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kSortThreads), 0, s, partial,
                     n_chunks);
  hipLaunchKernelGGL(scan_down_kernel, dim3(n_chunks), dim3(kSortThreads), 0, s, data, n,
                     partial);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

// The sort proper over caller-provided histogram and scan storage
// (sort_count_words / sort_partial_words words): no allocation.
template <typename K>
int radix_sort_impl(uint32_t* counts, uint32_t* partials, K* keys, uint32_t* vals,
                    K* keys_alt, uint32_t* vals_alt, int64_t n, int bits, hipStream_t s,
                    int* which, int lo_bit, const SortTile* tiles = nullptr,
                    int64_t n_seg_tiles = 0) {
  *which = 0;
  if (n <= 1 || bits <= 0) return PM_OK;
  const int64_t n_tiles = tiles ? n_seg_tiles : (n + kSortTile - 1) / kSortTile;
  K* kb[2] = {keys, keys_alt};
  uint32_t* vb[2] = {vals, vals_alt};
  int cur = 0;
  for (int shift = lo_bit; shift < bits; shift += 8) {
    hipLaunchKernelGGL(digit_hist_kernel<K>, dim3(n_tiles), dim3(kSortThreads), 0, s,
                       kb[cur], n, shift, counts, n_tiles, tiles);
    // uniform tiles: one row scan per digit (the digits' starts are added by
    // the scatter); a segmented table: the scan over the whole table
    const uint32_t* totals = nullptr;
    if (!tiles) {
      hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(kSortThreads), 0, s, counts, n_tiles,
                         partials);
      totals = partials;
    } else if (int rc = exclusive_scan_impl(partials, counts, n_tiles * 256, s)) {
      return rc;
    }
    hipLaunchKernelGGL(digit_scatter_kernel<K>, dim3(n_tiles), dim3(kSortThreads), 0, s,
                       kb[cur], vb[cur], kb[cur ^ 1], vb[cur ^ 1], n, shift, counts,
                       n_tiles, tiles, totals);
    PM_HIP(hipGetLastError());
    cur ^= 1;
  }
  *which = cur;
  return PM_OK;
}

}  // namespace

int exclusive_scan_u32(uint32_t* partial, uint32_t* data, int64_t n, hipStream_t s) {
  if (n <= 0) return PM_OK;
  return exclusive_scan_impl(partial, data, n, s);
}

int64_t scan_partial_words(int64_t n) {
  return (std::max<int64_t>(n, 1) + kScanChunk - 1) / kScanChunk + 1;
}

int64_t sort_count_words(int64_t n) {
  const int64_t n_tiles = (std::max<int64_t>(n, 1) + kSortTile - 1) / kSortTile;
  return n_tiles * 256;
}

int64_t sort_partial_words(int64_t n) {
  // the scan partials of the table, and at least the 256 digit totals
  return std::max<int64_t>((sort_count_words(n) + kScanChunk - 1) / kScanChunk + 1, 256);
}

int64_t sort_seg_tiles(int64_t n, int n_seg) {
  // each segment's last tile may be partial: at most n_seg more than uniform
  return (std::max<int64_t>(n, 1) + kSortTile - 1) / kSortTile + std::max(n_seg, 1);
}

int64_t sort_seg_count_words(int64_t n, int n_seg) { return sort_seg_tiles(n, n_seg) * 256; }

int64_t sort_seg_partial_words(int64_t n, int n_seg) {
  return (sort_seg_count_words(n, n_seg) + kScanChunk - 1) / kScanChunk + 1;
}

int radix_sort_seg_u32_raw(uint32_t* counts, uint32_t* partials, const SortTile* tiles,
                           int64_t n_tiles, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                           uint32_t* vals_alt, int64_t n, int bits, hipStream_t s, int* which,
                           int lo_bit) {
  return radix_sort_impl<uint32_t>(counts, partials, keys, vals, keys_alt, vals_alt, n, bits, s,
                                   which, lo_bit, tiles, n_tiles);
}

int sort_reserve(Workspace* ws, int64_t n) {
  if (int rc = ws->counts.reserve(static_cast<size_t>(sort_count_words(n)) * sizeof(uint32_t)))
    return rc;
  return ws->partials.reserve(static_cast<size_t>(sort_partial_words(n)) * sizeof(uint32_t));
}

int radix_sort_u32_raw(uint32_t* counts, uint32_t* partials, uint32_t* keys, uint32_t* vals,
                       uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                       hipStream_t s, int* which, int lo_bit) {
  return radix_sort_impl<uint32_t>(counts, partials, keys, vals, keys_alt, vals_alt, n, bits,
                                   s, which, lo_bit);
}


int radix_sort_u32(Workspace* ws, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                   uint32_t* vals_alt, int64_t n, int bits, hipStream_t s, int* which,
                   int lo_bit) {
  if (int rc = sort_reserve(ws, n)) return rc;
  return radix_sort_u32_raw(ws->counts.as<uint32_t>(), ws->partials.as<uint32_t>(), keys, vals,
                            keys_alt, vals_alt, n, bits, s, which, lo_bit);
}


}  // namespace pm
