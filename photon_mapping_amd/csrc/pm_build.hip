// Photon-map construction and query preparation kernels.
//
// Build (replaces KdTree::buildTree, include/photon_map.h:61-105,157-164):
//   1. bounding box of the photon positions (and a finiteness check);
//   2. 30-bit Hilbert code per photon over a 10-bit grid per axis;
//   3. stable radix sort of (code, original index);
//   4. reorder into slot order: pos4 = (x, y, z, index bits),
//      wif4 = (wi, flux.r), flux2 = (flux.g, flux.b), plus index -> slot;
//   5. a balanced binary tree over leaf buckets of leaf_size consecutive
//      slots: leaf boxes, then each level's boxes as unions of its children
//      (median split by count over the Hilbert order);
//   6. radius anchors every anchor_stride slots (centroid + Hilbert code),
//      and a table from 16-bit code prefixes to anchors.
// The anchors' k-NN radius bounds are computed lazily per k by the anchor
// bound kernel (launched through the gather driver in anchor mode).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pm_internal.hpp"

namespace pm {
namespace {

constexpr int kThreads = 256;

// Spreads the low 21 bits of v to every third bit (bit i to bit 3i).
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t quantize(float v, float lo, float scale) {
  const float t = (v - lo) * scale;
  if (!(t > 0.0f)) return 0u;
  constexpr uint32_t kTop = (1u << kGridBits) - 1u;
  if (t >= static_cast<float>(kTop)) return kTop;
  return static_cast<uint32_t>(t);
}

// Hilbert index of the kGridBits-bit quantized position (Skilling 2004,
// "Programming the Hilbert curve", AIP Conf. Proc. 707: axes to transposed
// index, then the transpose's bits interleaved), cut to its top kCodeBits
// bits.  Consecutive indices are always
// face-adjacent cells, so runs of sorted photons or queries stay compact
// where a Morton (Z) order jumps.
__device__ __forceinline__ uint32_t hilbert3(float x, float y, float z, const float* lo,
                                             const float* scale) {
  uint32_t X[3] = {quantize(x, lo[0], scale[0]), quantize(y, lo[1], scale[1]),
                   quantize(z, lo[2], scale[2])};
  constexpr uint32_t M = 1u << (kGridBits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  const uint64_t h = (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
  return static_cast<uint32_t>(h >> (3 * kGridBits - kCodeBits));
}

// Per-block min/max of positions plus a count of non-finite coordinates.
__global__ __launch_bounds__(kThreads) void bbox_partial_kernel(
    const float* __restrict__ pos, int64_t n, float* __restrict__ partial,
    uint32_t* __restrict__ bad) {
  __shared__ float smin[3][kThreads];
  __shared__ float smax[3][kThreads];
  const int t = threadIdx.x;
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t nbad = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + t; i < n;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    for (int c = 0; c < 3; ++c) {
      const float v = pos[c * n + i];
      if (!isfinite(v)) { ++nbad; continue; }
      mn[c] = fminf(mn[c], v);
      mx[c] = fmaxf(mx[c], v);
    }
  }
  if (nbad) atomicAdd(bad, nbad);
  for (int c = 0; c < 3; ++c) { smin[c][t] = mn[c]; smax[c][t] = mx[c]; }
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (t < o) {
      for (int c = 0; c < 3; ++c) {
        smin[c][t] = fminf(smin[c][t], smin[c][t + o]);
        smax[c][t] = fmaxf(smax[c][t], smax[c][t + o]);
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int c = 0; c < 3; ++c) {
      partial[blockIdx.x * 6 + c] = smin[c][0];
      partial[blockIdx.x * 6 + 3 + c] = smax[c][0];
    }
  }
}

// The per-block boxes reduced by one block: each thread a strided share, then
// a tree over the threads.
__global__ __launch_bounds__(kThreads) void bbox_final_kernel(const float* __restrict__ partial,
                                                              int nb, float* __restrict__ out) {
  __shared__ float sr[6][kThreads];
  const int t = threadIdx.x;
  float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = t; b < nb; b += kThreads) {
    for (int c = 0; c < 3; ++c) {
      r[c] = fminf(r[c], partial[b * 6 + c]);
      r[3 + c] = fmaxf(r[3 + c], partial[b * 6 + 3 + c]);
    }
  }
  for (int c = 0; c < 6; ++c) sr[c][t] = r[c];
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (t < o) {
      for (int c = 0; c < 3; ++c) {
        sr[c][t] = fminf(sr[c][t], sr[c][t + o]);
        sr[3 + c][t] = fmaxf(sr[3 + c][t], sr[3 + c][t + o]);
      }
    }
    __syncthreads();
  }
  if (t < 6) out[t] = sr[t][0];
}

struct Quant {
  float lo[3];
  float scale[3];
};

__global__ __launch_bounds__(kThreads) void photon_codes_kernel(
    const float* __restrict__ pos, int64_t n, Quant qz, uint32_t* __restrict__ codes,
    uint32_t* __restrict__ vals) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  codes[i] = hilbert3(pos[i], pos[n + i], pos[2 * n + i], qz.lo, qz.scale);
  vals[i] = static_cast<uint32_t>(i);
}

// inv_perm[order[s]] = s: each photon's slot.
__global__ __launch_bounds__(kThreads) void inv_perm_kernel(const uint32_t* __restrict__ order,
                                                            int64_t n,
                                                            int32_t* __restrict__ inv_perm) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (s >= n) return;
  inv_perm[order[s]] = static_cast<int32_t>(s);
}

// The photons into slot order, taken in original order: the nine SoA reads
// are coalesced and each photon's three records are written at its slot
// (three scattered stores instead of nine scattered loads).
__global__ __launch_bounds__(kThreads) void reorder_kernel(
    const float* __restrict__ pos, const float* __restrict__ wi,
    const float* __restrict__ flux, int64_t n, const int32_t* __restrict__ inv_perm,
    float4* __restrict__ pos4, float4* __restrict__ wif4, float2* __restrict__ flux2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int32_t s = inv_perm[i];
  pos4[s] = make_float4(pos[i], pos[n + i], pos[2 * n + i], __uint_as_float(static_cast<uint32_t>(i)));
  wif4[s] = make_float4(wi[i], wi[n + i], wi[2 * n + i], flux[i]);
  flux2[s] = make_float2(flux[n + i], flux[2 * n + i]);
}

__global__ __launch_bounds__(kThreads) void update_flux_kernel(
    const float* __restrict__ flux, int64_t n, const int32_t* __restrict__ inv_perm,
    float4* __restrict__ wif4, float2* __restrict__ flux2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int32_t s = inv_perm[i];
  wif4[s].w = flux[i];
  flux2[s] = make_float2(flux[n + i], flux[2 * n + i]);
}

__global__ __launch_bounds__(kThreads) void leaf_box_kernel(
    const float4* __restrict__ pos4, int64_t n, int leaf_size, int n_leaves,
    int first_leaf, float4* __restrict__ boxes) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= first_leaf) return;
  float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  if (j < n_leaves) {
    const int64_t s0 = static_cast<int64_t>(j) * leaf_size;
    const int64_t s1 = min(s0 + leaf_size, n);
    for (int64_t s = s0; s < s1; ++s) {
      const float4 p = pos4[s];
      mn.x = fminf(mn.x, p.x); mn.y = fminf(mn.y, p.y); mn.z = fminf(mn.z, p.z);
      mx.x = fmaxf(mx.x, p.x); mx.y = fmaxf(mx.y, p.y); mx.z = fmaxf(mx.z, p.z);
    }
  }
  boxes[2 * (first_leaf + j)] = mn;
  boxes[2 * (first_leaf + j) + 1] = mx;
}

__global__ __launch_bounds__(kThreads) void level_box_kernel(int level_start,
                                                             float4* __restrict__ boxes) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= level_start) return;
  const int node = level_start + i;
  const float4 a0 = boxes[2 * (2 * node)], a1 = boxes[2 * (2 * node) + 1];
  const float4 b0 = boxes[2 * (2 * node + 1)], b1 = boxes[2 * (2 * node + 1) + 1];
  boxes[2 * node] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
  boxes[2 * node + 1] =
      make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
}

__global__ __launch_bounds__(kThreads) void anchor_init_kernel(
    const float4* __restrict__ pos4, const uint32_t* __restrict__ codes, int64_t n,
    int stride, int n_anchors, float4* __restrict__ anchors,
    uint32_t* __restrict__ anchor_code) {
  const int a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n_anchors) return;
  const int64_t s0 = static_cast<int64_t>(a) * stride;
  const int64_t s1 = min(s0 + stride, n);
  float cx = 0.f, cy = 0.f, cz = 0.f;
  for (int64_t s = s0; s < s1; ++s) {
    const float4 p = pos4[s];
    cx += p.x; cy += p.y; cz += p.z;
  }
  const float inv = 1.0f / static_cast<float>(s1 - s0);
  anchors[a] = make_float4(cx * inv, cy * inv, cz * inv, INFINITY);
  anchor_code[a] = codes[s0];
}

// Entry p of the anchor table is the first anchor whose Hilbert code is at or
// above p << (30 - kAnchorTabBits) (n_anchors when there is none), for
// p = 0 .. 2^kAnchorTabBits; the codes are ascending.
__global__ __launch_bounds__(kThreads) void anchor_tab_kernel(
    const uint32_t* __restrict__ anchor_code, int n_anchors, uint32_t* __restrict__ tab) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p > (1 << kAnchorTabBits)) return;
  const uint64_t start = static_cast<uint64_t>(p) << (kCodeBits - kAnchorTabBits);
  int lo = 0, hi = n_anchors;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (anchor_code[m] < start) lo = m + 1;
    else hi = m;
  }
  tab[p] = static_cast<uint32_t>(lo);
}

// ---- query preparation ---------------------------------------------------------

// Hilbert code of local query li of an instance, in the photons' order along
// the curve.  A query outside the supported domain (a non-finite coordinate,
// or one beyond +-2^60) gets the invalid bit: it sorts after every valid
// query of its instance and is left out of the gather (its outputs say so).
__device__ __forceinline__ uint32_t query_code(const GatherDesc& D, int64_t li) {
  float x, y, z;
  if (D.q) {
    const pm_query& q = D.q[li];
    x = q.pos[0]; y = q.pos[1]; z = q.pos[2];
  } else {
    x = D.q_soa[li]; y = D.q_soa[D.m + li]; z = D.q_soa[2 * D.m + li];
  }
  const bool ok = fabsf(x) <= kCoordLimit && fabsf(y) <= kCoordLimit && fabsf(z) <= kCoordLimit;
  return ok ? hilbert3(x, y, z, D.qlo, D.qscale) : kInvalidKeyBit;
}

// Nearest radius anchor of a sorted query's code: the last anchor whose
// first code is <= it (upper bound, one back); 0 for an invalid query, which
// the gather leaves out.  The upper bound lies inside the code's prefix
// bucket of the table: anchors before the bucket have smaller codes, those
// after it larger.
__device__ __forceinline__ uint32_t anchor_of(const GatherDesc& D, uint32_t code) {
  if (code & kInvalidKeyBit) return 0;
  int lo = 0, hi = 0;
  if (D.n_anchors > 0) {
    const uint32_t pre = code >> (kCodeBits - kAnchorTabBits);
    lo = static_cast<int>(D.anchor_tab[pre]);
    hi = static_cast<int>(D.anchor_tab[pre + 1]);
  }
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (D.anchor_code[m] <= code) lo = m + 1;
    else hi = m;
  }
  return static_cast<uint32_t>(lo > 0 ? lo - 1 : 0);
}

// Sort key of every query of one instance; the value is the query index.
__global__ __launch_bounds__(kThreads) void query_keys_kernel(
    const GatherDesc* __restrict__ descs, int64_t total, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ vals) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= total) return;
  keys[i] = query_code(descs[0], i);
  vals[i] = static_cast<uint32_t>(i);
}

// Nearest anchor of every sorted query of one instance, over the sorted keys:
// reads and writes are coalesced and neighbouring queries search the same
// table bucket; q_anchor is indexed by sorted position.
__global__ __launch_bounds__(kThreads) void query_anchor_kernel(
    const GatherDesc* __restrict__ descs, int64_t total, const uint32_t* __restrict__ sorted_keys,
    uint32_t* __restrict__ q_anchor) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= total) return;
  q_anchor[i] = anchor_of(descs[0], sorted_keys[i]);
}

// ---- several instances: one segmented sort -------------------------------------
// The queries of instance i occupy [q_prefix[i], q_prefix[i + 1]) before and
// after the sort, so each is a segment of one sort over 32-bit codes, cut
// into tiles that never straddle two instances; a tile's instance then comes
// with it, with no search per query.

// Tile table of the segmented sort (one block): instance s gets
// ceil(m_s / kSortTileElems) tiles, its counts the table range from 256 times
// its first tile, digit-major within it; the tiles past the last instance's
// are empty and count into a range of their own.
__global__ __launch_bounds__(kThreads) void sort_tiles_kernel(const int64_t* __restrict__ q_prefix,
                                                              int n_seg, SortTile* __restrict__ tiles,
                                                              int64_t n_tiles) {
  __shared__ uint32_t wsum[kThreads / kWave];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int w = t / kWave;
  uint32_t run = 0;  // tiles of the instances before this chunk
  for (int s0 = 0; s0 < n_seg; s0 += kThreads) {
    const int sg = s0 + t;
    int64_t b = 0, m = 0;
    if (sg < n_seg) {
      b = q_prefix[sg];
      m = q_prefix[sg + 1] - b;
    }
    const uint32_t nt = static_cast<uint32_t>((m + kSortTileElems - 1) / kSortTileElems);
    uint32_t incl = nt;
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, chunk = 0;
    for (int i = 0; i < kThreads / kWave; ++i) {
      if (i < w) before += wsum[i];
      chunk += wsum[i];
    }
    const uint32_t first = run + before + incl - nt;
    for (uint32_t j = 0; j < nt; ++j) {
      SortTile tl;
      tl.base = b + static_cast<int64_t>(j) * kSortTileElems;
      tl.n = static_cast<uint32_t>(min<int64_t>(kSortTileElems, m - static_cast<int64_t>(j) * kSortTileElems));
      tl.cbase = first * 256u + j;
      tl.stride = nt;
      tl.seg = static_cast<uint32_t>(sg);
      tiles[first + j] = tl;
    }
    run += chunk;
    __syncthreads();
  }
  for (int64_t i = static_cast<int64_t>(run) + t; i < n_tiles; i += kThreads) {
    SortTile tl;
    tl.base = 0;
    tl.n = 0;
    tl.cbase = static_cast<uint32_t>(i) * 256u;
    tl.stride = 1;
    tl.seg = 0;
    tiles[i] = tl;
  }
}

__global__ __launch_bounds__(kThreads) void query_keys_seg_kernel(
    const GatherDesc* __restrict__ descs, const int64_t* __restrict__ q_prefix,
    const SortTile* __restrict__ tiles, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const SortTile tl = tiles[blockIdx.x];
  const GatherDesc& D = descs[tl.seg];
  const int64_t off = tl.base - q_prefix[tl.seg];
  for (uint32_t j = threadIdx.x; j < tl.n; j += kThreads) {
    keys[tl.base + j] = query_code(D, off + j);
    vals[tl.base + j] = static_cast<uint32_t>(tl.base + j);
  }
}

__global__ __launch_bounds__(kThreads) void query_anchor_seg_kernel(
    const GatherDesc* __restrict__ descs, const SortTile* __restrict__ tiles,
    const uint32_t* __restrict__ sorted_keys, uint32_t* __restrict__ q_anchor) {
  const SortTile tl = tiles[blockIdx.x];
  const GatherDesc& D = descs[tl.seg];
  for (uint32_t j = threadIdx.x; j < tl.n; j += kThreads)
    q_anchor[tl.base + j] = anchor_of(D, sorted_keys[tl.base + j]);
}

// ---- kNN row sort: neighbour indices, farthest first ------------------------------
// One wave per query row: recompute each neighbour's (d^2, index) key from
// the original-order positions, bitonic-sort the row descending, and write
// the indices back (slots past the set hold -1).  The row's keys live in
// LDS (p2 = the power of two >= k, up to kRowSortLds keys) or, for longer
// rows, in a global scratch of p2 keys per row (kGlobal).
template <bool kGlobal>
__global__ __launch_bounds__(kWave) void knn_rowsort_kernel(const GatherDesc* __restrict__ dp,
                                                            int32_t* __restrict__ out_idx,
                                                            uint64_t* __restrict__ scratch,
                                                            int p2) {
  extern __shared__ uint64_t lds_key[];
  const GatherDesc& D = *dp;
  const int64_t q = blockIdx.x;
  uint64_t* key = kGlobal ? scratch + q * p2 : lds_key;
  const int lane = threadIdx.x;
  const int k = D.k;
  const int64_t n = D.n;
  const float qx = D.q_soa[q], qy = D.q_soa[D.m + q], qz = D.q_soa[2 * D.m + q];
  int32_t* row = out_idx + q * k;
  for (int j = lane; j < p2; j += kWave) {
    uint64_t kk = 0;  // padding sorts last (descending); real keys are +1
    if (j < k) {
      const int32_t idx = row[j];
      if (idx >= 0) {
        const float dx = qx - D.pos_orig[idx];
        const float dy = qy - D.pos_orig[n + idx];
        const float dz = qz - D.pos_orig[2 * n + idx];
        float acc = dx * dx;
        acc = acc + dy * dy;
        acc = acc + dz * dz;
        kk = ((static_cast<uint64_t>(__float_as_uint(acc)) << 32) |
              static_cast<uint32_t>(idx)) + 1ull;
      }
    }
    key[j] = kk;
  }
  __syncthreads();
  // bitonic sort, descending
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < p2 / 2; t += kWave) {
        const int i = 2 * t - (t & (stride - 1));
        const int j = i + stride;
        const bool desc = ((i & size) == 0);
        const uint64_t a = key[i], b = key[j];
        if ((a < b) == desc) {
          key[i] = b;
          key[j] = a;
        }
      }
      __syncthreads();
    }
  }
  for (int j = lane; j < k; j += kWave) {
    const uint64_t kk = key[j];
    row[j] = kk != 0 ? static_cast<int32_t>((kk - 1ull) & 0xffffffffull) : -1;
  }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_f32_kernel(float* p, float v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// out[c n + i] += acc[c n + inv_perm[i]] for the three channels: gradients
// accumulated in slot order (the adjoint's atomics land on neighbouring
// slots, a few cache lines per wave) are added to the caller's array in
// original photon order.  Writes are coalesced; reads gather by slot.
__global__ __launch_bounds__(kThreads) void slot_grad_add_kernel(
    const float* __restrict__ acc, const int32_t* __restrict__ inv_perm, int64_t n,
    float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t sl = inv_perm[i];
  out[i] += acc[sl];
  out[n + i] += acc[n + sl];
  out[2 * n + i] += acc[2 * n + sl];
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

int launch_slot_grad_add(const float* acc, const int32_t* inv_perm, int64_t n, float* out,
                         hipStream_t s) {
  if (n <= 0) return PM_OK;
  hipLaunchKernelGGL(slot_grad_add_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s,
                     acc, inv_perm, n, out);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return PM_OK;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, p,
                     v, n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

int launch_fill_f32(float* p, float v, int64_t n, hipStream_t s) {
  if (n <= 0) return PM_OK;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, p,
                     v, n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

// Writes one descriptor (passed by value in the kernel arguments) and its
// query and group prefixes {0, q_total}, {0, g_total} to device memory.
__global__ void store_desc_kernel(GatherDesc d, GatherDesc* __restrict__ dst,
                                  int64_t* __restrict__ prefix, int64_t q_total,
                                  int64_t g_total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *dst = d;
    prefix[0] = 0;
    prefix[1] = q_total;
    prefix[2] = 0;
    prefix[3] = g_total;
  }
}

int launch_store_desc(const GatherDesc& d, GatherDesc* dst, int64_t* prefix, int64_t q_total,
                      int64_t g_total, hipStream_t s) {
  hipLaunchKernelGGL(store_desc_kernel, dim3(1), dim3(64), 0, s, d, dst, prefix, q_total,
                     g_total);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

void fill_desc_map_part(const PhotonMapImpl* map, GatherDesc* d) {
  d->pos4 = map->pos4.as<float4>();
  d->wif4 = map->wif4.as<float4>();
  d->flux2 = map->flux2.as<float2>();
  d->boxes = map->boxes.as<float4>();
  d->anchors = map->anchors.as<float4>();
  d->anchor_code = map->anchor_code.as<uint32_t>();
  d->anchor_tab = map->anchor_tab.as<uint32_t>();
  d->inv_perm = map->inv_perm.as<int32_t>();
  d->pos_orig = map->pos_soa.as<float>();
  for (int c = 0; c < 3; ++c) {
    d->qlo[c] = map->qlo[c];
    d->qscale[c] = map->qscale[c];
  }
  d->n = static_cast<int32_t>(map->n);
  d->leaf_size = map->leaf_size;
  d->n_leaves = map->n_leaves;
  d->first_leaf = map->first_leaf;
  d->n_anchors = map->n_anchors;
  d->anchor_stride = map->anchor_stride;
}

int build_photon_map(PhotonMapImpl* map, hipStream_t s) {
  Context* ctx = map->ctx;
  const int64_t n = map->n;
  map->built = false;
  for (auto& a : map->aset) a.k = -1;
  if (n == 0) {
    map->n_leaves = 0;
    map->first_leaf = 1;
    map->n_anchors = 0;
    map->built = true;
    return PM_OK;
  }
  const float* pos = map->pos_soa.as<float>();

  // 1. bounding box + finiteness
  const int nb = static_cast<int>(std::min<int64_t>(ceil_div(n, kThreads), 1024));
  if (int rc = ctx->ws_tmp.reserve(static_cast<size_t>(nb) * 6 * sizeof(float) + 64)) return rc;
  float* partial = ctx->ws_tmp.as<float>();
  float* box_out = partial + nb * 6;  // 6 floats
  uint32_t* bad = reinterpret_cast<uint32_t*>(box_out + 6);
  if (int rc = launch_fill_u32(bad, 0u, 1, s)) return rc;
  hipLaunchKernelGGL(bbox_partial_kernel, dim3(nb), dim3(kThreads), 0, s, pos, n, partial,
                     bad);
  hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(kThreads), 0, s, partial, nb, box_out);
  PM_HIP(hipGetLastError());
  float hb[7];
  PM_HIP(hipMemcpyAsync(hb, box_out, 7 * sizeof(float), hipMemcpyDeviceToHost, s));
  PM_HIP(hipStreamSynchronize(s));
  uint32_t nbad;
  std::memcpy(&nbad, &hb[6], sizeof(nbad));
  if (nbad != 0) return fail(PM_ERR_INVALID_ARG, "photon positions must be finite");
  for (int c = 0; c < 6; ++c)
    if (!(std::fabs(hb[c]) <= kCoordLimit))
      return fail(PM_ERR_INVALID_ARG, "photon coordinates must lie within +-2^60");
  for (int c = 0; c < 3; ++c) {
    const float ext = hb[3 + c] - hb[c];
    map->qlo[c] = hb[c];
    map->qscale[c] = ext > 0.0f ? static_cast<float>(1u << kGridBits) / ext : 1.0f;
  }

  // 2-3. Hilbert codes, stable sort (over the context's query-sort buffers:
  // the query order they held is gone)
  ctx->last_order = Context::QueryOrder{};
  if (int rc = ctx->ws.keys.reserve(n * sizeof(uint32_t))) return rc;
  if (int rc = ctx->ws.vals.reserve(n * sizeof(uint32_t))) return rc;
  if (int rc = ctx->ws.keys2.reserve(n * sizeof(uint32_t))) return rc;
  if (int rc = ctx->ws.vals2.reserve(n * sizeof(uint32_t))) return rc;
  Quant qz;
  for (int c = 0; c < 3; ++c) { qz.lo[c] = map->qlo[c]; qz.scale[c] = map->qscale[c]; }
  uint32_t* k0 = ctx->ws.keys.as<uint32_t>();
  uint32_t* v0 = ctx->ws.vals.as<uint32_t>();
  uint32_t* k1 = ctx->ws.keys2.as<uint32_t>();
  uint32_t* v1 = ctx->ws.vals2.as<uint32_t>();
  hipLaunchKernelGGL(photon_codes_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s,
                     pos, n, qz, k0, v0);
  PM_HIP(hipGetLastError());
  int which = 0;
  if (int rc = radix_sort_u32(&ctx->ws, k0, v0, k1, v1, n, kCodeBits, s, &which)) return rc;
  const uint32_t* sorted_codes = which ? k1 : k0;
  const uint32_t* order = which ? v1 : v0;

  // 4. reorder into slot order
  if (int rc = map->pos4.reserve(n * sizeof(float4))) return rc;
  if (int rc = map->wif4.reserve(n * sizeof(float4))) return rc;
  if (int rc = map->flux2.reserve(n * sizeof(float2))) return rc;
  if (int rc = map->inv_perm.reserve(n * sizeof(int32_t))) return rc;
  if (int rc = map->codes.reserve(n * sizeof(uint32_t))) return rc;
  hipLaunchKernelGGL(inv_perm_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, order,
                     n, map->inv_perm.as<int32_t>());
  hipLaunchKernelGGL(reorder_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, pos,
                     map->wi_soa.as<float>(), map->flux_soa.as<float>(), n,
                     map->inv_perm.as<int32_t>(), map->pos4.as<float4>(),
                     map->wif4.as<float4>(), map->flux2.as<float2>());
  PM_HIP(hipMemcpyAsync(map->codes.ptr, sorted_codes, n * sizeof(uint32_t),
                        hipMemcpyDeviceToDevice, s));

  // 5. tree boxes
  map->n_leaves = static_cast<int>(ceil_div(n, map->leaf_size));
  int fl = 1;
  while (fl < map->n_leaves) fl <<= 1;
  map->first_leaf = fl;
  if (int rc = map->boxes.reserve(static_cast<size_t>(2 * fl) * 2 * sizeof(float4))) return rc;
  float4* boxes = map->boxes.as<float4>();
  hipLaunchKernelGGL(leaf_box_kernel, dim3(ceil_div(fl, kThreads)), dim3(kThreads), 0, s,
                     map->pos4.as<float4>(), n, map->leaf_size, map->n_leaves, fl, boxes);
  for (int level = fl >> 1; level >= 1; level >>= 1) {
    hipLaunchKernelGGL(level_box_kernel, dim3(ceil_div(level, kThreads)), dim3(kThreads), 0,
                       s, level, boxes);
  }
  PM_HIP(hipGetLastError());

  // 6. anchors
  map->anchor_stride = ctx->anchor_stride;
  map->n_anchors = static_cast<int>(ceil_div(n, map->anchor_stride));
  if (int rc = map->anchors.reserve(static_cast<size_t>(map->n_anchors) * sizeof(float4)))
    return rc;
  if (int rc = map->anchor_code.reserve(static_cast<size_t>(map->n_anchors) * sizeof(uint32_t)))
    return rc;
  hipLaunchKernelGGL(anchor_init_kernel, dim3(ceil_div(map->n_anchors, kThreads)),
                     dim3(kThreads), 0, s, map->pos4.as<float4>(),
                     map->codes.as<uint32_t>(), n, map->anchor_stride, map->n_anchors,
                     map->anchors.as<float4>(), map->anchor_code.as<uint32_t>());
  PM_HIP(hipGetLastError());
  constexpr int kTab = (1 << kAnchorTabBits) + 1;
  if (int rc = map->anchor_tab.reserve(static_cast<size_t>(kTab) * sizeof(uint32_t))) return rc;
  hipLaunchKernelGGL(anchor_tab_kernel, dim3(ceil_div(kTab, kThreads)), dim3(kThreads), 0, s,
                     map->anchor_code.as<uint32_t>(), map->n_anchors,
                     map->anchor_tab.as<uint32_t>());
  PM_HIP(hipGetLastError());
  map->built = true;
  return PM_OK;
}

int update_flux(PhotonMapImpl* map, const float* flux_dev, hipStream_t s) {
  if (map->n == 0 || !map->built) return PM_OK;
  hipLaunchKernelGGL(update_flux_kernel, dim3(ceil_div(map->n, kThreads)), dim3(kThreads), 0,
                     s, flux_dev, map->n, map->inv_perm.as<int32_t>(),
                     map->wif4.as<float4>(), map->flux2.as<float2>());
  PM_HIP(hipGetLastError());
  return PM_OK;
}

GatherDesc anchor_desc(const PhotonMapImpl* map, int k_eff, float* out) {
  GatherDesc d{};
  fill_desc_map_part(map, &d);
  d.m = map->n_anchors;
  d.q_base = 0;
  d.group_base = 0;
  d.n_groups = ceil_div(map->n_anchors, kWave);
  d.k = k_eff;
  d.k_eff = k_eff;
  d.all_members = 0;
  d.anchor_out = out;
  return d;
}

const float4* cached_anchors(PhotonMapImpl* map, int k_eff) {
  for (auto& a : map->aset)
    if (a.k == k_eff) {
      a.stamp = ++map->anchor_clock;
      return a.buf.as<float4>();
    }
  return nullptr;
}

int launch_anchor_bounds(Context* ctx, const GatherDesc* dev_desc, const int64_t* dev_gp,
                         int64_t n_anchors, uint32_t* counters, hipStream_t s) {
  GatherLaunch L{};
  L.descs = dev_desc;
  L.group_prefix = dev_gp;
  L.perm = nullptr;
  L.q_prefix = nullptr;
  L.counters = counters;
  L.total_groups = n_anchors;  // the anchor kernel takes one anchor per wave at a time
  L.n_inst = 1;
  L.xcd_aware = 0;
  L.count_stats = 0;
  L.run_split = ctx->run_split;
  L.mode = kModeAnchor;
  return launch_gather(ctx, L, s);
}

int ensure_anchors(PhotonMapImpl* map, int k_eff, hipStream_t s, const float4** out,
                   uint64_t pin_epoch) {
  *out = map->anchors.as<float4>();
  if (map->n == 0 || k_eff >= map->n) return PM_OK;  // every photon is a neighbour
  for (auto& a : map->aset)
    if (a.k == k_eff) {
      a.stamp = ++map->anchor_clock;
      if (pin_epoch) a.pin = pin_epoch;
      *out = a.buf.as<float4>();
      return PM_OK;
    }
  // the least recently used set not pinned by this batch (an empty one first)
  PhotonMapImpl::AnchorSet* slot = nullptr;
  for (auto& a : map->aset) {
    if (pin_epoch && a.pin == pin_epoch) continue;
    if (!slot || a.k < 0 || (slot->k >= 0 && a.stamp < slot->stamp)) slot = &a;
    if (slot->k < 0) break;
  }
  if (!slot) {
    map->aset.emplace_back();
    slot = &map->aset.back();
  }
  slot->pin = pin_epoch;
  Context* ctx = map->ctx;
  const size_t bytes = static_cast<size_t>(map->n_anchors) * sizeof(float4);
  slot->k = -1;
  if (int rc = slot->buf.reserve(bytes)) return rc;
  PM_HIP(hipMemcpyAsync(slot->buf.ptr, map->anchors.ptr, bytes, hipMemcpyDeviceToDevice, s));
  const GatherDesc d = anchor_desc(map, k_eff, slot->buf.as<float>());
  if (int rc = ctx->ws.counters.reserve(64 * sizeof(uint32_t))) return rc;
  int64_t gp[2] = {0, d.n_groups};
  if (int rc = stage_begin(ctx, sizeof(GatherDesc) + sizeof(gp) + 256, s)) return rc;
  const auto* dd = static_cast<const GatherDesc*>(stage_put(ctx, &d, sizeof(d)));
  const auto* dgp = static_cast<const int64_t*>(stage_put(ctx, gp, sizeof(gp)));
  if (int rc = stage_commit(ctx, s)) return rc;
  if (int rc = launch_anchor_bounds(ctx, dd, dgp, d.m, ctx->ws.counters.as<uint32_t>(), s))
    return rc;
  slot->k = k_eff;
  slot->stamp = ++map->anchor_clock;
  *out = slot->buf.as<float4>();
  return PM_OK;
}

int prepare_reserve(Workspace* ws, int n_inst, int64_t total_q, int64_t total_groups) {
  (void)total_groups;
  const int64_t nq = std::max<int64_t>(total_q, 1);
  if (int rc = ws->keys.reserve(nq * sizeof(uint32_t))) return rc;
  if (int rc = ws->keys2.reserve(nq * sizeof(uint32_t))) return rc;
  if (int rc = ws->vals.reserve(nq * sizeof(uint32_t))) return rc;
  if (int rc = ws->vals2.reserve(nq * sizeof(uint32_t))) return rc;
  if (int rc = ws->q_anchor.reserve(nq * sizeof(uint32_t))) return rc;
  if (int rc = ws->counters.reserve(64 * sizeof(uint32_t))) return rc;
  if (n_inst <= 1) return sort_reserve(ws, nq);
  if (int rc = ws->tiles.reserve(static_cast<size_t>(sort_seg_tiles(nq, n_inst)) *
                                 sizeof(SortTile)))
    return rc;
  if (int rc = ws->counts.reserve(static_cast<size_t>(sort_seg_count_words(nq, n_inst)) *
                                  sizeof(uint32_t)))
    return rc;
  return ws->partials.reserve(static_cast<size_t>(sort_seg_partial_words(nq, n_inst)) *
                              sizeof(uint32_t));
}

int prepare_queries(Workspace* ws, int n_inst, const GatherDesc* dev_descs,
                    const int64_t* dev_q_prefix, const int64_t* dev_g_prefix,
                    int64_t total_q, int64_t total_groups, hipStream_t s,
                    const uint32_t** perm_out, bool need_anchors) {
  (void)dev_g_prefix;
  if (total_q == 0) {
    *perm_out = nullptr;
    return PM_OK;
  }
  if (int rc = prepare_reserve(ws, n_inst, total_q, total_groups)) return rc;
  QueryScratch qs;
  qs.keys = ws->keys.ptr;
  qs.keys2 = ws->keys2.ptr;
  qs.vals = ws->vals.as<uint32_t>();
  qs.vals2 = ws->vals2.as<uint32_t>();
  qs.q_anchor = ws->q_anchor.as<uint32_t>();
  qs.counts = ws->counts.as<uint32_t>();
  qs.partials = ws->partials.as<uint32_t>();
  qs.tiles = n_inst > 1 ? ws->tiles.as<SortTile>() : nullptr;
  return prepare_queries_raw(qs, n_inst, dev_descs, dev_q_prefix, total_q, s, perm_out,
                             need_anchors);
}

int prepare_queries_raw(const QueryScratch& qs, int n_inst, const GatherDesc* dev_descs,
                        const int64_t* dev_q_prefix, int64_t total_q, hipStream_t s,
                        const uint32_t** perm_out, bool need_anchors) {
  if (total_q == 0) {
    *perm_out = nullptr;
    return PM_OK;
  }
  uint32_t* k0 = static_cast<uint32_t*>(qs.keys);
  uint32_t* k1 = static_cast<uint32_t*>(qs.keys2);
  uint32_t* v0 = qs.vals;
  uint32_t* v1 = qs.vals2;
  int which = 0;
  if (n_inst <= 1) {
    const dim3 grid(ceil_div(total_q, kThreads));
    hipLaunchKernelGGL(query_keys_kernel, grid, dim3(kThreads), 0, s, dev_descs, total_q, k0, v0);
    PM_HIP(hipGetLastError());
    if (int rc = radix_sort_u32_raw(qs.counts, qs.partials, k0, v0, k1, v1, total_q, 32, s,
                                    &which, kQuerySortLoBit))
      return rc;
    if (need_anchors)
      hipLaunchKernelGGL(query_anchor_kernel, grid, dim3(kThreads), 0, s, dev_descs, total_q,
                         which ? k1 : k0, qs.q_anchor);
  } else {
    if (!qs.tiles) return fail(PM_ERR_INVALID_ARG, "a batch needs the sort's tile table");
    const int64_t n_tiles = sort_seg_tiles(total_q, n_inst);
    hipLaunchKernelGGL(sort_tiles_kernel, dim3(1), dim3(kThreads), 0, s, dev_q_prefix, n_inst,
                       qs.tiles, n_tiles);
    hipLaunchKernelGGL(query_keys_seg_kernel, dim3(n_tiles), dim3(kThreads), 0, s, dev_descs,
                       dev_q_prefix, qs.tiles, k0, v0);
    PM_HIP(hipGetLastError());
    if (int rc = radix_sort_seg_u32_raw(qs.counts, qs.partials, qs.tiles, n_tiles, k0, v0, k1, v1,
                                        total_q, 32, s, &which, kQuerySortLoBit))
      return rc;
    if (need_anchors)
      hipLaunchKernelGGL(query_anchor_seg_kernel, dim3(n_tiles), dim3(kThreads), 0, s, dev_descs,
                         qs.tiles, which ? k1 : k0, qs.q_anchor);
  }
  PM_HIP(hipGetLastError());
  *perm_out = which ? v1 : v0;
  return PM_OK;
}

int launch_knn_rowsort(const GatherDesc* dev_desc, int64_t m, int k, int32_t* out_idx,
                       DevBuf* scratch, hipStream_t s) {
  if (m <= 0) return PM_OK;
  int p2 = kWave;
  while (p2 < k) p2 <<= 1;
  if (p2 <= kRowSortLds) {
    hipLaunchKernelGGL(knn_rowsort_kernel<false>, dim3(m), dim3(kWave),
                       static_cast<unsigned>(p2 * sizeof(uint64_t)), s, dev_desc, out_idx,
                       nullptr, p2);
  } else {
    if (int rc = scratch->reserve(static_cast<size_t>(m) * p2 * sizeof(uint64_t))) return rc;
    hipLaunchKernelGGL(knn_rowsort_kernel<true>, dim3(m), dim3(kWave), 0, s, dev_desc, out_idx,
                       scratch->as<uint64_t>(), p2);
  }
  PM_HIP(hipGetLastError());
  return PM_OK;
}

}  // namespace pm
