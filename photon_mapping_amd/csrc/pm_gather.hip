// Persistent kNN photon-gather kernel for gfx950.
//
// One wavefront owns a group of 64 queries adjacent in Hilbert order (one per lane).
// The k nearest neighbours are never kept in a heap.  Each lane selects the
// exact k-th smallest key
//   key = (bits(d^2) << 32) | original photon index
// which orders photons exactly like the reference's
// std::priority_queue<std::pair<float,int>> (include/photon_map.h:107-126).
//
// The lanes of a group are cut into runs of nearby queries (usually one run
// per group).  For a run with centre c and query spread delta:
//   centre  a wave-parallel histogram of d^2(c, p) over the photons of a
//           guaranteed bound gives r_c, a radius holding at least k photons
//           around c.  By the triangle inequality the k nearest photons of
//           every lane q lie within R_q = min(r_c + |q - c|, r0_q) of q (r0_q
//           the lane's own anchor bound), and all of them within the largest
//           R_q + |q - c| of c: that sphere's photons are the run's candidate
//           list (a few hundred at leaf size 8);
//   pass A  per lane, a histogram of d^2 over [0, R_q^2] in kBins linear bins
//           (LDS) finds the bin holding the k-th key, i.e. a key window;
//           the list then shrinks to the sphere that holds every window;
//   pass B  photons below the window are neighbours and are consumed on the
//           spot (radiance sum, index write, ...); photons inside it are
//           stashed in LDS (a handful), and a rank selection inside the
//           stash finishes the set.
// The neighbour set is {key <= tau}; tau's d^2 is r^2.  Bins that hold more
// than kStashCap keys are split further (rare) until they fit.
//
// Candidate lists are found by a breadth-first walk of the implicit photon
// tree that expands several levels per step so that one vector load per lane
// tests up to 64 nodes at once.  Candidates of the per-lane passes are
// staged into LDS 64 at a time by a coalesced vector load, the next chunk
// being fetched into registers while the current one is consumed through
// broadcast LDS reads.  The consume loops are branch-free: the rare per-lane
// events (stash, index write) are recorded in a 64-bit mask per chunk and
// drained once per chunk.
//
// d^2 and the Lambert sign tests use the reference's FP32 operation order
// (include/photon_map.h:33-41, include/core.h:206-209); the file is compiled
// with -ffp-contract=off and denormals preserved, so they are bit-exact.
// Every radius bound carries an upward relative margin (>= 1e-5) that covers
// the rounding of the FP32 arithmetic it is derived with.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "pm_internal.hpp"

// Read-only data reached through the constant address space is fetched with
// scalar loads when the address is wave-uniform.  The qualifier only means
// something to the gfx950 device pass; the host pass merely parses these
// device functions and rejects class copies out of a foreign address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define PM_CONST __attribute__((address_space(4)))
#else
#define PM_CONST
#endif

namespace pm {
namespace {

constexpr float kPiF = 3.14159265359f;  // include/core.h:7
constexpr uint64_t kKeyMax = ~0ull;
constexpr int kChunk = 64;        // candidate photons staged per LDS chunk
#ifndef PM_MAX_RANGES
#define PM_MAX_RANGES 192
#endif
#ifndef PM_SLOT_CAP
#define PM_SLOT_CAP 384
#endif
constexpr int kMaxRanges = PM_MAX_RANGES;  // candidate subtrees per list
// deepest level whose nodes always fit a list: 2^kCapDepth <= kMaxRanges
constexpr int kCapDepth = kMaxRanges >= 256 ? 8 : kMaxRanges >= 128 ? 7 : kMaxRanges >= 64 ? 6 : 5;
static_assert(kMaxRanges >= 32, "a depth-capped walk must fit the list");
static_assert((1 << kCapDepth) <= kMaxRanges, "depth cap");
constexpr int kFront = 256;       // breadth-first frontier capacity
constexpr int kNarrowTries = 12;  // lists tried when narrowing a loose sphere
constexpr int kSparseRuns = 3;        // runs of a group at which runs are cut wider
constexpr float kSparseSplit = 1.67f;  // ... by this factor on the run split
constexpr int kDrhoSlots = 128;   // adjoint: materials accumulated in LDS
constexpr int kCentreBins = 256;  // centre histogram bins
constexpr int kSlotCap = PM_SLOT_CAP;  // photons of an explicit slot list
// A candidate list is either nr subtrees (sh.rl.rnode / sh.rl.rpre) or, when nr
// is this value, an explicit list of photon slots (sh.cslot).
constexpr uint32_t kSlotList = 0xffffffffu;
// build_list's answer, when asked to give up rather than coarsen, that the
// sphere needs more subtrees than a list holds.
constexpr uint32_t kNoList = 0xfffffffeu;
// What build_list does when a list would overflow (after the retry with whole
// inner nodes, for kListGiveUp; at once, for kListLeavesOnly).
constexpr int kListCoarsen = 0;
constexpr int kListGiveUp = 1;
constexpr int kListLeavesOnly = 2;

template <typename T>
__device__ __forceinline__ const PM_CONST T* cptr(const T* p) {
  return (const PM_CONST T*)(p);
}

// Descriptor fields are read straight from constant memory where they are
// used, rather than held in registers across the whole group.
using DescRef = const PM_CONST GatherDesc&;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Subtree list: roots and the exclusive prefix of their slot counts.
struct RangeList {
  uint32_t rnode[kMaxRanges];
  uint32_t rpre[kMaxRanges];
};

// Pass B's use of the lane-bin storage: the per-lane stash of keys (entry e
// of lane l at [e * kWave + l]; a key's photon slot is found again through
// the map's inverse permutation) and the slots of the staged chunk.
struct StashArea {
  uint64_t key[kStashCap * kWave];
  uint32_t chunk_slot[kChunk];
};

// One wave's LDS (one-wave workgroups).  Buffers whose lifetimes do not
// overlap share storage, which keeps the block at 10 KB so that 16 waves
// fit a CU's 160 KB: the candidate list and the staged positions live
// throughout; the lane bins (pass A, refinement, centre bins, the adjoint's
// query staging), pass B's staged flux and stash, collect_slots' output and
// the breadth-first frontier of a list walk take turns in the rest.
struct GatherShared {
  // A candidate list is a subtree list or an explicit slot list, never both.
  union {
    RangeList rl;
    uint32_t cslot[kSlotCap];
  };
  float4 pos[kChunk];  // staged candidate positions
  union {
    uint32_t hist[(kBins + 2) * kWave];  // per-lane bins
    struct {
      float4 wif[kChunk];  // with a shared normal: effective flux (rgb)
      float2 flux[kChunk];
      StashArea sb;
    } b;                                 // pass B / finish
    uint32_t cscratch[kSlotCap];         // collect_slots output before it is accepted
    uint32_t front[2][kFront];           // list walks
  };
};
// The adjoint kernel also accumulates dL/drho per material in LDS.
struct AdjointShared : GatherShared {
  float drho[3 * kDrhoSlots];
};
static_assert(kSlotCap * sizeof(uint32_t) <= sizeof(RangeList), "slot list fits the range list");
static_assert(kCentreBins <= (kBins + 2) * kWave, "centre bins alias the lane bins");

__device__ __forceinline__ float d2_ref(float qx, float qy, float qz, float px,
                                        float py, float pz) {
  const float dx = qx - px;
  const float dy = qy - py;
  const float dz = qz - pz;
  float acc = dx * dx;
  acc = acc + dy * dy;
  acc = acc + dz * dz;
  return acc;
}

__device__ __forceinline__ float dot_ref(float ax, float ay, float az, float bx,
                                         float by, float bz) {
  float s = ax * bx;
  s = s + ay * by;
  s = s + az * bz;
  return s;
}

// Conservative radius^2 for culling: every photon whose FP32 d^2 is <= r2
// lies within sqrt(cull_bound(r2)) of the query, despite the rounding of d^2.
__device__ __forceinline__ float cull_bound(float r2) {
  return r2 * 1.00002f + 1e-30f;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx) {
  return (static_cast<uint64_t>(__float_as_uint(d2)) << 32) | idx;
}

// Histogram bin of d^2 for a lane bracket [lo2, hi2): with inv_w =
// kBins / (hi2 - lo2) and c1 = 1 - lo2 * inv_w, t = fma(d2, inv_w, c1) is
// 1 + kBins (d2 - lo2) / (hi2 - lo2) up to rounding; bin 0 holds d^2 below
// the bracket, bins 1..kBins split it linearly, bin kBins + 1 holds the rest.
// Monotone in d^2, which is all the selection relies on: a key in bin 0 is
// below every key of the higher bins.  A lane outside the current run uses
// inv_w = 0, c1 = 64 (always bin kBins + 1); an unbounded lane inv_w = 0,
// c1 = 1 (always bin 1).
__device__ __forceinline__ uint32_t bin_of(float d2, float inv_w, float c1) {
  const float t = fmaf(d2, inv_w, c1);
  return static_cast<uint32_t>(fminf(fmaxf(t, 0.0f), static_cast<float>(kBins + 1)));
}

// Smallest non-negative finite float (as bits) whose bin is >= b, or the
// bits of +inf when there is none.  The quotient (b - c1) / inv_w lands
// within a few ulps of the answer, which is then found by stepping; a
// bisection covers any case the steps do not settle.
__device__ __forceinline__ uint32_t bin_start_bits(uint32_t b, float inv_w, float c1) {
  constexpr uint32_t kInfBits = 0x7f800000u;
  if (b == 0) return 0;
  auto ge = [&](uint32_t u) { return u < kInfBits && bin_of(__uint_as_float(u), inv_w, c1) >= b; };
  const float x = (static_cast<float>(b) - c1) / inv_w;
  uint32_t u = (x >= 0.0f) ? __float_as_uint(x) : 0u;  // NaN -> 0
  u = u < kInfBits ? u : kInfBits;
  for (int s = 0; s < 8; ++s) {
    if (u > 0 && ge(u - 1)) { --u; continue; }
    if (u < kInfBits && !ge(u)) { ++u; continue; }
    return u;
  }
  uint32_t lo = 0;
  uint32_t hi = kInfBits;
  for (int it = 0; it < 32 && lo < hi; ++it) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (ge(mid)) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// ---- wave-wide scans and reductions over DPP row shifts and broadcasts ------
// (every lane of the wave must be active).  After row_shr 1, 2, 4, 8 each
// lane holds the prefix of its 16-lane row; row_bcast 15 and 31 carry the
// row totals across rows, so lane 63 ends with the total.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(old), static_cast<int>(v), CTRL, ROW_MASK, 0xf, false));
}

// Inclusive prefix sum over the lanes.
__device__ __forceinline__ uint32_t wave_scan_add(uint32_t v) {
  v += dpp_u32<0x111, 0xf>(0u, v);
  v += dpp_u32<0x112, 0xf>(0u, v);
  v += dpp_u32<0x114, 0xf>(0u, v);
  v += dpp_u32<0x118, 0xf>(0u, v);
  v += dpp_u32<0x142, 0xa>(0u, v);
  v += dpp_u32<0x143, 0xc>(0u, v);
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, dpp_u32<0x111, 0xf>(~0u, v));
  v = min(v, dpp_u32<0x112, 0xf>(~0u, v));
  v = min(v, dpp_u32<0x114, 0xf>(~0u, v));
  v = min(v, dpp_u32<0x118, 0xf>(~0u, v));
  v = min(v, dpp_u32<0x142, 0xa>(~0u, v));
  v = min(v, dpp_u32<0x143, 0xc>(~0u, v));
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), kWave - 1));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, dpp_u32<0x111, 0xf>(0u, v));
  v = max(v, dpp_u32<0x112, 0xf>(0u, v));
  v = max(v, dpp_u32<0x114, 0xf>(0u, v));
  v = max(v, dpp_u32<0x118, 0xf>(0u, v));
  v = max(v, dpp_u32<0x142, 0xa>(0u, v));
  v = max(v, dpp_u32<0x143, 0xc>(0u, v));
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), kWave - 1));
}

// Order-preserving map of floats onto unsigned integers (and back).
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Wave-uniform minimum / maximum of a float.
__device__ __forceinline__ float wave_min(float v) { return funkey(wave_min_u32(fkey(v))); }
__device__ __forceinline__ float wave_max(float v) { return funkey(wave_max_u32(fkey(v))); }

// Wave-uniform sum of a float, in a fixed order (the same scan as above).
__device__ __forceinline__ float wave_sum(float v) {
  auto add = [](float x, uint32_t y) { return x + __uint_as_float(y); };
  v = add(v, dpp_u32<0x111, 0xf>(0u, __float_as_uint(v)));
  v = add(v, dpp_u32<0x112, 0xf>(0u, __float_as_uint(v)));
  v = add(v, dpp_u32<0x114, 0xf>(0u, __float_as_uint(v)));
  v = add(v, dpp_u32<0x118, 0xf>(0u, __float_as_uint(v)));
  v = add(v, dpp_u32<0x142, 0xa>(0u, __float_as_uint(v)));
  v = add(v, dpp_u32<0x143, 0xc>(0u, __float_as_uint(v)));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), kWave - 1));
}

// Adds a lane's (r, g, b) to the albedo gradient of material mat (nothing
// for mat < 0).  Materials below kDrhoSlots accumulate in the block's LDS
// copy.  The others go to global memory, summed first over the lanes that
// share the material: many lanes of a wave usually share one (a large face,
// or a light quad typed Lambert), and one address taking an atomic per lane
// from every wave serialises them.  Call with every lane of the wave active.
__device__ __forceinline__ void add_albedo_grad(float* lds, float* glob, int mat, float r,
                                                float g, float b, int lane) {
  if (mat >= 0 && mat < kDrhoSlots) {
    atomicAdd(&lds[3 * mat + 0], r);
    atomicAdd(&lds[3 * mat + 1], g);
    atomicAdd(&lds[3 * mat + 2], b);
  }
  bool pending = mat >= kDrhoSlots;
  uint64_t pm = __ballot(pending);
  while (pm != 0ull) {
    const int leader = static_cast<int>(__builtin_ctzll(pm));
    const int m0 = __builtin_amdgcn_readlane(mat, leader);
    const bool mine = pending && mat == m0;
    const float sr = wave_sum(mine ? r : 0.0f);
    const float sg = wave_sum(mine ? g : 0.0f);
    const float sb = wave_sum(mine ? b : 0.0f);
    if (lane == leader) {
      atomicAdd(&glob[3 * m0 + 0], sr);
      atomicAdd(&glob[3 * m0 + 1], sg);
      atomicAdd(&glob[3 * m0 + 2], sb);
    }
    pending = pending && !mine;
    pm = __ballot(pending);
  }
}

// Square roots for radius bounds, from the hardware instruction (1 ulp);
// every bound carries a relative margin far above that.  The upper form
// never underestimates a denormal argument, the lower never overestimates.
__device__ __forceinline__ float sqrt_up(float x) {
  return __builtin_amdgcn_sqrtf(fmaxf(x, FLT_MIN));
}
__device__ __forceinline__ float sqrt_down(float x) {
  return x < FLT_MIN ? 0.0f : __builtin_amdgcn_sqrtf(x);
}

// (the hardware lane counts: lane is the wave lane, threadIdx.x of a one-wave
// workgroup)
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask, int lane) {
  (void)lane;
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Per-lane selection state.
struct SelState {
  uint64_t klo;   // keys < klo are neighbours
  uint64_t khi;   // keys in [klo, khi) are stashed
  uint32_t need;  // 1-based rank of the k-th key inside [klo, khi)
  uint32_t cnt;   // number of candidate keys inside [klo, khi)
};

// Per-lane stash of the keys inside [klo, khi) and their photons' slots, in
// LDS (sh.sb; the lane bins are dead while it is in use).  Keys past
// kStashCap are dropped; the selection keeps the count of the stashed bin at
// or below it.
struct Stash {
  uint32_t n;
  __device__ __forceinline__ void reset() { n = 0; }
  __device__ __forceinline__ void put(GatherShared& sh, int lane, uint64_t k) {
    if (n < static_cast<uint32_t>(kStashCap)) {
      sh.b.sb.key[n * kWave + static_cast<uint32_t>(lane)] = k;
      ++n;
    }
  }
};

// Squared distances from c to the nearest point of the box [mn, mx] (0 when
// c is inside) and to its farthest corner; +inf for an empty box.
__device__ __forceinline__ void box_dist2(const float4& mn, const float4& mx, float cx,
                                          float cy, float cz, float& near2, float& far2) {
  const float ax = fmaxf(fmaxf(mn.x - cx, cx - mx.x), 0.0f);
  const float ay = fmaxf(fmaxf(mn.y - cy, cy - mx.y), 0.0f);
  const float az = fmaxf(fmaxf(mn.z - cz, cz - mx.z), 0.0f);
  near2 = ax * ax + ay * ay + az * az;
  const float fx = fmaxf(cx - mn.x, mx.x - cx);
  const float fy = fmaxf(cy - mn.y, mx.y - cy);
  const float fz = fmaxf(cz - mn.z, mx.z - cz);
  far2 = fx * fx + fy * fy + fz * fz;
}

// Slot range [s0, s1) of the photons under tree node f.
__device__ __forceinline__ void node_slots(DescRef D, uint32_t f, int32_t& s0, int32_t& s1) {
  const uint32_t first_leaf = static_cast<uint32_t>(D.first_leaf);
  const int depth_leaf = 31 - __builtin_clz(first_leaf);
  const int depth = 31 - __builtin_clz(f);
  const int sh = depth_leaf - depth;
  const uint32_t n_leaves = static_cast<uint32_t>(D.n_leaves);
  uint32_t lo = (f << sh) - first_leaf;
  uint32_t hi = ((f + 1u) << sh) - first_leaf;
  lo = lo < n_leaves ? lo : n_leaves;
  hi = hi < n_leaves ? hi : n_leaves;
  const int64_t a = static_cast<int64_t>(lo) * D.leaf_size;
  const int64_t b = static_cast<int64_t>(hi) * D.leaf_size;
  s0 = static_cast<int32_t>(a < D.n ? a : D.n);
  s1 = static_cast<int32_t>(b < D.n ? b : D.n);
}

// Writes the exclusive prefix of the slot counts of the nr list entries in
// sh.rl.rnode to sh.rl.rpre; returns nr, the total slot count in *total.
__device__ uint32_t list_prefix(DescRef D, GatherShared& sh, int lane, uint32_t nr,
                                uint32_t* total) {
  uint32_t run = 0;
  for (uint32_t base = 0; base < nr; base += kWave) {
    const uint32_t i = base + lane;
    uint32_t cnt = 0;
    if (i < nr) {
      int32_t s0, s1;
      node_slots(D, sh.rl.rnode[i], s0, s1);
      cnt = static_cast<uint32_t>(s1 - s0);
    }
    const uint32_t incl = wave_scan_add(cnt);
    if (i < nr) sh.rl.rpre[i] = run + incl - cnt;
    run += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
  }
  __syncthreads();
  *total = static_cast<uint32_t>(uni(static_cast<int>(run)));
  return static_cast<uint32_t>(uni(static_cast<int>(nr)));
}

// Candidate list of the sphere (c, sqrt(r2)), where r2 bounds the squared
// distance (exact arithmetic) of every photon wanted: the subtrees of the
// implicit heap-ordered tree (root 1, children 2f and 2f+1, leaves from
// first_leaf) that meet the sphere, disjoint, holding every photon inside it.
// The walk is breadth-first and expands e levels per step, e as large as
// keeps (frontier << e) within one wave, so one vector load per lane tests up
// to 64 descendants at once.  A leaf, or (unless leaf_only) a node entirely
// inside the sphere, is emitted.  r2 = +inf lists the root.  Writes the
// subtree roots to sh.rl.rnode and the exclusive prefix of their slot counts to
// sh.rl.rpre; returns the number of subtrees, the total slot count in *total.
// When the list would overflow, the walk restarts with whole inner nodes
// allowed (unless give_up is kListLeavesOnly, which returns kNoList and no
// list at once).  If that overflows too, kListGiveUp returns kNoList;
// kListCoarsen restarts the walk with a depth cap above the level that
// overflowed (at most kCapDepth levels: 2^kCapDepth <= kMaxRanges subtrees, so
// it fits).  A frontier that would overflow counts as an overflow when giving
// up, else emits its nodes whole; a walk that still overflows lists the root
// (never a truncated list).  Giving up also returns kNoList for a list of
// more than max_total photons (whole inner nodes can make one large).  *fine (when given)
// tells whether every entry is a single leaf (asked with leaf_only and no
// coarsening was needed).  A returned list coarser than the sphere asks adds
// 1 to *stat (when given), 1000 when it is the whole map.
__device__ uint32_t build_list(DescRef D, float cx, float cy, float cz, float r2,
                               GatherShared& sh, int lane, uint32_t* total, uint32_t* stat,
                               bool leaf_only = false, bool* fine = nullptr,
                               int give_up = kListCoarsen, uint32_t max_total = 0xffffffffu) {
  const bool split = give_up != kListCoarsen;
  const uint32_t first_leaf = static_cast<uint32_t>(D.first_leaf);
  const int depth_leaf = 31 - __builtin_clz(first_leaf);
  uint32_t nr = 0;
  bool whole = !(r2 < INFINITY) || depth_leaf == 0;
  bool coarse = whole;
  int max_depth = depth_leaf;
  const float hit_lim = r2 * 1.00001f;
  const float in_lim = r2 * 0.99999f;
  for (int attempt = 0; attempt < 3 && !whole; ++attempt) {
    nr = 0;
    bool full = false;
    int full_depth = 0;
    if (lane == 0) sh.front[0][0] = 1u;
    __syncthreads();
    uint32_t ncur = 1;
    int cur = 0;
    int depth = 0;
    while (ncur > 0 && !full) {
      int e = 1;
      while (depth + e < max_depth && (ncur << (e + 1)) <= static_cast<uint32_t>(kWave)) ++e;
      const uint32_t width = ncur << e;
      const bool last_lvl = depth + e == max_depth;
      const uint32_t sub = (1u << e) - 1u;
      uint32_t nnext = 0;
      for (uint32_t base = 0; base < width; base += kWave) {
        const uint32_t i = base + lane;
        uint32_t f = 0;
        bool emit = false, push = false;
        if (i < width) {
          f = (sh.front[cur][i >> e] << e) | (i & sub);
          const float4 mn = D.boxes[2 * f];
          const float4 mx = D.boxes[2 * f + 1];
          float n2, f2;
          box_dist2(mn, mx, cx, cy, cz, n2, f2);
          if (n2 <= hit_lim) {
            if (last_lvl || (!leaf_only && f2 <= in_lim)) emit = true;
            else push = true;
          }
        }
        uint64_t em = __ballot(emit);
        uint64_t pmk = __ballot(push);
        const bool front_full =
            nnext + static_cast<uint32_t>(__popcll(pmk)) > static_cast<uint32_t>(kFront);
        if (front_full && !split) {
          emit = emit || push;
          push = false;
          em = __ballot(emit);
          pmk = 0ull;
          coarse = true;
        }
        const uint32_t ne = static_cast<uint32_t>(__popcll(em));
        if ((front_full && split) || nr + ne > static_cast<uint32_t>(kMaxRanges)) {
          full = true;
          full_depth = depth + e;
          break;
        }
        if (emit) sh.rl.rnode[nr + lanes_below(em, lane)] = f;
        if (push) sh.front[cur ^ 1][nnext + lanes_below(pmk, lane)] = f;
        nr += ne;
        nnext += static_cast<uint32_t>(__popcll(pmk));
      }
      __syncthreads();
      cur ^= 1;
      ncur = nnext;
      depth += e;
    }
    if (!full) break;
    if (attempt == 2) {
      whole = true;  // cannot happen with the depth cap; never return a cut list
    } else if (leaf_only && give_up != kListLeavesOnly) {
      leaf_only = false;
    } else if (split) {
      return kNoList;
    } else {
      coarse = true;
      max_depth = min(full_depth - 1, kCapDepth);
      if (max_depth < 1) whole = true;
    }
  }
  if (whole) {
    nr = 1;
    if (lane == 0) sh.rl.rnode[0] = 1u;
    __syncthreads();
  }
  if (stat && lane == 0 && coarse && r2 < INFINITY && depth_leaf > 0) atomicAdd(stat, whole ? 1000u : 1u);
  if (fine) *fine = leaf_only && !coarse && !whole && max_depth == depth_leaf;
  nr = list_prefix(D, sh, lane, nr, total);
  return split && *total > max_total ? kNoList : nr;
}

// Drops the entries of the current list whose boxes miss the sphere
// (c, sqrt(r2)) (r2 as for build_list) and rebuilds the prefix.
__device__ uint32_t filter_list(DescRef D, GatherShared& sh, int lane, uint32_t nr, float cx,
                                float cy, float cz, float r2, uint32_t* total) {
  const float hit_lim = r2 * 1.00001f;
  uint32_t out = 0;
  for (uint32_t base = 0; base < nr; base += kWave) {
    const uint32_t i = base + lane;
    uint32_t f = 0;
    bool keep = false;
    if (i < nr) {
      f = sh.rl.rnode[i];
      float n2, f2;
      box_dist2(D.boxes[2 * f], D.boxes[2 * f + 1], cx, cy, cz, n2, f2);
      keep = n2 <= hit_lim;
    }
    const uint64_t km = __ballot(keep);
    // every lane's entry was read above; writes land below base + kWave
    if (keep) sh.rl.rnode[out + lanes_below(km, lane)] = f;
    out += static_cast<uint32_t>(__popcll(km));
  }
  __syncthreads();
  return list_prefix(D, sh, lane, out, total);
}

// Slot of the t-th candidate of the current list; r is the lane's cursor
// into a subtree list and only moves forward (t must not decrease between
// calls).
__device__ __forceinline__ int32_t list_slot(DescRef D, const GatherShared& sh, uint32_t nr,
                                             uint32_t t, uint32_t& r) {
  if (nr == kSlotList) return static_cast<int32_t>(sh.cslot[t]);
  while (r + 1 < nr && sh.rl.rpre[r + 1] <= t) ++r;
  int32_t s0, s1;
  node_slots(D, sh.rl.rnode[r], s0, s1);
  return s0 + static_cast<int32_t>(t - sh.rl.rpre[r]);
}

// What the candidate loop stages per photon besides its position.
enum Stage : int {
  kStagePos = 0,   // position only
  kStageFull = 1,  // + wi and flux, and the slot
  kStageEff = 2,   // + the flux the run's shared Lambert sign lets through, and the slot
};

// A normal shared (within eps) by the accumulating lanes of a run, and the
// margin term m = eps * 1.0001 + 1e-6 * max |n| of the sign decision.
struct SharedNormal {
  float x, y, z, m;
};

// Iterates the candidate photons of the current list: body(cnt, amb) is
// called once per staged chunk, after sh.pos (and with kStageFull also
// sh.b.wif/flux, with kStageEff the effective flux in sh.b.wif, its w one
// when the sign lets the photon through) hold cnt candidates; amb has bit j
// set for each staged photon j whose sign kStageEff left open.  The next
// chunk's loads are in flight while body runs.
//
// kStageEff decides each photon's Lambert sign once for all lanes: with d0 =
// dot(wi, n0) in the reference's FP32 order, |d0| > |wi|_1 * m implies that
// every lane's own FP32 dot(wi, n) has the sign of d0 (|n - n0| <= eps moves
// the exact dot by at most |wi| eps, and each FP32 dot is within 1.8e-7
// |wi|_1 |n| of exact).  The staged flux is then the photon's (positive sign)
// or zero; otherwise it is zero and the photon is left to the lane's own
// test (its bit in amb).
template <int ST, typename Body>
__device__ __forceinline__ void for_candidates_st(DescRef D, GatherShared& sh, int lane,
                                                  uint32_t nr, uint32_t total,
                                                  const SharedNormal& sn, Body&& body,
                                                  uint32_t begin = 0) {
  if (total <= begin) return;
  constexpr bool kWF = ST != kStagePos;
  const float4* __restrict__ P = D.pos4;
  const float4* __restrict__ W = D.wif4;
  const float2* __restrict__ F = D.flux2;
  uint32_t t = begin + static_cast<uint32_t>(lane);
  uint32_t r = 0;
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f), w = p;
  float2 f = make_float2(0.f, 0.f);
  int32_t slot = 0;
  if (t < total) {
    slot = list_slot(D, sh, nr, t, r);
    p = P[slot];
    if (kWF) {
      w = W[slot];
      f = F[slot];
    }
  }
  for (uint32_t o = begin; o < total; o += kChunk) {
    sh.pos[lane] = p;
    if (ST == kStageFull) {
      sh.b.wif[lane] = w;
      sh.b.flux[lane] = f;
    }
    uint64_t amb_mask = 0ull;
    if (ST == kStageEff) {
      const float d0 = dot_ref(w.x, w.y, w.z, sn.x, sn.y, sn.z);
      const float mw = (fabsf(w.x) + fabsf(w.y) + fabsf(w.z)) * sn.m + 1e-30f;
      const bool amb = !(fabsf(d0) > mw);
      const bool inc = !amb && d0 > 0.0f;
      sh.b.wif[lane] = make_float4(inc ? w.w : 0.0f, inc ? f.x : 0.0f, inc ? f.y : 0.0f,
                                   inc ? 1.0f : 0.0f);
      amb_mask = __ballot(amb && o + static_cast<uint32_t>(lane) < total);
    }
    if (kWF) sh.b.sb.chunk_slot[lane] = static_cast<uint32_t>(slot);  // (aliases the bins)
    __syncthreads();
    t += kChunk;
    if (t < total) {
      slot = list_slot(D, sh, nr, t, r);
      p = P[slot];
      if (kWF) {
        w = W[slot];
        f = F[slot];
      }
    }
    const uint32_t cnt = total - o < static_cast<uint32_t>(kChunk) ? total - o
                                                                    : static_cast<uint32_t>(kChunk);
    body(cnt, amb_mask);
    __syncthreads();
  }
}

template <bool WF, typename Body>
__device__ __forceinline__ void for_candidates(DescRef D, GatherShared& sh, int lane,
                                               uint32_t nr, uint32_t total, Body&& body,
                                               uint32_t begin = 0) {
  const SharedNormal none{0.0f, 0.0f, 0.0f, 0.0f};
  for_candidates_st<WF ? kStageFull : kStagePos>(
      D, sh, lane, nr, total, none, [&](uint32_t cnt, uint64_t) { body(cnt); }, begin);
}

// The normal of the lanes with on set, when they agree: true (and *sn set,
// with its margin term) when every such lane's normal lies within 5% of its
// length of the first one's.
__device__ __forceinline__ bool shared_normal(bool on, float nx, float ny, float nz,
                                              SharedNormal* sn) {
  *sn = SharedNormal{0.0f, 0.0f, 0.0f, 0.0f};
  const uint64_t am = __ballot(on);
  if (am == 0ull) return false;
  const int l0 = __builtin_ctzll(am);
  sn->x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nx), l0));
  sn->y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ny), l0));
  sn->z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nz), l0));
  const float ex = nx - sn->x, ey = ny - sn->y, ez = nz - sn->z;
  const float dev = on ? sqrtf(ex * ex + ey * ey + ez * ez) : 0.0f;
  const float nl = on ? sqrtf(nx * nx + ny * ny + nz * nz) : 0.0f;
  const float eps = wave_max(dev);
  const float nmax = wave_max(nl);
  if (!(eps <= 0.05f * nmax && nmax <= FLT_MAX)) return false;
  sn->m = eps * 1.0001f + 1e-6f * nmax;
  return true;
}

// Lane-parallel sweep of the current list: every lane takes every 64th
// candidate, four loads in flight at a time, and calls fn(u, ok, position,
// slot) for its four candidates of a step with every lane active (ok tells
// whether the lane holds a candidate).
template <typename Fn>
__device__ __forceinline__ void sweep_lanes(DescRef D, const GatherShared& sh, int lane,
                                            uint32_t nr, uint32_t total, Fn&& fn) {
  const float4* __restrict__ P = D.pos4;
  uint32_t r = 0;
  for (uint32_t base = 0; base < total; base += 4 * kWave) {
    float4 p[4];
    int32_t slot[4];
    bool ok[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = base + static_cast<uint32_t>(u * kWave + lane);
      ok[u] = t < total;
      p[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      slot[u] = 0;
      if (ok[u]) {
        slot[u] = list_slot(D, sh, nr, t, r);
        p[u] = P[slot[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) fn(u, ok[u], p[u], slot[u]);
  }
}

// Lane-parallel sweep of the current list like sweep_lanes; when every entry
// is a single leaf (leaves), a lane takes a whole leaf instead, eight
// contiguous slots per call round, so no lane walks the list.
template <typename Fn>
__device__ __forceinline__ void sweep_list(DescRef D, const GatherShared& sh, int lane,
                                           uint32_t nr, uint32_t total, bool leaves, Fn&& fn) {
  if (!leaves) {
    sweep_lanes(D, sh, lane, nr, total, fn);
    return;
  }
  const float4* __restrict__ P = D.pos4;
  const int32_t leaf_size = D.leaf_size;
  for (uint32_t base = 0; base < nr; base += kWave) {
    const uint32_t i = base + static_cast<uint32_t>(lane);
    int32_t s0 = 0, s1 = 0;
    if (i < nr) node_slots(D, sh.rl.rnode[i], s0, s1);
    for (int32_t j0 = 0; j0 < leaf_size; j0 += 8) {
      float4 p[8];
      bool ok[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int32_t sl = s0 + j0 + u;
        ok[u] = sl < s1;
        p[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok[u]) p[u] = P[sl];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) fn(u, ok[u], p[u], s0 + j0 + u);
    }
  }
}

// Explicit slot list of the photons of the current subtree list whose FP32
// d^2 from c is <= lim2.  It is gathered in sh.cscratch and, when it fits,
// copied to sh.cslot, which replaces the subtree list; returns its length.
// When more than kSlotCap photons qualify, returns kSlotList and the subtree
// list is left as it was.
__device__ uint32_t collect_slots(DescRef D, GatherShared& sh, int lane, uint32_t nr,
                                  uint32_t total, bool leaves, float cx, float cy, float cz,
                                  float lim2) {
  uint32_t n = 0;
  bool over = false;
  sweep_list(D, sh, lane, nr, total, leaves, [&](int, bool ok, const float4& p, int32_t slot) {
    const bool keep = ok && d2_ref(cx, cy, cz, p.x, p.y, p.z) <= lim2;
    const uint64_t km = __ballot(keep);
    const uint32_t at = n + lanes_below(km, lane);
    if (keep && at < static_cast<uint32_t>(kSlotCap)) sh.cscratch[at] = static_cast<uint32_t>(slot);
    n += static_cast<uint32_t>(__popcll(km));
    over = over || n > static_cast<uint32_t>(kSlotCap);
  });
  __syncthreads();
  if (over) return kSlotList;
  const uint32_t cnt = static_cast<uint32_t>(uni(static_cast<int>(n)));
  for (uint32_t i = lane; i < cnt; i += kWave) sh.cslot[i] = sh.cscratch[i];
  __syncthreads();
  return cnt;
}

// Keeps the entries of the explicit slot list whose photon has FP32 d^2 from
// c <= lim2, in order, and returns the new length.  Four batches of 64
// entries are loaded before any is tested, so their loads overlap; a batch's
// kept entries land below the next batch, which is read before it can be
// overwritten.
__device__ uint32_t filter_slots(DescRef D, GatherShared& sh, int lane, uint32_t total,
                                 float cx, float cy, float cz, float lim2) {
  const float4* __restrict__ P = D.pos4;
  uint32_t out = 0;
  for (uint32_t base = 0; base < total; base += 4 * kWave) {
    uint32_t sl[4];
    float4 p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = base + static_cast<uint32_t>(u * kWave + lane);
      sl[u] = 0;
      p[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (t < total) {
        sl[u] = sh.cslot[t];
        p[u] = P[sl[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = base + static_cast<uint32_t>(u * kWave + lane);
      const bool keep = t < total && d2_ref(cx, cy, cz, p[u].x, p[u].y, p[u].z) <= lim2;
      const uint64_t km = __ballot(keep);
      if (keep) sh.cslot[out + lanes_below(km, lane)] = sl[u];
      out += static_cast<uint32_t>(__popcll(km));
    }
  }
  __syncthreads();
  return static_cast<uint32_t>(uni(static_cast<int>(out)));
}

// The run's core: the photons of the explicit slot list that lie below every
// lane's window by the triangle inequality, d(c, p) < min over the run's
// lanes of (sqrt(bound_q) - |q - c|) with bound_q a lane's lower key bound
// (its squared distance part).  Such a photon's FP32 d^2 from every lane of
// the run is below that lane's bound (the 1e-4 relative margin covers the
// FP32 roundings), so a histogram pass counts it in every lane's lowest bin
// and a selection pass finds it below klo for every lane, without its
// pairwise d^2.  Returns the squared core radius (0 when there is none).
__device__ __forceinline__ float core_radius2(bool part, float bound_d2, float dq) {
  float cr = INFINITY;
  if (part) cr = (bound_d2 > 0.0f && bound_d2 <= FLT_MAX) ? sqrtf(bound_d2) * 0.9999f - dq : 0.0f;
  cr = wave_min(cr);
  return cr > 0.0f ? cr * cr : 0.0f;
}

// Keeps the entries of the explicit slot list whose photon has FP32 d^2 from
// c <= lim2, the core ones (d^2 from c < core2) first, each part in list
// order; returns the new length and the core's in *n_core.  Core entries are
// compacted in place (an entry never moves up past one not yet read, as in
// filter_slots), the others go through cscratch, which is free whenever a
// slot list is complete.
__device__ uint32_t partition_slots(DescRef D, GatherShared& sh, int lane, uint32_t total,
                                    float cx, float cy, float cz, float lim2, float core2,
                                    uint32_t* n_core) {
  const float4* __restrict__ P = D.pos4;
  uint32_t nc = 0, nk = 0;
  for (uint32_t base = 0; base < total; base += 4 * kWave) {
    uint32_t sl[4];
    float4 p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = base + static_cast<uint32_t>(u * kWave + lane);
      sl[u] = 0;
      p[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (t < total) {
        sl[u] = sh.cslot[t];
        p[u] = P[sl[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = base + static_cast<uint32_t>(u * kWave + lane);
      const float d2 = d2_ref(cx, cy, cz, p[u].x, p[u].y, p[u].z);
      const bool core = t < total && d2 < core2;
      const bool keep = t < total && !core && d2 <= lim2;
      const uint64_t cm = __ballot(core);
      if (core) sh.cslot[nc + lanes_below(cm, lane)] = sl[u];
      nc += static_cast<uint32_t>(__popcll(cm));
      const uint64_t km = __ballot(keep);
      if (keep) sh.cscratch[nk + lanes_below(km, lane)] = sl[u];
      nk += static_cast<uint32_t>(__popcll(km));
    }
  }
  __syncthreads();
  for (uint32_t i = static_cast<uint32_t>(lane); i < nk; i += kWave) sh.cslot[nc + i] = sh.cscratch[i];
  __syncthreads();
  *n_core = static_cast<uint32_t>(uni(static_cast<int>(nc)));
  return static_cast<uint32_t>(uni(static_cast<int>(nc + nk)));
}

// A staged candidate as the consume loops see it.
struct Cand {
  float4 p;   // position, original index bits
  float4 w;   // wi, flux.r   (staged with WF only)
  float2 f;   // flux.g, flux.b
};

template <int ST>
__device__ __forceinline__ Cand load_cand(const GatherShared& sh, uint32_t j) {
  Cand c;
  c.p = sh.pos[j];
  c.w = make_float4(0.f, 0.f, 0.f, 0.f);
  c.f = make_float2(0.f, 0.f);
  if (ST != kStagePos) c.w = sh.b.wif[j];
  if (ST == kStageFull) c.f = sh.b.flux[j];
  return c;
}

// Calls use(cand) for the cnt staged candidates, reading them in blocks
// (8 positions, or 4 records when more is staged) into registers first, so that the
// LDS reads of a block are in flight together.  use returns a per-lane flag;
// the flags form a 64-bit mask over the chunk, which is returned for
// draining.
template <int ST, typename Use>
__device__ __forceinline__ uint64_t consume_st(const GatherShared& sh, uint32_t cnt, Use&& use) {
  constexpr uint32_t kBlk = ST == kStagePos ? 8 : 4;
  uint32_t mlo = 0, mhi = 0;
  uint32_t j = 0;
  for (; j + kBlk <= cnt; j += kBlk) {
    Cand c[kBlk];
#pragma unroll
    for (uint32_t u = 0; u < kBlk; ++u) c[u] = load_cand<ST>(sh, j + u);
    uint32_t bits = 0;
#pragma unroll
    for (uint32_t u = 0; u < kBlk; ++u) bits |= use(c[u]) ? (1u << u) : 0u;
    if (j < 32) mlo |= bits << j;
    else mhi |= bits << (j - 32);
  }
  for (; j < cnt; ++j) {
    const uint32_t bit = use(load_cand<ST>(sh, j)) ? 1u : 0u;
    if (j < 32) mlo |= bit << j;
    else mhi |= bit << (j - 32);
  }
  return (static_cast<uint64_t>(mhi) << 32) | mlo;
}

template <bool WF, typename Use>
__device__ __forceinline__ uint64_t consume(const GatherShared& sh, uint32_t cnt, Use&& use) {
  return consume_st<WF ? kStageFull : kStagePos>(sh, cnt, use);
}

// Calls fn(j) for every set bit j of each lane's mask (per-lane loop; the
// wave runs as many rounds as its busiest lane).
template <typename Fn>
__device__ __forceinline__ void drain(uint64_t m, Fn&& fn) {
  while (wave_any(m != 0ull)) {
    if (m != 0ull) {
      const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(m));
      m &= m - 1ull;
      fn(j);
    }
  }
}

__device__ __forceinline__ void clear_hist(GatherShared& sh, int lane) {
  uint4* h4 = reinterpret_cast<uint4*>(sh.hist);
  for (int i = lane; i < (kBins + 2) * kWave / 4; i += kWave) h4[i] = make_uint4(0, 0, 0, 0);
}

__device__ __forceinline__ void lds_inc(uint32_t* p) {
  __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void hist_add(GatherShared& sh, uint32_t b, int lane) {
  lds_inc(&sh.hist[b * kWave + lane]);
}

// Bracket [t_lo, t_hi] of the k-th smallest FP32 d^2(c, p) over the photons
// of the current list: fewer than k_eff of them have d^2 < t_lo, at least
// k_eff have d^2 < t_hi.  A wave-parallel histogram of kCentreBins linear
// bins over [0, rb); returns false when fewer than k_eff photons fall below
// rb.
__device__ bool centre_bound(DescRef D, GatherShared& sh, int lane, uint32_t nr,
                             uint32_t total, bool leaves, float cx, float cy, float cz,
                             float rb, uint32_t k_eff, float* t_lo, float* t_hi) {
  uint32_t* ch = sh.hist;
  for (int i = lane; i < kCentreBins; i += kWave) ch[i] = 0u;
  __syncthreads();
  const float inv_w = static_cast<float>(kCentreBins) / rb;
  sweep_list(D, sh, lane, nr, total, leaves, [&](int, bool ok, const float4& p, int32_t) {
    const float t = d2_ref(cx, cy, cz, p.x, p.y, p.z) * inv_w;
    if (ok && t < static_cast<float>(kCentreBins)) lds_inc(&ch[static_cast<uint32_t>(t)]);
  });
  __syncthreads();
  static_assert(kCentreBins == 4 * kWave, "four centre bins per lane");
  const uint4 h = reinterpret_cast<const uint4*>(ch)[lane];
  const uint32_t s = h.x + h.y + h.z + h.w;
  const uint32_t incl = wave_scan_add(s);
  const uint64_t hit = __ballot(incl >= k_eff);
  __syncthreads();
  if (hit == 0ull) return false;
  const int fl = __builtin_ctzll(hit);
  uint32_t b = 4u * static_cast<uint32_t>(lane);
  uint32_t cum = incl - s;
  if (cum + h.x < k_eff) {
    cum += h.x;
    ++b;
    if (cum + h.y < k_eff) {
      cum += h.y;
      ++b;
      if (cum + h.z < k_eff) ++b;
    }
  }
  b = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), fl));
  // photons of bins <= b have d^2 * inv_w < b + 1 before rounding of the
  // product, hence d^2 < (b + 1) / inv_w up to the margin; photons with
  // d^2 < b / inv_w (less the margin) all lie in bins < b
  *t_hi = (static_cast<float>(b + 1u) / inv_w) * 1.00001f + 1e-30f;
  *t_lo = (static_cast<float>(b) / inv_w) * 0.99999f;
  return true;
}

// Radius r with at least k_eff photons within r of c (FP32 d^2, up to the
// margins), from leaf lists only, so that no coarse subtree is ever swept.
// cap2 is a squared radius known to hold k_eff photons.  The search starts at
// the squared radius rb; a sphere whose leaf list overflows shrinks (4-fold,
// or to the geometric mean with the largest sphere seen holding fewer than
// k_eff photons), one holding fewer grows; few and over seed those two ends.
// A sphere that holds k_eff photons gives the bracket's upper end, narrowed by
// a second histogram over it.  Returns sqrt(cap2) (rounded up) when 7 lists
// give nothing better.
__device__ float leaf_bracket(DescRef D, GatherShared& sh, int lane, float cx, float cy,
                              float cz, float rb, float few, float over, float cap2,
                              uint32_t k_eff) {
  float r = sqrt_up(cap2) * 1.00002f + 1e-30f;
  for (int attempt = 0; attempt < 7 && rb > 0.0f; ++attempt) {
    uint32_t total = 0;
    bool fine = false;
    const uint32_t nr = build_list(D, cx, cy, cz, rb, sh, lane, &total, nullptr, true, &fine,
                                   kListLeavesOnly);
    if (nr == kNoList || !fine) {
      over = rb;
      rb = few > 0.0f ? sqrtf(few * over) : rb * 0.25f;
      continue;
    }
    float t_lo = 0.0f, t_hi = INFINITY;
    if (centre_bound(D, sh, lane, nr, total, fine, cx, cy, cz, rb, k_eff, &t_lo, &t_hi)) {
      const float rb2 = t_hi;
      r = fminf(r, sqrt_up(rb2) * 1.00002f + 1e-30f);
      if (centre_bound(D, sh, lane, nr, total, fine, cx, cy, cz, rb2, k_eff, &t_lo, &t_hi)) {
        r = fminf(r, sqrt_up(t_hi) * 1.00002f + 1e-30f);
      }
      break;
    }
    if (!(rb < cap2)) break;  // cap2 held its k_eff photons: cannot happen
    few = rb;
    rb = over < INFINITY ? sqrtf(few * over) : fminf(rb * 4.0f, cap2);
  }
  return r;
}

// Wave-cycle accounting of the phases of a group, for statistics runs only
// (acc == nullptr otherwise).  Phases: 0 query and radius bound, 1 centre
// bound, 2 candidate lists, 3 pass A, 4 thresholds and refinement, 5 pass B,
// 6 stash, scatter and outputs, 7 work queue.
struct PhaseTimer {
  uint64_t* acc;
  uint64_t t;
  __device__ __forceinline__ void mark(int phase, int lane) {
    if (acc) {
      const uint64_t now = clock64();
      if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[phase]),
                               static_cast<unsigned long long>(now - t));
      t = now;
    }
  }
};

// Most photons a run's candidate list may hold before the run is halved (or,
// for one lane, its sphere narrowed): the halves cost about what the run
// costs once its sphere holds more than a few k photons per lane, and one
// lane needs more than that only when its bound is loose.
__device__ __forceinline__ uint32_t run_list_cap(bool part, uint32_t k_eff) {
  const uint64_t n_part = static_cast<uint64_t>(__popcll(__ballot(part)));
  const uint64_t per = 2ull * k_eff + 128ull;
  const uint64_t cap = n_part > 1 ? n_part * per : 4ull * per;
  return static_cast<uint32_t>(cap < 0xffffffffull ? cap : 0xffffffffull);
}

// Most photons the list of a run of several lanes may hold in the adjoint
// with reused keys before the run is halved.  Both scatter passes cost about
// (list photons) x (run lanes), and the run's sphere holds every lane's tau
// ball: a lane with a wide ball next to a dense region (a caustic's edge)
// makes the list many times what the other lanes need, while its halves
// keep the wide ball to few lanes.  A list of a compact run holds about
// 2 k_eff photons plus the slack of its leaves.
#ifndef PM_REUSE_CAP_K
#define PM_REUSE_CAP_K 4
#endif
__device__ __forceinline__ uint32_t reuse_list_cap(uint32_t k_eff) {
  if (PM_REUSE_CAP_K <= 0) return 0xffffffffu;
  const uint64_t cap = static_cast<uint64_t>(PM_REUSE_CAP_K) * k_eff + 256ull;
  return static_cast<uint32_t>(cap < 0xffffffffull ? cap : 0xffffffffull);
}

// Selection for the lanes of one run (part), with centre c, lane distances
// dq >= |q - c| and their maximum dmax.  r0sq is the lane's guaranteed bound
// on its k-th FP32 d^2.  Afterwards the neighbour set of a part lane is
// {key < klo} plus the `need` smallest keys of [klo, khi), and [klo, khi)
// holds at most kStashCap keys.  Leaves the run's candidate list (covering
// every key below the lane bounds) in sh.rl.rnode/sh.rl.rpre: returns its entry
// count and its photon count in *total.  When the centre sweep's sphere needs
// more subtrees than a list holds, or holds more than 2 k + 128 photons per
// lane (run_list_cap), a run of several lanes gets kNoList and nothing is selected:
// the caller halves it.  A run of one lane then has a loose bound (its own):
// the sphere is narrowed by leaf_bracket first, so no coarse or oversized
// list is swept.
__device__ uint32_t run_select(DescRef D, float qx, float qy, float qz, bool part, float r0sq,
                               float cx, float cy, float cz, float dq, float dmax,
                               uint32_t k_eff, GatherShared& sh, int lane, uint32_t* counters,
                               int count_stats, SelState& st, PhaseTimer& pt,
                               uint32_t* total) {
  uint32_t* cut_stat = count_stats ? &counters[kStatOverflow] : nullptr;
  // ---- centre bracket: lane q's k nearest photons lie within X_q of q, hence
  // within X_q + |q - c| of c, so the smallest such sum bounds the centre's
  // k-NN radius; a histogram around c then brackets it
  const float xq = part ? sqrt_up(cull_bound(r0sq)) * 1.00002f + dq : INFINITY;
  float r0c = wave_min(xq) * 1.00002f;
  float rc_hi = INFINITY, rc_lo = 0.0f;
  uint32_t nr = 0;
  bool fine = false;
  if (r0c < INFINITY) {
    // The list of the centre sweep also covers the run's final sphere, which
    // is then a filter of it: that sphere reaches no farther than r_c + 2
    // dmax (r_c <= r0c up to the margins), nor than the largest lane bound
    // plus its |q - c| (each lane's upper end is at most its own bound).
    const float xmax = wave_max(part ? xq : 0.0f);
    // A run of one lane whose list overflows (or holds more than
    // run_list_cap photons) has a loose bound (its own, c = q, dmax = 0): its
    // sphere is narrowed by a geometric search (halved while the list
    // overflows, then bisected in log radius between the largest sphere seen
    // holding fewer than k photons and the smallest that overflowed), when
    // a leaf list can hold k photons at all.  After kNarrowTries lists, or
    // without that, the lane's own bound gets a coarse list.  Each pass of
    // the loop either ends it or moves on, and the coarse pass always ends it.
    const float r0c_in = r0c;
    const bool narrowable =
        2ull * k_eff <= static_cast<uint64_t>(kMaxRanges) * static_cast<uint64_t>(D.leaf_size);
    float few = 0.0f, over = INFINITY;  // radii: fewer than k photons / list overflowed
    bool narrowing = false;
    int give = kListGiveUp;
    for (int t = 0;; ++t) {
      const float rl = fminf((r0c * 1.0001f + 2.0f * dmax) * 1.0004f, xmax * 1.0005f) + 2e-30f;
      nr = build_list(D, cx, cy, cz, rl * rl + 1e-30f, sh, lane, total, cut_stat, true, &fine,
                      give, run_list_cap(part, k_eff));
      if (nr == kNoList) {
        if (__popcll(__ballot(part)) > 1) return kNoList;
        if (t == 0 && count_stats && lane == 0) atomicAdd(&counters[kStatNarrowed], 1u);
        over = r0c;
        narrowing = narrowable && t < kNarrowTries;
        if (narrowing) {
          r0c = few > 0.0f ? sqrtf(few * over) : 0.5f * r0c;
        } else {
          r0c = r0c_in;
          give = kListCoarsen;
        }
        continue;
      }
      float t_lo = 0.0f, t_hi = INFINITY;
      if (centre_bound(D, sh, lane, nr, *total, fine, cx, cy, cz, r0c * r0c * 1.0001f + 1e-30f,
                       k_eff, &t_lo, &t_hi)) {
        rc_hi = sqrt_up(t_hi) * 1.00001f + 1e-30f;
        rc_lo = sqrt_down(t_lo) * 0.99999f;
        break;
      }
      if (!narrowing) break;  // the lane's own bound: no bracket
      few = r0c;  // narrowed too far
      narrowing = t < kNarrowTries;
      if (narrowing) {
        r0c = sqrtf(few * over);
      } else {
        r0c = r0c_in;
        give = kListCoarsen;
      }
    }
    if (count_stats && lane == 0) atomicAdd(&counters[kStatCand1], *total);
  }
  pt.mark(1, lane);

  // ---- lane brackets and the run's sphere ---------------------------------------
  // k >= k photons lie within rc_hi of c, so within rc_hi + |q - c| of q;
  // fewer than k lie within rc_lo of c, so fewer than k within
  // rc_lo - |q - c| of q.  Every neighbour of every lane lies within
  // rc_hi + 2 dmax of c.
  // The lane's own radius bound r0sq (from the anchors) may be the smaller
  // upper end; the run's sphere then only has to reach every lane's.  Only
  // with a centre bracket: valid lane bounds always give one, and without it
  // the run falls back to unbounded brackets, so that the neighbour sets
  // never depend on the lane bounds being right.
  const bool capped = rc_hi < INFINITY;
  const float rq = (rc_hi + dq) * 1.00002f;
  float hi2 = rq * rq * 1.0001f + 1e-30f;
  if (part && capped) hi2 = fminf(hi2, r0sq * 1.0001f + 1e-30f);
  const float lo = fmaxf(rc_lo - dq, 0.0f) * 0.99999f;
  float lo2 = part ? lo * lo * 0.99999f : 0.0f;
  const float xs = part ? sqrt_up(hi2) * 1.00002f + dq : 0.0f;
  const float xsm = wave_max(xs);
  const float rs = capped ? fminf((rc_hi + 2.0f * dmax) * 1.0002f, xsm * 1.0002f) : INFINITY;
  float rs2 = rs * rs + 1e-30f;

  st.klo = 0; st.khi = 0; st.need = 0; st.cnt = 0;
  uint32_t n_core = 0;  // leading core photons of the slot list (pass A skips them)
  for (int attempt = 0; attempt < 2; ++attempt) {
    n_core = 0;
    uint32_t ns = kSlotList;
    if (attempt == 0 && rs2 < INFINITY) {
      // exactly the photons of the run's sphere, when they fit
      ns = collect_slots(D, sh, lane, nr, *total, fine, cx, cy, cz, rs2 * 1.00001f);
    }
    if (ns != kSlotList) {
      nr = kSlotList;
      *total = ns;
    } else if (attempt == 0 && rs2 < INFINITY && fine) {
      nr = filter_list(D, sh, lane, nr, cx, cy, cz, rs2, total);
    } else {
      // (a run of several lanes is halved rather than given a coarse list)
      const bool halve = attempt == 0 && __popcll(__ballot(part)) > 1;
      nr = build_list(D, cx, cy, cz, rs2, sh, lane, total, cut_stat, false, nullptr,
                      halve ? kListGiveUp : kListCoarsen, run_list_cap(part, k_eff));
      if (nr == kNoList) return kNoList;
    }
    if (count_stats && lane == 0) atomicAdd(&counters[kStatCand2], *total);
    pt.mark(2, lane);
    // ---- pass A: histogram of d^2 over the lane bracket [lo2, hi2) ---------------
    float inv_w = 0.0f, c1 = 64.0f;  // lanes outside the run: always bin kBins + 1
    if (part) {
      inv_w = static_cast<float>(kBins) / (hi2 - lo2);
      c1 = 1.0f - lo2 * inv_w;
      if (!(inv_w < INFINITY) || !(fabsf(c1) < INFINITY)) {
        inv_w = 0.0f;  // unbounded (or degenerate) bracket: every key in bin 1
        c1 = 1.0f;
      }
    }
    if (nr == kSlotList) {
      // the core goes first and counts in every lane's lowest bin unbinned
      // (a lane with a degenerate bracket has none)
      const float core2 = core_radius2(part, inv_w > 0.0f ? lo2 : 0.0f, dq);
      if (core2 > 0.0f)
        *total = partition_slots(D, sh, lane, *total, cx, cy, cz, INFINITY, core2, &n_core);
    }
    clear_hist(sh, lane);
    __syncthreads();
    for_candidates<false>(D, sh, lane, nr, *total, [&](uint32_t cnt) {
      consume<false>(sh, cnt, [&](const Cand& c) {
        hist_add(sh, bin_of(d2_ref(qx, qy, qz, c.p.x, c.p.y, c.p.z), inv_w, c1), lane);
        return false;
      });
    }, n_core);
    if (part) sh.hist[lane] += n_core;  // bin 0
    __syncthreads();
    pt.mark(3, lane);
    uint32_t cum = 0, below = 0, cb = 0, bsel = kBins + 1;
#pragma unroll
    for (int j = 0; j <= kBins; ++j) {
      const uint32_t h = sh.hist[j * kWave + lane];
      if (bsel == kBins + 1 && cum + h >= k_eff) {
        bsel = j;
        below = cum;
        cb = h;
      }
      cum += h;
    }
    // the k-th key must fall inside the bracket (bins 1..kBins)
    const bool lane_ok = !part || (bsel >= 1 && bsel <= kBins);
    if (!wave_any(!lane_ok)) {
      if (part) {
        st.klo = static_cast<uint64_t>(bin_start_bits(bsel, inv_w, c1)) << 32;
        st.khi = static_cast<uint64_t>(bin_start_bits(bsel + 1, inv_w, c1)) << 32;
        st.need = k_eff - below;
        st.cnt = cb;
      }
      break;
    }
    // The brackets are guaranteed, so this is a guard only: retry the lanes
    // that missed unbounded over every photon.
    if (count_stats && lane == 0) atomicAdd(&counters[kStatFallback], 1u);
    if (!lane_ok) {
      hi2 = INFINITY;
      lo2 = 0.0f;
    }
    rs2 = INFINITY;
  }

  // ---- refinement: split [klo, khi) until it holds <= kStashCap keys --------
  // Each round shrinks the key span of a refining lane 32-fold, so 64-bit
  // spans end within 13 rounds; the cap only bounds the loop if the counts
  // were ever inconsistent.
  for (int round = 0;
       round < 16 && wave_any(part && st.cnt > static_cast<uint32_t>(kStashCap)); ++round) {
    if (count_stats && lane == 0) atomicAdd(&counters[kStatRefine], 1u);
    const bool mine = part && st.cnt > static_cast<uint32_t>(kStashCap);
    const uint64_t span = st.khi - st.klo;
    const uint64_t w = mine ? (span + kBins - 1) / kBins : 1ull;
    clear_hist(sh, lane);
    __syncthreads();
    for_candidates<false>(D, sh, lane, nr, *total, [&](uint32_t cnt) {
      for (uint32_t j = 0; j < cnt; ++j) {
        const float4 p = sh.pos[j];
        const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
        const uint64_t key = make_key(d2, __float_as_uint(p.w));
        if (mine && key >= st.klo && key < st.khi) {
          hist_add(sh, static_cast<uint32_t>((key - st.klo) / w), lane);
        }
      }
    }, n_core);  // the core lies below every lane's window
    __syncthreads();
    if (mine) {
      uint32_t cum = 0;
      for (int j = 0; j < kBins; ++j) {
        const uint32_t h = sh.hist[j * kWave + lane];
        if (cum + h >= st.need) {
          const uint64_t nlo = st.klo + static_cast<uint64_t>(j) * w;
          const uint64_t nhi_raw = nlo + w;
          st.khi = (nhi_raw < st.khi && nhi_raw > nlo) ? nhi_raw : st.khi;
          st.klo = nlo;
          st.need -= cum;
          st.cnt = h;
          break;
        }
        cum += h;
      }
    }
    __syncthreads();
  }
  pt.mark(4, lane);
  return nr;
}

// The flux of the run's core (the first n entries of the slot list, each
// below every lane's klo) that passes the Lambert sign test of a lane with on
// set, added to (ar, ag, ab): the sign decided once with the shared normal
// as kStageEff decides it, photons it leaves open by each lane's own test.
__device__ void add_core_flux(DescRef D, const GatherShared& sh, int lane, uint32_t n,
                              const SharedNormal& sn, bool on, float nx, float ny, float nz,
                              float& ar, float& ag, float& ab) {
  float sr = 0.0f, sg = 0.0f, sb = 0.0f;
  for (uint32_t b = 0; b < n; b += kWave) {
    const uint32_t t = b + static_cast<uint32_t>(lane);
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 f = make_float2(0.f, 0.f);
    if (t < n) {
      const uint32_t s = sh.cslot[t];
      w = D.wif4[s];
      f = D.flux2[s];
    }
    const float d0 = dot_ref(w.x, w.y, w.z, sn.x, sn.y, sn.z);
    const float mw = (fabsf(w.x) + fabsf(w.y) + fabsf(w.z)) * sn.m + 1e-30f;
    const bool amb = t < n && !(fabsf(d0) > mw);
    if (t < n && !amb && d0 > 0.0f) {
      sr += w.w;
      sg += f.x;
      sb += f.y;
    }
    for (uint64_t am = __ballot(amb); am != 0ull; am &= am - 1ull) {
      const uint32_t s = sh.cslot[b + static_cast<uint32_t>(__builtin_ctzll(am))];
      const float4 wj = D.wif4[s];
      const float2 fj = D.flux2[s];
      if (on && !(dot_ref(wj.x, wj.y, wj.z, nx, ny, nz) < 0.0f)) {
        ar = ar + wj.w;
        ag = ag + fj.x;
        ab = ab + fj.y;
      }
    }
  }
  sr = wave_sum(sr);
  sg = wave_sum(sg);
  sb = wave_sum(sb);
  if (on) {
    ar += sr;
    ag += sg;
    ab += sb;
  }
}

template <int K>
using Kind = std::integral_constant<int, K>;

template <int MODE>
__device__ void process_group(DescRef D, int64_t lg, const uint32_t* __restrict__ perm,
                              GatherShared& sh, float* drho, int lane, uint32_t* counters,
                              int count_stats, float split, PhaseTimer& pt) {
  const int64_t qloc = lg * kWave + lane;
  const bool in_range = qloc < D.m;
  const uint32_t k_eff = static_cast<uint32_t>(D.k_eff);
  pt.mark(7, lane);

  // ---- the lane's query ----------------------------------------------------
  int64_t qi = 0;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  float nx = 0.f, ny = 0.f, nz = 0.f;
  float cos_o = 0.f;
  bool lambert = false;
  int32_t mat = -1;
  float rho_r = 0.f, rho_g = 0.f, rho_b = 0.f;
  const bool reuse = (MODE == kModeAdjoint) && D.key_in != nullptr;
  const bool all = D.all_members != 0;
  // The anchors around the lane's query depend only on its sorted position,
  // so they are loaded before its record: both chains of loads are then in
  // flight together.
  constexpr int kAn = 2 * kAnchorWindow + 1;
  float4 anq[kAn];
  if (!reuse && !all) {
    const float4* __restrict__ A = D.anchors;
    const int32_t center = static_cast<int32_t>(D.q_anchor[in_range ? qloc : lg * kWave]);
    const int32_t a0 = max(center - kAnchorWindow, 0);
    const int32_t a1 = min(center + kAnchorWindow, D.n_anchors - 1);
#pragma unroll
    for (int o = 0; o < kAn; ++o) anq[o] = A[min(max(center + o - kAnchorWindow, a0), a1)];
  }
  {
    const int64_t sp = D.q_base + (in_range ? qloc : lg * kWave);
    qi = static_cast<int64_t>(perm[sp]) - D.q_base;
    if (MODE == kModeKnn) {
      qx = D.q_soa[qi];
      qy = D.q_soa[D.m + qi];
      qz = D.q_soa[2 * D.m + qi];
    } else {
      const pm_query qr = D.q[qi];
      qx = qr.pos[0]; qy = qr.pos[1]; qz = qr.pos[2];
      nx = qr.ns[0]; ny = qr.ns[1]; nz = qr.ns[2];
      cos_o = dot_ref(qr.wo[0], qr.wo[1], qr.wo[2], nx, ny, nz);
      if (qr.material >= 0 && qr.material < D.n_mats) {
        const pm_material mt = D.mats[qr.material];
        lambert = mt.type == PM_MATERIAL_LAMBERT;
        mat = qr.material;
        rho_r = mt.rho[0]; rho_g = mt.rho[1]; rho_b = mt.rho[2];
      }
    }
  }
  // Queries outside the supported domain (a non-finite coordinate or one
  // beyond +-2^60) were sorted after the valid ones of their instance, so
  // the active lanes of a group stay a prefix of it.
  const bool active = in_range && fabsf(qx) <= kCoordLimit && fabsf(qy) <= kCoordLimit &&
                      fabsf(qz) <= kCoordLimit;

  // ---- radius bound: r0sq >= the lane's k-th FP32 d^2 --------------------------
  uint64_t tau_in = 0;
  float r0sq = -1.0f;  // inactive lanes never match anything
  if (reuse) {
    // the forward pass's key is the largest key of the neighbour set
    if (active) tau_in = D.key_in[qi];
    r0sq = __uint_as_float(static_cast<uint32_t>(tau_in >> 32));
  } else if (all) {
    r0sq = INFINITY;
  } else {
    // Triangle inequality on precomputed anchors: the k nearest photons of
    // anchor a lie within r_a of its centroid c_a, hence within
    // r_a + |q - c_a| of q.  The anchor radii are already rounded upward.
    float best = INFINITY;
#pragma unroll
    for (int o = 0; o < kAn; ++o) {
      const float4 an = anq[o];
      const float dx = qx - an.x, dy = qy - an.y, dz = qz - an.z;
      best = fminf(best, an.w + sqrt_up(dx * dx + dy * dy + dz * dz));
    }
    best *= 1.00002f;
    r0sq = best * best * 1.00002f + 1e-30f;
  }
  if (!active) r0sq = -1.0f;
  pt.mark(0, lane);

  constexpr bool kAccum = MODE == kModeRadiance;
  constexpr bool kRows = MODE == kModeKnn;
  // the estimate only counts photons when the outgoing side passes
  const bool lane_lambert = lambert && !(cos_o < 0.0f);
  float acc_r = 0.f, acc_g = 0.f, acc_b = 0.f;
  uint32_t n_written = 0;
  uint64_t key_max = 0;
  uint64_t tau = 0;
  int32_t* row = nullptr;
  if (kRows) row = D.out_slots + qi * D.k;

  const uint32_t n_act = static_cast<uint32_t>(__popcll(__ballot(active)));
  if (k_eff > 0 && n_act > 0) {
    // ---- runs of nearby lanes ------------------------------------------------
    const float scale = all ? INFINITY
                            : wave_min(active ? sqrt_up(cull_bound(r0sq)) : INFINITY);
    // A group whose queries fall apart into several runs (sparse queries)
    // pays a run's fixed cost often for few lanes: its runs are cut at a
    // larger spread, trading candidates per run for fewer runs (measured on
    // uneven workloads; neutral on dense ones, whose groups stay whole).
    float thr = split * scale;
    const float px = __shfl_up(qx, 1, kWave);
    const float py = __shfl_up(qy, 1, kWave);
    const float pz = __shfl_up(qz, 1, kWave);
    const float gap = fmaxf(fabsf(qx - px), fmaxf(fabsf(qy - py), fabsf(qz - pz)));
    uint64_t starts = __ballot(active && lane > 0 && gap > thr) | 1ull;
    if (__popcll(starts) >= kSparseRuns) {
      thr *= kSparseSplit;
      starts = __ballot(active && lane > 0 && gap > thr) | 1ull;
    }
    uint32_t c0 = 0;
    while (c0 < n_act) {
      const uint64_t later = c0 >= 63 ? 0ull : (starts >> (c0 + 1)) << (c0 + 1);
      uint32_t c1 = later ? static_cast<uint32_t>(__builtin_ctzll(later)) : n_act;
      c1 = c1 < n_act ? c1 : n_act;
      const bool part = static_cast<uint32_t>(lane) >= c0 && static_cast<uint32_t>(lane) < c1;
      const float lx = wave_min(part ? qx : INFINITY);
      const float ly = wave_min(part ? qy : INFINITY);
      const float lz = wave_min(part ? qz : INFINITY);
      const float hx = wave_max(part ? qx : -INFINITY);
      const float hy = wave_max(part ? qy : -INFINITY);
      const float hz = wave_max(part ? qz : -INFINITY);
      const float ext = fmaxf(hx - lx, fmaxf(hy - ly, hz - lz));
      if (c1 - c0 > 1 && ext > thr) {
        starts |= 1ull << ((c0 + c1) >> 1);
        continue;
      }
      const float cx = lx + 0.5f * (hx - lx);
      const float cy = ly + 0.5f * (hy - ly);
      const float cz = lz + 0.5f * (hz - lz);
      const float dq = part ? sqrt_up(d2_ref(qx, qy, qz, cx, cy, cz)) * 1.00002f + 1e-30f : 0.0f;
      const float dmax = wave_max(dq);

      // ---- selection state and candidate list of the run --------------------
      Stash stash;  // a lane stashes only in its own run
      stash.reset();
      SelState st;
      uint32_t total = 0;
      uint32_t nr = 0;
      // When the accumulating lanes' normals agree within 5%, each photon's
      // Lambert sign is decided once per chunk (kStageEff) instead of per
      // lane; photons it leaves open take the lane's own test in the drain.
      const bool acc_lane = part && lane_lambert;
      SharedNormal sn{0.0f, 0.0f, 0.0f, 0.0f};
      bool shared = false;
      uint32_t n_core = 0;  // leading core photons of the slot list (pass B skips them)
      if (reuse) {
        st.klo = tau_in + 1; st.khi = tau_in + 1; st.need = 0; st.cnt = 0;
        const float xs = part ? sqrt_up(cull_bound(r0sq)) * 1.00002f + dq : 0.0f;
        const float rs = wave_max(xs) * 1.0002f;
        const float rs2 = rs * rs + 1e-30f;
        bool fine = false;
        const bool several = c1 - c0 > 1;
        nr = build_list(D, cx, cy, cz, rs2, sh, lane, &total,
                        count_stats ? &counters[kStatOverflow] : nullptr, true, &fine,
                        several ? kListGiveUp : kListCoarsen,
                        several ? reuse_list_cap(k_eff) : 0xffffffffu);
        if (nr == kNoList) {
          if (count_stats && lane == 0) atomicAdd(&counters[kStatSplit], 1u);
          starts |= 1ull << ((c0 + c1) >> 1);
          continue;
        }
        if (count_stats && lane == 0) atomicAdd(&counters[kStatCand1], total);
        // both scatter passes read the list: keep exactly the photons of
        // the run's sphere when they fit the slot list
        const uint32_t ns = collect_slots(D, sh, lane, nr, total, fine, cx, cy, cz,
                                          rs2 * 1.00001f);
        if (ns != kSlotList) {
          nr = kSlotList;
          total = ns;
        }
        pt.mark(2, lane);
      } else if (all) {
        // every photon is a neighbour; pass B tracks the largest key
        st.klo = kKeyMax; st.khi = kKeyMax; st.need = 0; st.cnt = 0;
        nr = build_list(D, cx, cy, cz, INFINITY, sh, lane, &total, nullptr);
        pt.mark(2, lane);
      } else {
        nr = run_select(D, qx, qy, qz, part, r0sq, cx, cy, cz, dq, dmax, k_eff, sh, lane,
                        counters, count_stats, st, pt, &total);
        // A run whose sphere needs more subtrees than a list holds would
        // sweep coarse subtrees far beyond it: its halves go separately.
        if (nr == kNoList) {
          if (count_stats && lane == 0) atomicAdd(&counters[kStatSplit], 1u);
          starts |= 1ull << ((c0 + c1) >> 1);
          continue;
        }
        // With the lane windows known, every key below a lane's khi lies
        // within sqrt(d^2 of khi) + |q - c| of c: the slot list shrinks to
        // the sphere that holds them all before the later passes read it.
        shared = kAccum && shared_normal(acc_lane, nx, ny, nz, &sn);
        if (nr == kSlotList) {
          const float kd2 = __uint_as_float(static_cast<uint32_t>(st.khi >> 32));
          float xr = 0.0f;
          if (part) xr = kd2 <= FLT_MAX ? sqrt_up(cull_bound(kd2)) * 1.00002f + dq : INFINITY;
          const float rb = wave_max(xr) * 1.0002f;
          const float lim2 = rb < INFINITY ? rb * rb + 1e-30f : INFINITY;
          // The core, below every lane's klo, is put first: pass B adds its
          // flux once for the run when the sign test is shared (the adjoint's
          // selection has nothing to do with it at all).
          float core2 = 0.0f;
          if (shared || MODE == kModeAdjoint) {
            core2 = core_radius2(part, __uint_as_float(static_cast<uint32_t>(st.klo >> 32)), dq);
          }
          if (core2 > 0.0f) {
            total = partition_slots(D, sh, lane, total, cx, cy, cz, lim2, core2, &n_core);
          } else if (rb < INFINITY) {
            total = filter_slots(D, sh, lane, total, cx, cy, cz, lim2);
          }
          if (count_stats && lane == 0) atomicAdd(&counters[kStatCand3], total);
        }
      }

      if (count_stats && lane == 0) atomicAdd(&counters[kStatRuns], 1u);

      // ---- pass B: consume neighbours below klo, stash [klo, khi) ---------------
      // (the adjoint runs it only to finish the selection; its scatter pass
      // follows once tau is known)
      if (!reuse) {
        auto pass_b = [&](auto kind_tag, auto shared_tag) {
          constexpr int kind = decltype(kind_tag)::value;  // 0 all, 1 keys, 2 d^2 bits
          constexpr bool kShared = decltype(shared_tag)::value;
          constexpr int kSt = kShared ? kStageEff : (kAccum ? kStageFull : kStagePos);
          const uint64_t span = st.khi - st.klo;
          const uint32_t lo32 = static_cast<uint32_t>(st.klo >> 32);
          const uint32_t span32 = static_cast<uint32_t>(st.khi >> 32) - lo32;
          if (kShared && n_core > 0u) {
            add_core_flux(D, sh, lane, n_core, sn, acc_lane, nx, ny, nz, acc_r, acc_g, acc_b);
          }
          for_candidates_st<kSt>(D, sh, lane, nr, total, sn, [&](uint32_t cnt, uint64_t amb) {
            const uint64_t m = consume_st<kSt>(sh, cnt, [&](const Cand& c) {
              const float d2 = d2_ref(qx, qy, qz, c.p.x, c.p.y, c.p.z);
              bool nb, flag;
              if constexpr (kind == 0) {
                nb = true;
                const uint64_t key = make_key(d2, __float_as_uint(c.p.w));
                key_max = (part && key > key_max) ? key : key_max;
                flag = kRows && part;
              } else if constexpr (kind == 1) {
                const uint64_t key = make_key(d2, __float_as_uint(c.p.w));
                nb = key < st.klo;
                flag = part && ((key - st.klo) < span || (kRows && nb));
              } else {
                const uint32_t db = __float_as_uint(d2);
                nb = db < lo32;
                flag = part && ((db - lo32) < span32 || (kRows && nb));
              }
              if constexpr (kShared) {
                const float sel = (acc_lane && nb) ? 1.0f : 0.0f;
                acc_r = fmaf(sel, c.w.x, acc_r);
                acc_g = fmaf(sel, c.w.y, acc_g);
                acc_b = fmaf(sel, c.w.z, acc_b);
              } else if (kAccum) {
                const float cos_i = dot_ref(c.w.x, c.w.y, c.w.z, nx, ny, nz);
                const float sel = (acc_lane && nb && !(cos_i < 0.0f)) ? 1.0f : 0.0f;
                acc_r = fmaf(sel, c.w.w, acc_r);
                acc_g = fmaf(sel, c.f.x, acc_g);
                acc_b = fmaf(sel, c.f.y, acc_b);
              }
              return flag;
            });
            drain(m, [&](uint32_t j) {
              const float4 p = sh.pos[j];
              const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
              const uint32_t idx = __float_as_uint(p.w);
              const uint64_t key = make_key(d2, idx);
              if (kRows && key < st.klo) {
                if (n_written < static_cast<uint32_t>(D.k))
                  row[n_written++] = static_cast<int32_t>(idx);
              } else {
                stash.put(sh, lane, key);
              }
            });
            // photons whose sign the shared normal left open (rare): each
            // neighbouring lane's own test
            while (kShared && amb != 0ull) {
              const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(amb));
              amb &= amb - 1ull;
              const float4 p = sh.pos[j];
              const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
              if (acc_lane && make_key(d2, __float_as_uint(p.w)) < st.klo) {
                const uint32_t s = sh.b.sb.chunk_slot[j];
                const float4 w = D.wif4[s];
                const float2 f = D.flux2[s];
                if (!(dot_ref(w.x, w.y, w.z, nx, ny, nz) < 0.0f)) {
                  acc_r = acc_r + w.w;
                  acc_g = acc_g + f.x;
                  acc_b = acc_b + f.y;
                }
              }
            }
          }, n_core);
        };
        using No = std::false_type;
        using Yes = std::true_type;
        if (all) {
          pass_b(Kind<0>{}, No{});
        } else if (wave_any(part && ((st.klo | st.khi) & 0xffffffffull) != 0ull)) {
          if (shared) pass_b(Kind<1>{}, Yes{});
          else pass_b(Kind<1>{}, No{});
        } else {
          if (shared) pass_b(Kind<2>{}, Yes{});
          else pass_b(Kind<2>{}, No{});
        }
        pt.mark(5, lane);
      }

      // ---- finish the set inside the stash -----------------------------------
      // (wave-wide steps run with every lane active; only part lanes hold
      // stash entries)
      if (part) tau = reuse ? tau_in : (all ? key_max : tau);
      if (!reuse && !all) {
        // rank of each stashed key among the stash; the need smallest are
        // neighbours, the need-th is tau (unused entries hold kKeyMax)
        uint32_t members = 0;
        const uint32_t n_max = static_cast<uint32_t>(uni(static_cast<int>(
            wave_max(part ? static_cast<float>(stash.n) : 0.0f))));
        uint64_t sk[kStashCap];  // the lane's stash; unused entries hold kKeyMax
#pragma unroll
        for (int e = 0; e < kStashCap; ++e) {
          sk[e] = static_cast<uint32_t>(e) < stash.n ? sh.b.sb.key[e * kWave + lane] : kKeyMax;
        }
#pragma unroll 1
        for (uint32_t i = 0; i < n_max; ++i) {
          // (past the lane's own entries this is a stale key; it is not used)
          const uint64_t ki = sh.b.sb.key[i * kWave + lane];
          uint32_t rank = 0;
#pragma unroll
          for (int j = 0; j < kStashCap; ++j) rank += sk[j] < ki ? 1u : 0u;
          const bool valid = part && i < stash.n;
          if (valid && rank + 1 == st.need) tau = ki;
          if (valid && rank < st.need) members |= 1u << i;
        }
        if (kAccum) {
          // members' photons, four per round so their loads overlap
          constexpr int kRound = 4;
          uint32_t mm = members;
          while (wave_any(mm != 0u)) {
            uint32_t im[kRound];
            float4 wm[kRound];
            float2 fm[kRound];
#pragma unroll
            for (int u = 0; u < kRound; ++u) {
              im[u] = mm ? static_cast<uint32_t>(__builtin_ctz(mm)) : 32u;
              mm &= mm - 1u;
              wm[u] = make_float4(0.f, 0.f, 0.f, 0.f);
              fm[u] = make_float2(0.f, 0.f);
              if (im[u] < 32u) {
                const uint32_t idx = static_cast<uint32_t>(sh.b.sb.key[im[u] * kWave + lane]);
                const int32_t sl = D.inv_perm[idx];
                wm[u] = D.wif4[sl];
                fm[u] = D.flux2[sl];
              }
            }
#pragma unroll
            for (int u = 0; u < kRound; ++u) {
              const float cs = dot_ref(wm[u].x, wm[u].y, wm[u].z, nx, ny, nz);
              const float sel = (im[u] < 32u && lane_lambert && !(cs < 0.0f)) ? 1.0f : 0.0f;
              acc_r = fmaf(sel, wm[u].w, acc_r);
              acc_g = fmaf(sel, fm[u].x, acc_g);
              acc_b = fmaf(sel, fm[u].y, acc_b);
            }
          }
        } else if (kRows) {
#pragma unroll
          for (int i = 0; i < kStashCap; ++i) {
            if (((members >> i) & 1u) && n_written < static_cast<uint32_t>(D.k))
              row[n_written++] = static_cast<int32_t>(sk[i] & 0xffffffffull);
          }
        }
      }

      if constexpr (MODE == kModeAdjoint) {
        // ---- scatter: dL/dflux and dL/drho over the neighbour set {key <= tau} --------
        const float r2 = __uint_as_float(static_cast<uint32_t>(tau >> 32));
        const float denom = (D.n_emit * kPiF) * r2;
        // a query contributes only when its own outgoing side passes
        const bool contributes = part && lane_lambert;
        float g_r = 0.f, g_g = 0.f, g_b = 0.f;
        float fg_r = 1.f, fg_g = 1.f, fg_b = 1.f;
        int32_t fgm = -1;
        if (contributes) {
          g_r = D.dl_drgb[3 * qi + 0];
          g_g = D.dl_drgb[3 * qi + 1];
          g_b = D.dl_drgb[3 * qi + 2];
          if (D.fg_mat) {
            fgm = D.fg_mat[qi];
            if (fgm >= 0 && fgm < D.n_mats) {
              const pm_material fm = D.mats[fgm];
              fg_r = fm.rho[0]; fg_g = fm.rho[1]; fg_b = fm.rho[2];
            } else {
              fgm = -1;
            }
          }
        }
        // gradient reaching the estimate: scaled by the primary albedo of a
        // final-gathering bounce
        const float ge_r = g_r * fg_r, ge_g = g_g * fg_g, ge_b = g_b * fg_b;
        const float c_r = ((rho_r / kPiF) * ge_r) / denom;
        const float c_g = ((rho_g / kPiF) * ge_g) / denom;
        const float c_b = ((rho_b / kPiF) * ge_b) / denom;

        // One sweep over the run's candidates gives (1) each lane's flux sum
        // S of its estimate, for dL/drho, and (2) per staged chunk, which
        // contributing queries hold each of its photons: lane q's flags over
        // the chunk (the same FP32 key and sign tests as the estimate) go to
        // LDS as row q, and lane j adds up the coefficients of the queries
        // whose row has bit j, in lane order, then makes one atomic per
        // channel for photon j of the chunk.
        float s_r = 0.f, s_g = 0.f, s_b = 0.f;
        const uint32_t t_hi = static_cast<uint32_t>(tau >> 32);
        const uint32_t t_lo = static_cast<uint32_t>(tau & 0xffffffffull);
        // (the Lambert sign is decided per chunk when the contributing
        // lanes share their normal, as in the forward selection pass)
        SharedNormal sn;
        const bool shared = shared_normal(contributes, nx, ny, nz, &sn);
        // the rows and the coefficients live where the stash was: the stash
        // is no longer read, and the chunk staging does not touch it
        float4* qcoef = reinterpret_cast<float4*>(sh.b.sb.key);
        uint32_t* rows32 = reinterpret_cast<uint32_t*>(qcoef + kWave);  // row q: words 2q, 2q + 1
        static_assert(kWave * (sizeof(float4) + 2 * sizeof(uint32_t)) <= sizeof(sh.b.sb.key),
                      "query coefficients and rows fit the stash");
        qcoef[lane] = make_float4(c_r, c_g, c_b, 0.0f);
        __syncthreads();
        float* __restrict__ dflux = D.dl_dflux;
        const int64_t n = D.n;
        const uint32_t half = lane >= 32 ? 1u : 0u;
        auto sweep = [&](auto shared_tag) {
          constexpr bool kSh = decltype(shared_tag)::value;
          constexpr int kSt = kSh ? kStageEff : kStageFull;
          for_candidates_st<kSt>(D, sh, lane, nr, total, sn, [&](uint32_t cnt, uint64_t amb) {
            uint64_t m = consume_st<kSt>(sh, cnt, [&](const Cand& c) {
              const float d2 = d2_ref(qx, qy, qz, c.p.x, c.p.y, c.p.z);
              const uint32_t db = __float_as_uint(d2);
              const bool nb = db < t_hi || (db == t_hi && __float_as_uint(c.p.w) <= t_lo);
              bool on;
              if constexpr (kSh) {
                // c.w: the flux when the shared sign lets it through (w set
                // to one), else zeros
                on = contributes && nb && c.w.w != 0.0f;
                const float sel = (contributes && nb) ? 1.0f : 0.0f;
                s_r = fmaf(sel, c.w.x, s_r);
                s_g = fmaf(sel, c.w.y, s_g);
                s_b = fmaf(sel, c.w.z, s_b);
              } else {
                const float cos_i = dot_ref(c.w.x, c.w.y, c.w.z, nx, ny, nz);
                on = contributes && nb && !(cos_i < 0.0f);
                const float sel = on ? 1.0f : 0.0f;
                s_r = fmaf(sel, c.w.w, s_r);
                s_g = fmaf(sel, c.f.x, s_g);
                s_b = fmaf(sel, c.f.y, s_b);
              }
              return on;
            });
            while (kSh && amb != 0ull) {
              const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(amb));
              amb &= amb - 1ull;
              const float4 p = sh.pos[j];
              const uint32_t db = __float_as_uint(d2_ref(qx, qy, qz, p.x, p.y, p.z));
              const bool nb = db < t_hi || (db == t_hi && __float_as_uint(p.w) <= t_lo);
              if (contributes && nb) {
                const uint32_t sl = sh.b.sb.chunk_slot[j];
                const float4 w = D.wif4[sl];
                const float2 f = D.flux2[sl];
                if (!(dot_ref(w.x, w.y, w.z, nx, ny, nz) < 0.0f)) {
                  s_r = s_r + w.w;
                  s_g = s_g + f.x;
                  s_b = s_b + f.y;
                  m |= 1ull << j;
                }
              }
            }
            rows32[2 * lane] = static_cast<uint32_t>(m);
            rows32[2 * lane + 1] = static_cast<uint32_t>(m >> 32);
            __syncthreads();
            // queries whose set holds some photon of the chunk (only
            // contributing lanes flag any)
            const uint64_t held = __ballot(m != 0ull);
            if (held != 0ull) {
              float a_r = 0.f, a_g = 0.f, a_b = 0.f;
              for (uint64_t hq = held; hq != 0ull; hq &= hq - 1ull) {
                const uint32_t q = static_cast<uint32_t>(__builtin_ctzll(hq));
                // the half of row q that holds this lane's bit
                const uint32_t word = rows32[2 * q + half];
                const float4 C = qcoef[q];
                const float sel = static_cast<float>((word >> (lane & 31)) & 1u);
                a_r = fmaf(sel, C.x, a_r);
                a_g = fmaf(sel, C.y, a_g);
                a_b = fmaf(sel, C.z, a_b);
              }
              if (static_cast<uint32_t>(lane) < cnt && (a_r != 0.0f || a_g != 0.0f || a_b != 0.0f)) {
                const uint32_t slot = sh.b.sb.chunk_slot[lane];
                atomicAdd(&dflux[slot], a_r);
                atomicAdd(&dflux[n + slot], a_g);
                atomicAdd(&dflux[2 * n + slot], a_b);
              }
            }
          });
        };
        if (shared) sweep(std::true_type{});
        else sweep(std::false_type{});
        float dr = 0.f, dg = 0.f, db = 0.f, fr = 0.f, fgv = 0.f, fb = 0.f;
        if (contributes) {
          const float inv = 1.0f / (kPiF * denom);
          dr = ge_r * s_r * inv; dg = ge_g * s_g * inv; db = ge_b * s_b * inv;
          if (fgm >= 0) {
            // forward value of the estimate L = (rho/pi) * S / denom
            const float lr = ((rho_r / kPiF) * s_r) / denom;
            const float lgv = ((rho_g / kPiF) * s_g) / denom;
            const float lb = ((rho_b / kPiF) * s_b) / denom;
            fr = g_r * lr; fgv = g_g * lgv; fb = g_b * lb;
          }
        }
        add_albedo_grad(drho, D.dl_dalbedo, contributes ? mat : -1, dr, dg, db, lane);
        add_albedo_grad(drho, D.dl_dalbedo, contributes ? fgm : -1, fr, fgv, fb, lane);
        __syncthreads();
      }
      pt.mark(6, lane);
      c0 = c1;
    }
  }

  if (MODE == kModeAdjoint || !in_range) return;
  if (!active) {
    // a query outside the domain: no neighbour set, r^2 and radiance NaN
    const float nan = __uint_as_float(0x7fc00000u);
    if (D.out_r2) D.out_r2[qi] = nan;
    if (D.out_key) D.out_key[qi] = ~0ull;
    if (kRows) {
      for (int32_t j = 0; j < D.k; ++j) row[j] = -1;
    } else if (kAccum) {
      D.out_rgb[3 * qi + 0] = nan;
      D.out_rgb[3 * qi + 1] = nan;
      D.out_rgb[3 * qi + 2] = nan;
    }
    return;
  }
  const float r2 = (k_eff > 0) ? __uint_as_float(static_cast<uint32_t>(tau >> 32)) : 0.0f;
  if (D.out_r2) D.out_r2[qi] = r2;
  if (D.out_key) D.out_key[qi] = (k_eff > 0) ? tau : 0ull;
  if (kRows) {
    for (int32_t j = static_cast<int32_t>(n_written); j < D.k; ++j) row[j] = -1;
    return;
  }
  if (kAccum) {
    float lr = 0.f, lgv = 0.f, lb = 0.f;
    if (lambert && k_eff > 0) {
      const float denom = (D.n_emit * kPiF) * r2;
      lr = ((rho_r / kPiF) * acc_r) / denom;
      lgv = ((rho_g / kPiF) * acc_g) / denom;
      lb = ((rho_b / kPiF) * acc_b) / denom;
    }
    D.out_rgb[3 * qi + 0] = lr;
    D.out_rgb[3 * qi + 1] = lgv;
    D.out_rgb[3 * qi + 2] = lb;
  }
}

__device__ __forceinline__ int find_instance(const PM_CONST int64_t* prefix, int n_inst,
                                             int64_t g) {
  int lo = 0, hi = n_inst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return uni(lo);
}

// Resident waves per SIMD the register allocation of the selection kernels
// must allow (a tuning macro): four, as many as their LDS allows.  The
// adjoint kernel's larger LDS holds it at three, so it is left free.
#ifndef PM_GATHER_MIN_WAVES
#define PM_GATHER_MIN_WAVES 4
#endif

template <int MODE>
__global__ __launch_bounds__(kWave, MODE == kModeAdjoint ? 1 : PM_GATHER_MIN_WAVES) void
gather_kernel(GatherLaunch L) {
  __shared__ std::conditional_t<MODE == kModeAdjoint, AdjointShared, GatherShared> sh;
  const int lane = threadIdx.x;
  float* drho = nullptr;
  if constexpr (MODE == kModeAdjoint) {
    drho = sh.drho;
    for (int i = lane; i < 3 * kDrhoSlots; i += kWave) drho[i] = 0.0f;
    __syncthreads();
  }
  const int nchunks = L.xcd_aware ? kXcds : 1;
  const int home = L.xcd_aware ? static_cast<int>(blockIdx.x % kXcds) : 0;
  const PM_CONST int64_t* gprefix = cptr(L.group_prefix);
  const PM_CONST GatherDesc* descs = cptr(L.descs);
  PhaseTimer pt;
  pt.acc = L.count_stats ? reinterpret_cast<uint64_t*>(L.counters + kPhaseBase) : nullptr;
  pt.t = clock64();
  for (int c = 0; c < nchunks; ++c) {
    const int chunk = (home + c) % nchunks;
    const int64_t cbeg = L.total_groups * chunk / nchunks;
    const int64_t cend = L.total_groups * (chunk + 1) / nchunks;
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(&L.counters[chunk], 1u);
      t = static_cast<uint32_t>(uni(static_cast<int>(t)));
      const int64_t g = cbeg + static_cast<int64_t>(t);
      if (g >= cend) break;
      const int inst = (L.n_inst > 1) ? find_instance(gprefix, L.n_inst, g) : 0;
      DescRef D = descs[inst];
      process_group<MODE>(D, g - gprefix[inst], L.perm, sh, drho, lane, L.counters,
                          L.count_stats, L.run_split, pt);
    }
  }
  if (MODE == kModeAdjoint) {
    __syncthreads();
    DescRef D = descs[0];
    for (int i = lane; i < 3 * min(D.n_mats, kDrhoSlots); i += kWave) {
      const float v = drho[i];
      if (v != 0.0f) atomicAdd(&D.dl_dalbedo[i], v);
    }
  }
}

// Radius bound r_a of every anchor for k_eff: the k_eff nearest photons of
// the anchor's centroid lie within r_a.  One anchor per wave at a time.
//   * Any list of photons of which at least k_eff have FP32 d^2(c, p) below
//     some t gives the bound t: missing photons could only be nearer.  A
//     wave-parallel histogram of d^2 over a candidate list, 256 bins over
//     [0, rb), finds the smallest such bin end below rb when there is one.
//   * rb starts from the local photon density: the spread s of the anchor's
//     own photons, scaled to k_eff photons on a surface (s^2 k_eff / stride,
//     times 2.25).  Consecutive slots can straddle a jump of the curve (it
//     leaves a surface and re-enters it elsewhere), which inflates s, so the
//     search keeps a bracket: a sphere whose leaf list overflows (a coarse
//     list, slow to sweep) shrinks it, one holding fewer than k_eff photons
//     grows it, and no coarse list is ever swept: in a dark region next to
//     a bright one a single wave would otherwise sweep tens of thousands of
//     photons.  rb never exceeds the distance to the farthest of the k_eff
//     consecutive slots centred on the anchor (they are k_eff photons), and
//     that distance is the bound when no sphere succeeds within 7 tries.
//   * A second histogram over [0, first bound) narrows the bound to one of
//     its bins (about 0.2% of the radius).
// A bound is all the gather needs from an anchor (it is used through the
// triangle inequality), so no exact selection is run.
__global__ __launch_bounds__(kWave) void anchor_bound_kernel(GatherLaunch L) {
  __shared__ GatherShared sh;
  const int lane = threadIdx.x;
  DescRef D = cptr(L.descs)[0];
  const uint32_t k_eff = static_cast<uint32_t>(D.k_eff);
  const int64_t n = D.n;
  const int64_t stride = D.anchor_stride;
  const float4* __restrict__ P = D.pos4;
  for (int64_t a = blockIdx.x; a < D.m; a += gridDim.x) {
    const float4 an = D.anchors[a];
    const float cx = an.x, cy = an.y, cz = an.z;
    // the guaranteed bound: the farthest of k_eff consecutive slots
    const int64_t mid = a * stride + stride / 2;
    int64_t s0 = mid - static_cast<int64_t>(k_eff / 2);
    s0 = max(min(s0, n - static_cast<int64_t>(k_eff)), static_cast<int64_t>(0));
    float far2 = 0.0f;
    for (int64_t sl = s0 + lane; sl < s0 + static_cast<int64_t>(k_eff); sl += kWave) {
      const float4 p = P[sl];
      far2 = fmaxf(far2, d2_ref(cx, cy, cz, p.x, p.y, p.z));
    }
    const float r0sq = wave_max(far2) * 1.0001f + 1e-30f;
    // the density guess
    const int64_t a0 = a * stride;
    const int64_t na = min(a0 + stride, n) - a0;
    float sp2 = 0.0f;
    for (int64_t sl = a0 + lane; sl < a0 + na; sl += kWave) {
      const float4 p = P[sl];
      sp2 = fmaxf(sp2, d2_ref(cx, cy, cz, p.x, p.y, p.z));
    }
    sp2 = wave_max(sp2);
    float rb = fminf(sp2 * (static_cast<float>(k_eff) / static_cast<float>(na)) * 2.25f, r0sq);
    const float r = leaf_bracket(D, sh, lane, cx, cy, cz, rb, 0.0f, INFINITY, r0sq, k_eff);
    if (lane == 0) D.anchor_out[4 * a + 3] = r;
    __syncthreads();
  }
}

__global__ void zero_counters_kernel(uint32_t* c, int n) {
  const int i = threadIdx.x;
  if (i < n) c[i] = 0;
}

}  // namespace

int launch_gather(Context* ctx, const GatherLaunch& L, hipStream_t s) {
  if (L.total_groups <= 0) return PM_OK;
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, L.counters,
                     kCounterWords);
  int64_t blocks = static_cast<int64_t>(ctx->n_cu) * ctx->waves_per_cu;
  if (blocks > L.total_groups) blocks = L.total_groups;
  if (blocks < 1) blocks = 1;
  const dim3 grid(static_cast<unsigned>(blocks));
  switch (L.mode) {
    case kModeRadiance:
      hipLaunchKernelGGL(gather_kernel<kModeRadiance>, grid, dim3(kWave), 0, s, L);
      break;
    case kModeKnn:
      hipLaunchKernelGGL(gather_kernel<kModeKnn>, grid, dim3(kWave), 0, s, L);
      break;
    case kModeAnchor:  // one anchor per wave at a time (total_groups = anchors)
      hipLaunchKernelGGL(anchor_bound_kernel, grid, dim3(kWave), 0, s, L);
      break;
    case kModeAdjoint:
      hipLaunchKernelGGL(gather_kernel<kModeAdjoint>, grid, dim3(kWave), 0, s, L);
      break;
    default:
      return fail(PM_ERR_INVALID_ARG, "unknown gather mode");
  }
  PM_HIP(hipGetLastError());
  return PM_OK;
}

}  // namespace pm
