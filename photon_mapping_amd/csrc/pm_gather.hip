// Persistent kNN photon-gather kernel for gfx950.
//
// One wavefront owns a group of 64 Morton-adjacent queries (one per lane).
// The k nearest neighbours are never kept in a heap.  Each lane selects the
// exact k-th smallest key
//   key = (bits(d^2) << 32) | original photon index
// which orders photons exactly like the reference's
// std::priority_queue<std::pair<float,int>> (include/photon_map.h:107-126):
//   pass 1  histogram of d^2 over [0, R0^2] in 32 linear bins (LDS), where R0
//           is a guaranteed upper bound on the k-NN radius from precomputed
//           radius anchors; the bin holding the k-th key is found;
//   pass 2  every photon below that bin is a neighbour and is consumed on the
//           spot (radiance sum, index write, ...), photons inside the bin are
//           stashed in LDS (a handful), and a rank selection inside the stash
//           finishes the set.
// The neighbour set is {key <= tau}; tau's d^2 is r^2.  Bins that hold more
// than kStashCap keys are split further (rare) until they fit.
//
// Memory-level parallelism: the candidate photons of a pass are the subtrees
// of the photon tree that meet the box of the lanes' query balls.  They are
// found by a breadth-first walk with one node per lane (one vector load per
// tree level) and kept in LDS as slot ranges.  A group whose queries straddle
// a jump of the Morton curve is cut into runs of nearby queries first, each
// with its own box.  Candidates are staged into LDS 64 at a time by a
// coalesced vector load, the next chunk being fetched into registers while
// the current one is consumed through broadcast LDS reads.  The consume loops
// are branch-free: the rare per-lane events (stash, index write) are recorded
// in a 64-bit mask per chunk and drained once per chunk.
//
// d^2 and the Lambert sign tests use the reference's FP32 operation order
// (include/photon_map.h:33-41, include/core.h:206-209); the file is compiled
// with -ffp-contract=off and denormals preserved, so they are bit-exact.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

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
constexpr int kMaxRanges = 256;   // candidate subtrees per pass
constexpr int kFront = 256;       // breadth-first frontier capacity
constexpr int kDrhoSlots = 64;    // adjoint: materials accumulated in LDS

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

struct GatherShared {
  uint32_t hist[(kBins + 1) * kWave];
  uint64_t stash[kStashCap * kWave];
  // The chunk staging buffers are dead while a candidate list is built, so
  // the breadth-first frontier shares their storage.
  union {
    struct {
      float4 pos[kChunk];
      float4 wif[kChunk];
      float2 flux[kChunk];
    } c;
    uint32_t front[2][kFront];
  } u;
  uint32_t rnode[kMaxRanges];
  uint32_t rpre[kMaxRanges + 1];
  float drho[3 * kDrhoSlots];
};

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
// lies within sqrt(cull_bound(r2)) of the query on every axis, despite the
// rounding of d^2 and of the box arithmetic.
__device__ __forceinline__ float cull_bound(float r2) {
  return r2 * 1.00002f + 1e-30f;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx) {
  return (static_cast<uint64_t>(__float_as_uint(d2)) << 32) | idx;
}

// Histogram bin of d^2 for the range bound r0sq: 32 linear bins over
// [0, r0sq); bin kBins collects everything beyond.  Monotone in d^2.  A
// photon with d^2 <= r0sq / (1 + 2^-20) always lands below kBins; the radius
// bounds carry a larger upward margin than that.
__device__ __forceinline__ uint32_t bin_of(float d2, float inv_w) {
  const uint32_t b = static_cast<uint32_t>(d2 * inv_w);
  return b < static_cast<uint32_t>(kBins) ? b : static_cast<uint32_t>(kBins);
}

// Smallest non-negative float (as bits) whose bin is >= b.  bin_of is
// monotone in d^2, so this is a lower-bound search over the bit patterns
// [0, +inf]; bin_of(+inf) = kBins.
__device__ __forceinline__ uint32_t bin_start_bits(uint32_t b, float inv_w) {
  uint32_t lo = 0;
  uint32_t hi = 0x7f800000u;
  for (int it = 0; it < 32 && lo < hi; ++it) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (bin_of(__uint_as_float(mid), inv_w) >= b) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return unif(v);
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return unif(v);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask, int lane) {
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  return static_cast<uint32_t>(__popcll(mask & lt));
}

// 1.0f when c holds, 0.0f otherwise (a select; keeps the consume loops free
// of exec-mask branches).
__device__ __forceinline__ float one_if(bool c) { return c ? 1.0f : 0.0f; }

// Per-lane selection state.
struct SelState {
  uint64_t klo;   // keys < klo are neighbours
  uint64_t khi;   // keys in [klo, khi) are stashed
  uint32_t need;  // 1-based rank of the k-th key inside [klo, khi)
  uint32_t cnt;   // number of candidate keys inside [klo, khi)
};

// Box containing every point within sqrt(r2) of any lane's query
// (wave-uniform).  Lanes that take no part pass r2 < 0.
struct GroupBox {
  float x0, y0, z0, x1, y1, z1;
};

__device__ __forceinline__ GroupBox group_box(float qx, float qy, float qz, float r2) {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (r2 >= 0.0f) {
    const float R = sqrtf(cull_bound(r2)) * 1.000004f;
    const float q[3] = {qx, qy, qz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float a = q[c] - R;
      float b = q[c] + R;
      lo[c] = a - (fabsf(a) * 4e-7f + 1e-30f);
      hi[c] = b + (fabsf(b) * 4e-7f + 1e-30f);
    }
  }
  GroupBox g;
  g.x0 = wave_min(lo[0]); g.y0 = wave_min(lo[1]); g.z0 = wave_min(lo[2]);
  g.x1 = wave_max(hi[0]); g.y1 = wave_max(hi[1]); g.z1 = wave_max(hi[2]);
  return g;
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

// Breadth-first walk of the implicit heap-ordered tree (root 1, children 2f
// and 2f+1, leaves from first_leaf), one frontier node per lane.  A node that
// misses the box is dropped; a leaf, or a node lying entirely inside the box,
// is emitted as a candidate subtree.  Emitted subtrees are disjoint and
// together hold every photon inside the box.  Writes the subtree roots to
// sh.rnode and the exclusive prefix of their slot counts to sh.rpre; returns
// the number of subtrees, the total slot count in *total.  When the frontier
// or the list would overflow, subtrees are taken whole (a superset; counted
// in *stat when given).
__device__ uint32_t build_candidates(DescRef D, const GroupBox& g, GatherShared& sh,
                                     int lane, uint32_t* total, uint32_t* stat) {
  const uint32_t first_leaf = static_cast<uint32_t>(D.first_leaf);
  if (lane == 0) sh.u.front[0][0] = 1u;
  __syncthreads();
  uint32_t ncur = 1, nr = 0;
  int cur = 0;
  bool overflow = false;
  while (ncur > 0 && !overflow) {
    uint32_t nnext = 0;
    for (uint32_t base = 0; base < ncur; base += kWave) {
      const uint32_t i = base + lane;
      const bool valid = i < ncur;
      const uint32_t f = valid ? sh.u.front[cur][i] : 1u;
      bool emit = false, push = false;
      if (valid) {
        const float4 mn = D.boxes[2 * f];
        const float4 mx = D.boxes[2 * f + 1];
        const bool hit = !(mx.x < g.x0 || mn.x > g.x1 || mx.y < g.y0 || mn.y > g.y1 ||
                           mx.z < g.z0 || mn.z > g.z1);
        const bool inside = mn.x >= g.x0 && mx.x <= g.x1 && mn.y >= g.y0 && mx.y <= g.y1 &&
                            mn.z >= g.z0 && mx.z <= g.z1;
        if (hit) {
          if (f >= first_leaf || inside) emit = true;
          else push = true;
        }
      }
      uint64_t em = __ballot(emit);
      uint64_t pmk = __ballot(push);
      if (nnext + 2u * static_cast<uint32_t>(__popcll(pmk)) > static_cast<uint32_t>(kFront)) {
        emit = emit || push;
        push = false;
        em = __ballot(emit);
        pmk = 0ull;
        if (stat && lane == 0) atomicAdd(stat, 1u);
      }
      const uint32_t ne = static_cast<uint32_t>(__popcll(em));
      if (nr + ne > static_cast<uint32_t>(kMaxRanges)) {
        overflow = true;
        if (stat && lane == 0) atomicAdd(stat, 1000u);
        break;
      }
      if (emit) sh.rnode[nr + lanes_below(em, lane)] = f;
      if (push) {
        const uint32_t at = nnext + 2u * lanes_below(pmk, lane);
        sh.u.front[cur ^ 1][at] = 2u * f;
        sh.u.front[cur ^ 1][at + 1] = 2u * f + 1u;
      }
      nr += ne;
      nnext += 2u * static_cast<uint32_t>(__popcll(pmk));
    }
    __syncthreads();
    cur ^= 1;
    ncur = nnext;
  }
  if (overflow) {
    nr = 1;
    if (lane == 0) sh.rnode[0] = 1u;
    __syncthreads();
  }
  // exclusive prefix of slot counts
  uint32_t run = 0;
  for (uint32_t base = 0; base < nr; base += kWave) {
    const uint32_t i = base + lane;
    uint32_t cnt = 0;
    if (i < nr) {
      int32_t s0, s1;
      node_slots(D, sh.rnode[i], s0, s1);
      cnt = static_cast<uint32_t>(s1 - s0);
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    if (i < nr) sh.rpre[i] = run + incl - cnt;
    run += static_cast<uint32_t>(__shfl(incl, kWave - 1, kWave));
  }
  if (lane == 0) sh.rpre[nr] = run;
  __syncthreads();
  *total = static_cast<uint32_t>(uni(static_cast<int>(run)));
  return static_cast<uint32_t>(uni(static_cast<int>(nr)));
}

// Iterates the candidate photons of the last build_candidates: body(cnt) is
// called once per staged chunk, after sh.u.c.pos (and with WF also
// sh.u.c.wif/flux) hold cnt candidates.  The next chunk's loads are in flight
// while body runs.
template <bool WF, typename Body>
__device__ __forceinline__ void for_candidates(DescRef D, GatherShared& sh, int lane,
                                               uint32_t nr, uint32_t total, Body&& body) {
  if (total == 0) return;
  const float4* __restrict__ P = D.pos4;
  const float4* __restrict__ W = D.wif4;
  const float2* __restrict__ F = D.flux2;
  // lane-private cursor: flat index t and its subtree r
  uint32_t t = lane;
  uint32_t r = 0;
  auto locate = [&]() -> int32_t {
    while (r + 1 < nr && sh.rpre[r + 1] <= t) ++r;
    int32_t s0, s1;
    node_slots(D, sh.rnode[r], s0, s1);
    return s0 + static_cast<int32_t>(t - sh.rpre[r]);
  };
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f), w = p;
  float2 f = make_float2(0.f, 0.f);
  if (t < total) {
    const int32_t slot = locate();
    p = P[slot];
    if (WF) {
      w = W[slot];
      f = F[slot];
    }
  }
  for (uint32_t o = 0; o < total; o += kChunk) {
    sh.u.c.pos[lane] = p;
    if (WF) {
      sh.u.c.wif[lane] = w;
      sh.u.c.flux[lane] = f;
    }
    __syncthreads();
    t += kChunk;
    if (t < total) {
      const int32_t slot = locate();
      p = P[slot];
      if (WF) {
        w = W[slot];
        f = F[slot];
      }
    }
    const uint32_t cnt = total - o < static_cast<uint32_t>(kChunk) ? total - o
                                                                    : static_cast<uint32_t>(kChunk);
    body(cnt);
    __syncthreads();
  }
}

// A staged candidate as the consume loops see it.
struct Cand {
  float4 p;   // position, original index bits
  float4 w;   // wi, flux.r   (staged with WF only)
  float2 f;   // flux.g, flux.b
};

template <bool WF>
__device__ __forceinline__ Cand load_cand(const GatherShared& sh, uint32_t j) {
  Cand c;
  c.p = sh.u.c.pos[j];
  if (WF) {
    c.w = sh.u.c.wif[j];
    c.f = sh.u.c.flux[j];
  } else {
    c.w = make_float4(0.f, 0.f, 0.f, 0.f);
    c.f = make_float2(0.f, 0.f);
  }
  return c;
}

// Calls use(cand) for the cnt staged candidates, reading them in blocks of 8
// into registers first (the LDS reads of a block are in flight together).
// use returns a per-lane flag; the flags form a 64-bit mask over the chunk,
// which is returned for draining.
template <bool WF, typename Use>
__device__ __forceinline__ uint64_t consume(const GatherShared& sh, uint32_t cnt, Use&& use) {
  uint32_t mlo = 0, mhi = 0;
  uint32_t j = 0;
  for (; j + 8 <= cnt; j += 8) {
    Cand c[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) c[u] = load_cand<WF>(sh, j + u);
    uint32_t bits = 0;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) bits |= use(c[u]) ? (1u << u) : 0u;
    if (j < 32) mlo |= bits << j;
    else mhi |= bits << (j - 32);
  }
  for (; j < cnt; ++j) {
    const uint32_t bit = use(load_cand<WF>(sh, j)) ? 1u : 0u;
    if (j < 32) mlo |= bit << j;
    else mhi |= bit << (j - 32);
  }
  return (static_cast<uint64_t>(mhi) << 32) | mlo;
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
  for (int i = lane; i < (kBins + 1) * kWave / 4; i += kWave) h4[i] = make_uint4(0, 0, 0, 0);
}

__device__ __forceinline__ void hist_add(GatherShared& sh, uint32_t b, int lane) {
  __hip_atomic_fetch_add(&sh.hist[b * kWave + lane], 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave-cycle accounting of the phases of a group, for statistics runs only
// (acc == nullptr otherwise).  Phases: 0 query and radius bound, 1 pass-1
// candidate lists, 2 pass-1 histogram, 3 thresholds and refinement, 4 pass-2
// candidate lists, 5 pass-2 selection, 6 stash and outputs, 7 work queue.
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

// Runs one pass over the candidates of every lane, where a lane's candidates
// are the photons within sqrt(r2) of its query (lanes with r2 < 0 take no
// part).  The Morton-sorted lanes are cut into runs wherever a query lies
// more than twice the largest radius away from its predecessor, so that a
// group straddling a jump of the Morton curve does not scan the space between
// its pieces.  For each run the candidate list of its lanes is built, then
// pass(part, nr, total) is called with part = "this lane is in the run".
template <typename Pass>
__device__ __forceinline__ void clustered(DescRef D, float qx, float qy, float qz, float r2,
                                          GatherShared& sh, int lane, uint32_t* counters,
                                          int count_stats, int cand_stat, PhaseTimer& pt,
                                          int list_phase, int pass_phase, Pass&& pass) {
  const bool in = r2 >= 0.0f;
  const float R = in ? sqrtf(cull_bound(r2)) : 0.0f;
  const float Rm = wave_max(R);
  const float px = __shfl_up(qx, 1, kWave);
  const float py = __shfl_up(qy, 1, kWave);
  const float pz = __shfl_up(qz, 1, kWave);
  const int prev_in = __shfl_up(in ? 1 : 0, 1, kWave);
  const float gap = fmaxf(fabsf(qx - px), fmaxf(fabsf(qy - py), fabsf(qz - pz)));
  const bool start = in && (lane == 0 || !prev_in || gap > 2.0f * Rm);
  uint64_t starts = __ballot(start);
  while (starts != 0ull) {
    const int c0 = __builtin_ctzll(starts);
    starts &= starts - 1ull;
    const int c1 = starts != 0ull ? __builtin_ctzll(starts) : kWave;
    const bool part = in && lane >= c0 && lane < c1;
    const GroupBox g = group_box(qx, qy, qz, part ? r2 : -1.0f);
    uint32_t total = 0;
    const uint32_t nr = build_candidates(D, g, sh, lane, &total,
                                         count_stats ? &counters[kStatOverflow] : nullptr);
    if (count_stats && lane == 0) atomicAdd(&counters[cand_stat], total);
    pt.mark(list_phase, lane);
    pass(part, nr, total);
    pt.mark(pass_phase, lane);
  }
}

// Determines the selection state of every lane: afterwards the neighbour set
// of a lane is {key < klo} plus the `need` smallest keys of [klo, khi), and
// [klo, khi) holds at most kStashCap keys.
__device__ void select_range(DescRef D, float qx, float qy, float qz,
                             bool active, float r0sq, uint32_t k_eff,
                             GatherShared& sh, int lane, uint32_t* counters,
                             int count_stats, SelState& st, PhaseTimer& pt) {
  st.klo = 0; st.khi = 0; st.need = 0; st.cnt = 0;
  // ---- pass 1: histogram of d^2 over [0, R0^2] -------------------------------
  for (int attempt = 0; attempt < 2; ++attempt) {
    // (the bins of inactive lanes are never read)
    const float inv_w = active ? static_cast<float>(kBins) / r0sq : 0.0f;
    clear_hist(sh, lane);
    __syncthreads();
    clustered(D, qx, qy, qz, active ? r0sq : -1.0f, sh, lane, counters, count_stats,
              kStatCand1, pt, 1, 2, [&](bool part, uint32_t nr, uint32_t total) {
      for_candidates<false>(D, sh, lane, nr, total, [&](uint32_t cnt) {
        consume<false>(sh, cnt, [&](const Cand& c) {
          const float d2 = d2_ref(qx, qy, qz, c.p.x, c.p.y, c.p.z);
          hist_add(sh, part ? bin_of(d2, inv_w) : static_cast<uint32_t>(kBins), lane);
          return false;
        });
      });
    });
    __syncthreads();
    uint32_t cum = 0, below = 0, cb = 0, bsel = kBins;
#pragma unroll
    for (int j = 0; j < kBins; ++j) {
      const uint32_t h = sh.hist[j * kWave + lane];
      if (bsel == kBins && cum + h >= k_eff) {
        bsel = j;
        below = cum;
        cb = h;
      }
      cum += h;
    }
    const bool lane_ok = !active || bsel < kBins;
    if (!wave_any(!lane_ok)) {
      if (active) {
        const uint32_t lo_bits = bin_start_bits(bsel, inv_w);
        const uint32_t hi_bits = bin_start_bits(bsel + 1, inv_w);
        st.klo = static_cast<uint64_t>(lo_bits) << 32;
        st.khi = static_cast<uint64_t>(hi_bits) << 32;
        st.need = k_eff - below;
        st.cnt = cb;
      }
      break;
    }
    // A lane found fewer than k photons inside its bound.  The bounds are
    // guaranteed, so this is a guard only: retry that lane unbounded.
    if (count_stats && lane == 0) atomicAdd(&counters[kStatFallback], 1u);
    if (active && !lane_ok) r0sq = INFINITY;
  }
  pt.mark(3, lane);

  // ---- refinement: split [klo, khi) until it holds <= kStashCap keys --------
  // Each round shrinks the key span of a refining lane 32-fold, so 64-bit
  // spans end within 13 rounds; the cap only bounds the loop if the counts
  // were ever inconsistent.
  for (int round = 0;
       round < 16 && wave_any(active && st.cnt > static_cast<uint32_t>(kStashCap));
       ++round) {
    if (count_stats && lane == 0) atomicAdd(&counters[kStatRefine], 1u);
    const bool mine = active && st.cnt > static_cast<uint32_t>(kStashCap);
    const uint64_t span = st.khi - st.klo;
    const uint64_t w = mine ? (span + kBins - 1) / kBins : 1ull;
    clear_hist(sh, lane);
    __syncthreads();
    const float r2hi = __uint_as_float(static_cast<uint32_t>(st.khi >> 32));
    clustered(D, qx, qy, qz, mine ? r2hi : -1.0f, sh, lane, counters, 0, kStatCand2, pt, 3,
              3, [&](bool part, uint32_t nr, uint32_t total) {
      for_candidates<false>(D, sh, lane, nr, total, [&](uint32_t cnt) {
        for (uint32_t j = 0; j < cnt; ++j) {
          const float4 p = sh.u.c.pos[j];
          const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
          const uint64_t key = make_key(d2, __float_as_uint(p.w));
          if (part && key >= st.klo && key < st.khi) {
            hist_add(sh, static_cast<uint32_t>((key - st.klo) / w), lane);
          }
        }
      });
    });
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
}

template <int MODE>
__device__ void process_group(DescRef D, int64_t lg, const uint32_t* __restrict__ perm,
                              GatherShared& sh, int lane, uint32_t* counters,
                              int count_stats, PhaseTimer& pt) {
  const int64_t qloc = lg * kWave + lane;
  const bool active = qloc < D.m;
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
  if (MODE == kModeAnchor) {
    qi = active ? qloc : lg * kWave;
    const float4 a = D.anchors[qi];
    qx = a.x; qy = a.y; qz = a.z;
  } else {
    const int64_t sp = D.q_base + (active ? qloc : lg * kWave);
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

  // ---- radius bound R0 ---------------------------------------------------------
  float r0sq = -1.0f;  // inactive lanes never match anything
  const bool reuse = (MODE == kModeAdjoint) && D.key_in != nullptr;
  if (D.all_members || reuse) {
    r0sq = INFINITY;
  } else if (MODE == kModeAnchor) {
    // Union box of the leaves around this anchor that hold >= k photons;
    // every one of those photons is inside it, so the farthest box corner
    // bounds the k-NN distance of the anchor centroid.
    const int32_t leaf_size = D.leaf_size;
    const int32_t leaf = static_cast<int32_t>((qi * D.anchor_stride) / leaf_size);
    int32_t j0 = leaf, j1 = leaf;
    auto count_of = [&](int32_t a, int32_t b) -> int64_t {
      const int64_t hi = min(static_cast<int64_t>(b + 1) * leaf_size, static_cast<int64_t>(D.n));
      return hi - static_cast<int64_t>(a) * leaf_size;
    };
    while (count_of(j0, j1) < static_cast<int64_t>(k_eff)) {
      if (j0 > 0) --j0;
      if (j1 < D.n_leaves - 1) ++j1;
    }
    float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
    float4 mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    for (int32_t j = j0; j <= j1; ++j) {
      const float4 a = D.boxes[2 * (D.first_leaf + j)];
      const float4 b = D.boxes[2 * (D.first_leaf + j) + 1];
      mn.x = fminf(mn.x, a.x); mn.y = fminf(mn.y, a.y); mn.z = fminf(mn.z, a.z);
      mx.x = fmaxf(mx.x, b.x); mx.y = fmaxf(mx.y, b.y); mx.z = fmaxf(mx.z, b.z);
    }
    const float ex = fmaxf(fabsf(qx - mn.x), fabsf(mx.x - qx));
    const float ey = fmaxf(fabsf(qy - mn.y), fabsf(mx.y - qy));
    const float ez = fmaxf(fabsf(qz - mn.z), fabsf(mx.z - qz));
    if (active) r0sq = (ex * ex + ey * ey + ez * ez) * 1.0001f + 1e-30f;
  } else {
    // Triangle inequality on precomputed anchors: the k nearest photons of
    // anchor a lie within r_a of its centroid c_a, hence within
    // r_a + |q - c_a| of q.  The anchor radii are already rounded upward.
    const PM_CONST float4* A = cptr(D.anchors);
    const int32_t center = uni(static_cast<int32_t>(D.group_anchor[lg]));
    const int32_t a0 = max(center - kAnchorWindow, 0);
    const int32_t a1 = min(center + kAnchorWindow, D.n_anchors - 1);
    float best = INFINITY;
    for (int32_t a = a0; a <= a1; ++a) {
      const float4 an = A[a];
      const float dx = qx - an.x, dy = qy - an.y, dz = qz - an.z;
      best = fminf(best, an.w + sqrtf(dx * dx + dy * dy + dz * dz));
    }
    best *= 1.00002f;
    if (active) r0sq = best * best * 1.00002f + 1e-30f;
  }
  if (!active) r0sq = -1.0f;
  pt.mark(0, lane);

  SelState st;
  if (reuse) {
    // The forward pass's key is the largest key of the neighbour set.
    const uint64_t tau = active ? D.key_in[qi] : 0ull;
    st.klo = tau + 1; st.khi = tau + 1; st.need = 0; st.cnt = 0;
  } else if (D.all_members) {
    // Every photon is a neighbour; pass 2 tracks the largest key.
    st.klo = kKeyMax; st.khi = kKeyMax; st.need = 0; st.cnt = 0;
  } else {
    select_range(D, qx, qy, qz, active, r0sq, k_eff, sh, lane, counters, count_stats, st,
                 pt);
  }

  // ---- pass 2: consume neighbours below klo, stash [klo, khi) -------------------
  // (the adjoint runs it only to finish the selection; its scatter pass
  // follows once tau is known)
  constexpr bool kAccum = MODE == kModeRadiance;
  constexpr bool kRows = MODE == kModeKnn;
  float acc_r = 0.f, acc_g = 0.f, acc_b = 0.f;
  uint32_t n_stash = 0;
  uint32_t n_written = 0;
  uint64_t key_max = 0;
  int32_t* row = nullptr;
  if (kRows) row = D.out_slots + qi * D.k;
  const float wt_o = one_if(lambert && !(cos_o < 0.0f));  // outgoing side passes
  if (!reuse) {
    const bool all = D.all_members != 0;
    const float r2hi = all ? INFINITY : __uint_as_float(static_cast<uint32_t>(st.khi >> 32));
    const uint64_t span = st.khi - st.klo;
    clustered(D, qx, qy, qz, active ? r2hi : -1.0f, sh, lane, counters, count_stats,
              kStatCand2, pt, 4, 5, [&](bool part, uint32_t nr, uint32_t total) {
      const float wt_part = part ? 1.0f : 0.0f;
      for_candidates<kAccum>(D, sh, lane, nr, total, [&](uint32_t cnt) {
        const uint64_t m = consume<kAccum>(sh, cnt, [&](const Cand& c) {
          const float d2 = d2_ref(qx, qy, qz, c.p.x, c.p.y, c.p.z);
          const uint64_t key = make_key(d2, __float_as_uint(c.p.w));
          const float wm = one_if(key < st.klo) * wt_part;  // neighbour of this lane
          if (kAccum) {
            const float cos_i = dot_ref(c.w.x, c.w.y, c.w.z, nx, ny, nz);
            const float sel = wm * wt_o * one_if(!(cos_i < 0.0f));
            acc_r = fmaf(sel, c.w.w, acc_r);
            acc_g = fmaf(sel, c.f.x, acc_g);
            acc_b = fmaf(sel, c.f.y, acc_b);
          }
          if (all) key_max = (part && key > key_max) ? key : key_max;
          // flagged: inside the stash range, or (kNN rows) a neighbour
          const bool in_stash = (key - st.klo) < span;
          return part && (in_stash || (kRows && key < st.klo));
        });
        drain(m, [&](uint32_t j) {
          const float4 p = sh.u.c.pos[j];
          const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
          const uint32_t idx = __float_as_uint(p.w);
          const uint64_t key = make_key(d2, idx);
          if (kRows && key < st.klo) {
            if (n_written < static_cast<uint32_t>(D.k)) row[n_written++] = static_cast<int32_t>(idx);
          } else if (n_stash < static_cast<uint32_t>(kStashCap)) {
            sh.stash[n_stash * kWave + lane] = key;
            ++n_stash;
          }
        });
      });
    });
  }

  // ---- finish the set inside the stash ---------------------------------------------
  uint64_t tau = 0;
  if (reuse) {
    tau = st.klo - 1;
  } else if (D.all_members) {
    tau = key_max;
  } else if (active) {
    for (uint32_t i = 0; i < n_stash; ++i) {
      const uint64_t ki = sh.stash[i * kWave + lane];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < n_stash; ++j) rank += sh.stash[j * kWave + lane] < ki ? 1u : 0u;
      if (rank + 1 == st.need) tau = ki;
      if (rank < st.need) {
        const uint32_t idx = static_cast<uint32_t>(ki & 0xffffffffu);
        if (kAccum) {
          const int32_t s = D.inv_perm[idx];
          const float4 w = D.wif4[s];
          const float2 f = D.flux2[s];
          const float cos_i = dot_ref(w.x, w.y, w.z, nx, ny, nz);
          const float sel = wt_o * one_if(!(cos_i < 0.0f));
          acc_r = fmaf(sel, w.w, acc_r);
          acc_g = fmaf(sel, f.x, acc_g);
          acc_b = fmaf(sel, f.y, acc_b);
        } else if (kRows) {
          if (n_written < static_cast<uint32_t>(D.k)) row[n_written++] = static_cast<int32_t>(idx);
        }
      }
    }
  }
  const float r2 = (k_eff > 0) ? __uint_as_float(static_cast<uint32_t>(tau >> 32)) : 0.0f;

  if (MODE == kModeAdjoint) {
    // ---- scatter pass: dL/dflux over the neighbour set {key <= tau} -------------
    const float denom = (D.n_emit * kPiF) * r2;
    float g_r = 0.f, g_g = 0.f, g_b = 0.f;
    float fg_r = 1.f, fg_g = 1.f, fg_b = 1.f;
    int32_t fgm = -1;
    if (active && lambert && k_eff > 0) {
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
    const bool contributes = active && lambert && k_eff > 0;
    const uint64_t klim = tau + 1;
    float s_r = 0.f, s_g = 0.f, s_b = 0.f;
    const int32_t n = D.n;
    float* __restrict__ dflux = D.dl_dflux;
    clustered(D, qx, qy, qz, contributes ? r2 : -1.0f, sh, lane, counters, 0, kStatCand2, pt,
              4, 5, [&](bool part, uint32_t nr, uint32_t total) {
      for_candidates<true>(D, sh, lane, nr, total, [&](uint32_t cnt) {
        for (uint32_t j = 0; j < cnt; ++j) {
          const float4 p = sh.u.c.pos[j];
          const float4 w = sh.u.c.wif[j];
          const float2 f = sh.u.c.flux[j];
          const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
          const uint64_t key = make_key(d2, __float_as_uint(p.w));
          const float cos_i = dot_ref(w.x, w.y, w.z, nx, ny, nz);
          const bool use = part && key < klim && !(cos_o < 0.0f || cos_i < 0.0f);
          if (wave_any(use)) {
            const float sel = use ? 1.0f : 0.0f;
            s_r = fmaf(sel, w.w, s_r);
            s_g = fmaf(sel, f.x, s_g);
            s_b = fmaf(sel, f.y, s_b);
            const float tr = wave_sum(use ? c_r : 0.0f);
            const float tg = wave_sum(use ? c_g : 0.0f);
            const float tb = wave_sum(use ? c_b : 0.0f);
            if (lane == 0) {
              const uint32_t idx = __float_as_uint(p.w);
              atomicAdd(&dflux[idx], tr);
              atomicAdd(&dflux[static_cast<int64_t>(n) + idx], tg);
              atomicAdd(&dflux[2 * static_cast<int64_t>(n) + idx], tb);
            }
          }
        }
      });
    });
    // dL/drho of the queried material, and of the final-gathering albedo
    if (contributes) {
      const float inv = 1.0f / (kPiF * denom);
      const float dr = ge_r * s_r * inv, dg = ge_g * s_g * inv, db = ge_b * s_b * inv;
      if (mat < kDrhoSlots) {
        atomicAdd(&sh.drho[3 * mat + 0], dr);
        atomicAdd(&sh.drho[3 * mat + 1], dg);
        atomicAdd(&sh.drho[3 * mat + 2], db);
      } else {
        atomicAdd(&D.dl_dalbedo[3 * mat + 0], dr);
        atomicAdd(&D.dl_dalbedo[3 * mat + 1], dg);
        atomicAdd(&D.dl_dalbedo[3 * mat + 2], db);
      }
      if (fgm >= 0) {
        // forward value of the estimate L = (rho/pi) * S / denom
        const float lr = ((rho_r / kPiF) * s_r) / denom;
        const float lgv = ((rho_g / kPiF) * s_g) / denom;
        const float lb = ((rho_b / kPiF) * s_b) / denom;
        if (fgm < kDrhoSlots) {
          atomicAdd(&sh.drho[3 * fgm + 0], g_r * lr);
          atomicAdd(&sh.drho[3 * fgm + 1], g_g * lgv);
          atomicAdd(&sh.drho[3 * fgm + 2], g_b * lb);
        } else {
          atomicAdd(&D.dl_dalbedo[3 * fgm + 0], g_r * lr);
          atomicAdd(&D.dl_dalbedo[3 * fgm + 1], g_g * lgv);
          atomicAdd(&D.dl_dalbedo[3 * fgm + 2], g_b * lb);
        }
      }
    }
    return;
  }

  if (!active) return;
  if (MODE == kModeAnchor) {
    D.anchor_out[4 * qi + 3] = sqrtf(r2) * 1.00002f + 1e-30f;
    return;
  }
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

template <int MODE>
__global__ __launch_bounds__(kWave) void gather_kernel(GatherLaunch L) {
  __shared__ GatherShared sh;
  const int lane = threadIdx.x;
  if (MODE == kModeAdjoint) {
    for (int i = lane; i < 3 * kDrhoSlots; i += kWave) sh.drho[i] = 0.0f;
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
      process_group<MODE>(D, g - gprefix[inst], L.perm, sh, lane, L.counters,
                          L.count_stats, pt);
      pt.mark(6, lane);
    }
  }
  if (MODE == kModeAdjoint) {
    __syncthreads();
    DescRef D = descs[0];
    for (int i = lane; i < 3 * min(D.n_mats, kDrhoSlots); i += kWave) {
      const float v = sh.drho[i];
      if (v != 0.0f) atomicAdd(&D.dl_dalbedo[i], v);
    }
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
    case kModeAnchor:
      hipLaunchKernelGGL(gather_kernel<kModeAnchor>, grid, dim3(kWave), 0, s, L);
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
