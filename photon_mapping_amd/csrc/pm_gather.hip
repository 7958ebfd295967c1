// Persistent kNN photon-gather kernel for gfx950.
//
// One wavefront owns a group of 64 Morton-adjacent queries (one per lane) and
// walks the photon tree once for the whole group: node boxes and photon
// records are read at wave-uniform addresses through the constant address
// space (scalar loads into SGPRs), and every lane tests them against its own
// query.  The k nearest neighbours are never kept in a heap.  Each lane
// selects the exact k-th smallest key
//   key = (bits(d^2) << 32) | original photon index
// which orders photons exactly like the reference's
// std::priority_queue<std::pair<float,int>> (include/photon_map.h:107-126):
//   pass 1  histogram of d^2 over [0, R0^2] in 32 linear bins (LDS), where R0
//           is a guaranteed upper bound on the k-NN radius from precomputed
//           radius anchors; the bin holding the k-th key is found;
//   pass 2  every photon below that bin is a neighbour and is consumed on the
//           spot (radiance sum, slot write, ...), photons inside the bin are
//           stashed in LDS (a handful), and a rank selection inside the stash
//           finishes the set.
// The neighbour set is {key <= tau}; tau's d^2 is r^2.  Bins that hold more
// than kStashCap keys are split further (rare) until they fit.
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
constexpr int kBatch = 8;         // photons fetched per scalar-load batch
constexpr int kDrhoSlots = 64;    // adjoint: materials accumulated in LDS

template <typename T>
__device__ __forceinline__ const PM_CONST T* cptr(const T* p) {
  return (const PM_CONST T*)(p);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

struct GatherShared {
  uint32_t hist[(kBins + 1) * kWave];
  uint64_t stash_key[kStashCap * kWave];
  uint32_t stash_slot[kStashCap * kWave];
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

// Squared distance from a point to a box (0 inside).  Empty boxes
// (min = +inf, max = -inf) give +inf.
__device__ __forceinline__ float box_d2(float qx, float qy, float qz, float4 mn,
                                        float4 mx) {
  const float ex = fmaxf(fmaxf(mn.x - qx, qx - mx.x), 0.0f);
  const float ey = fmaxf(fmaxf(mn.y - qy, qy - mx.y), 0.0f);
  const float ez = fmaxf(fmaxf(mn.z - qz, qz - mx.z), 0.0f);
  return ex * ex + ey * ey + ez * ez;
}

// Conservative culling radius: a node is skipped only when every photon in
// it has an FP32 d^2 strictly above r2 (the margin covers the rounding of the
// box distance and of d^2 itself).
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

// Per-lane selection state.
struct SelState {
  uint64_t klo;   // keys < klo are neighbours
  uint64_t khi;   // keys in [klo, khi) are stashed
  uint32_t need;  // 1-based rank of the k-th key inside [klo, khi)
  uint32_t cnt;   // number of candidate keys inside [klo, khi)
};

// Walks the implicit heap-ordered tree (root 1, children 2n and 2n+1, leaves
// from first_leaf) in a fixed left-to-right order and calls leaf_fn(s0, s1)
// for the slot range of every leaf whose box lies within sqrt(cull_r2) of at
// least one lane's query (lanes pass a negative cull_r2 to opt out).  The
// walk is stackless: after a leaf or a culled node it moves to the next node
// in depth-first order by stripping trailing one bits.
template <typename LeafFn>
__device__ __forceinline__ void walk_tree(const GatherDesc& D, float qx,
                                          float qy, float qz, float cull_r2,
                                          LeafFn&& leaf_fn) {
  const uint32_t first_leaf = static_cast<uint32_t>(D.first_leaf);
  const uint32_t n_leaves = static_cast<uint32_t>(D.n_leaves);
  const PM_CONST float4* boxes = cptr(D.boxes);
  const int32_t leaf_size = D.leaf_size;
  const int32_t n = D.n;
  uint32_t node = 1;
  for (;;) {
    bool descend = false;
    const float4 mn = boxes[2 * node];
    const float4 mx = boxes[2 * node + 1];
    const bool hit = wave_any(box_d2(qx, qy, qz, mn, mx) <= cull_r2);
    if (node >= first_leaf) {
      const uint32_t leaf = node - first_leaf;
      if (hit && leaf < n_leaves) {
        const int32_t s0 = static_cast<int32_t>(leaf) * leaf_size;
        const int32_t s1 = min(s0 + leaf_size, n);
        leaf_fn(s0, s1);
      }
    } else {
      descend = hit;
    }
    if (descend) {
      node = 2 * node;
    } else {
      node = (node >> __builtin_ctz(~node)) + 1u;
      if (node == 1u) break;
    }
  }
}

// Calls fn(pos4, slot) for every slot in [s0, s1), fetching kBatch records
// per batch with scalar loads so their latency overlaps.
template <typename Fn>
__device__ __forceinline__ void for_slots(const PM_CONST float4* P, int32_t s0, int32_t s1,
                                          Fn&& fn) {
  int32_t s = s0;
  for (; s + kBatch <= s1; s += kBatch) {
    float4 p[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) p[u] = P[s + u];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) fn(p[u], s + u);
  }
  for (; s < s1; ++s) fn(P[s], s);
}

// Same, also fetching the (wi, flux) records of each slot.
template <typename Fn>
__device__ __forceinline__ void for_slots_full(const PM_CONST float4* P,
                                               const PM_CONST float4* W,
                                               const PM_CONST float2* F, int32_t s0,
                                               int32_t s1, Fn&& fn) {
  constexpr int B = kBatch / 2;
  int32_t s = s0;
  for (; s + B <= s1; s += B) {
    float4 p[B], w[B];
    float2 f[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      p[u] = P[s + u];
      w[u] = W[s + u];
      f[u] = F[s + u];
    }
#pragma unroll
    for (int u = 0; u < B; ++u) fn(p[u], w[u], f[u], s + u);
  }
  for (; s < s1; ++s) fn(P[s], W[s], F[s], s);
}

__device__ __forceinline__ void clear_hist(GatherShared& sh, int lane) {
  uint4* h4 = reinterpret_cast<uint4*>(sh.hist);
  for (int i = lane; i < (kBins + 1) * kWave / 4; i += kWave) h4[i] = make_uint4(0, 0, 0, 0);
}

__device__ __forceinline__ void hist_add(GatherShared& sh, uint32_t b, int lane) {
  __hip_atomic_fetch_add(&sh.hist[b * kWave + lane], 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Determines the selection state of every lane: after this, the neighbour
// set of lane q is {key < klo} plus the need smallest keys of [klo, khi),
// and [klo, khi) holds at most kStashCap keys.
__device__ void select_range(const GatherDesc& D, float qx, float qy, float qz,
                             bool active, float r0sq, uint32_t k_eff,
                             GatherShared& sh, int lane, uint32_t* counters,
                             int count_stats, SelState& st) {
  const PM_CONST float4* P = cptr(D.pos4);
  st.klo = 0; st.khi = 0; st.need = 0; st.cnt = 0;
  // ---- pass 1: histogram of d^2 over [0, R0^2] -------------------------------
  for (int attempt = 0; attempt < 2; ++attempt) {
    // (the bins of inactive lanes are never read)
    const float inv_w = active ? static_cast<float>(kBins) / r0sq : 0.0f;
    clear_hist(sh, lane);
    __syncthreads();
    const float cull = active ? cull_bound(r0sq) : -1.0f;
    walk_tree(D, qx, qy, qz, cull, [&](int32_t s0, int32_t s1) {
      for_slots(P, s0, s1, [&](float4 p, int32_t) {
        const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
        hist_add(sh, bin_of(d2, inv_w), lane);
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
    if (count_stats && lane == 0) atomicAdd(&counters[2 * kXcds + 0], 1u);
    if (active && !lane_ok) r0sq = INFINITY;
  }

  // ---- refinement: split [klo, khi) until it holds <= kStashCap keys --------
  // Each round shrinks the key span of a refining lane 32-fold, so 64-bit
  // spans end within 13 rounds; the cap only bounds the loop if the counts
  // were ever inconsistent.
  for (int round = 0;
       round < 16 && wave_any(active && st.cnt > static_cast<uint32_t>(kStashCap));
       ++round) {
    if (count_stats && lane == 0) atomicAdd(&counters[2 * kXcds + 1], 1u);
    const bool mine = active && st.cnt > static_cast<uint32_t>(kStashCap);
    const uint64_t span = st.khi - st.klo;
    const uint64_t w = mine ? (span + kBins - 1) / kBins : 1ull;
    clear_hist(sh, lane);
    __syncthreads();
    const float cull =
        mine ? cull_bound(__uint_as_float(static_cast<uint32_t>(st.khi >> 32))) : -1.0f;
    walk_tree(D, qx, qy, qz, cull, [&](int32_t s0, int32_t s1) {
      for_slots(P, s0, s1, [&](float4 p, int32_t) {
        const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
        const uint64_t key = make_key(d2, __float_as_uint(p.w));
        if (mine && key >= st.klo && key < st.khi) {
          hist_add(sh, static_cast<uint32_t>((key - st.klo) / w), lane);
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
__device__ void process_group(const GatherDesc& D, int64_t lg,
                              const uint32_t* __restrict__ perm,
                              GatherShared& sh, int lane, uint32_t* counters,
                              int count_stats) {
  const int64_t qloc = lg * kWave + lane;
  const bool active = qloc < D.m;
  const uint32_t k_eff = static_cast<uint32_t>(D.k_eff);
  const PM_CONST float4* P = cptr(D.pos4);
  const PM_CONST float4* W = cptr(D.wif4);
  const PM_CONST float2* F = cptr(D.flux2);

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

  SelState st;
  if (reuse) {
    // The forward pass's key is the largest key of the neighbour set.
    const uint64_t tau = active ? D.key_in[qi] : 0ull;
    st.klo = tau + 1; st.khi = tau + 1; st.need = 0; st.cnt = 0;
  } else if (D.all_members) {
    // Every photon is a neighbour; pass 2 tracks the largest key.
    st.klo = kKeyMax; st.khi = kKeyMax; st.need = 0; st.cnt = 0;
  } else {
    select_range(D, qx, qy, qz, active, r0sq, k_eff, sh, lane, counters, count_stats, st);
  }

  // ---- pass 2: consume neighbours below klo, stash [klo, khi) -------------------
  // (the adjoint runs it only to finish the selection; its scatter pass
  // follows once tau is known)
  constexpr bool kAccum = MODE == kModeRadiance;
  float acc_r = 0.f, acc_g = 0.f, acc_b = 0.f;
  uint32_t n_stash = 0;
  uint32_t n_written = 0;
  uint64_t key_max = 0;
  int32_t* row = nullptr;
  if (MODE == kModeKnn) row = D.out_slots + qi * D.k;
  const bool need_pass2 = !reuse;
  if (need_pass2) {
    const float cull =
        !active ? -1.0f
                : (D.all_members ? INFINITY
                                 : cull_bound(__uint_as_float(static_cast<uint32_t>(st.khi >> 32))));
    auto visit = [&](float4 p, int32_t s, bool member, uint64_t key) {
      if (D.all_members) key_max = (active && key > key_max) ? key : key_max;
      if (MODE == kModeKnn) {
        if (member && n_written < static_cast<uint32_t>(D.k)) row[n_written++] = s;
      }
      if (active && key >= st.klo && key < st.khi && n_stash < static_cast<uint32_t>(kStashCap)) {
        sh.stash_key[n_stash * kWave + lane] = key;
        sh.stash_slot[n_stash * kWave + lane] = static_cast<uint32_t>(s);
        ++n_stash;
      }
    };
    if (kAccum) {
      walk_tree(D, qx, qy, qz, cull, [&](int32_t s0, int32_t s1) {
        for_slots_full(P, W, F, s0, s1, [&](float4 p, float4 w, float2 f, int32_t s) {
          const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
          const uint64_t key = make_key(d2, __float_as_uint(p.w));
          const bool member = active && key < st.klo;
          const float cos_i = dot_ref(w.x, w.y, w.z, nx, ny, nz);
          const bool use = member && !(cos_o < 0.0f || cos_i < 0.0f);
          const float sel = use ? 1.0f : 0.0f;
          acc_r = fmaf(sel, w.w, acc_r);
          acc_g = fmaf(sel, f.x, acc_g);
          acc_b = fmaf(sel, f.y, acc_b);
          visit(p, s, member, key);
        });
      });
    } else {
      walk_tree(D, qx, qy, qz, cull, [&](int32_t s0, int32_t s1) {
        for_slots(P, s0, s1, [&](float4 p, int32_t s) {
          const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
          const uint64_t key = make_key(d2, __float_as_uint(p.w));
          visit(p, s, active && key < st.klo, key);
        });
      });
    }
  }

  // ---- finish the set inside the stash ---------------------------------------------
  uint64_t tau = 0;
  if (reuse) {
    tau = st.klo - 1;
  } else if (D.all_members) {
    tau = key_max;
  } else if (active) {
    for (uint32_t i = 0; i < n_stash; ++i) {
      const uint64_t ki = sh.stash_key[i * kWave + lane];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < n_stash; ++j)
        rank += sh.stash_key[j * kWave + lane] < ki ? 1u : 0u;
      if (rank + 1 == st.need) tau = ki;
      if (rank < st.need) {
        const uint32_t s = sh.stash_slot[i * kWave + lane];
        if (kAccum) {
          const float4 w = D.wif4[s];
          const float2 f = D.flux2[s];
          const float cos_i = dot_ref(w.x, w.y, w.z, nx, ny, nz);
          if (!(cos_o < 0.0f || cos_i < 0.0f)) {
            acc_r += w.w;
            acc_g += f.x;
            acc_b += f.y;
          }
        } else if (MODE == kModeKnn) {
          if (n_written < static_cast<uint32_t>(D.k)) row[n_written++] = static_cast<int32_t>(s);
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
    const float cull = contributes ? cull_bound(r2) : -1.0f;
    walk_tree(D, qx, qy, qz, cull, [&](int32_t s0, int32_t s1) {
      for_slots_full(P, W, F, s0, s1, [&](float4 p, float4 w, float2 f, int32_t) {
        const float d2 = d2_ref(qx, qy, qz, p.x, p.y, p.z);
        const uint64_t key = make_key(d2, __float_as_uint(p.w));
        const float cos_i = dot_ref(w.x, w.y, w.z, nx, ny, nz);
        const bool use = contributes && key < klim && !(cos_o < 0.0f || cos_i < 0.0f);
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
  if (MODE == kModeKnn) {
    for (int32_t j = static_cast<int32_t>(n_written); j < D.k; ++j) row[j] = -1;
    return;
  }
  if (MODE == kModeRadiance) {
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
      const GatherDesc D = descs[inst];
      process_group<MODE>(D, g - gprefix[inst], L.perm, sh, lane, L.counters,
                          L.count_stats);
    }
  }
  if (MODE == kModeAdjoint) {
    __syncthreads();
    const GatherDesc D = descs[0];
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
                     2 * kXcds + 4);
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
