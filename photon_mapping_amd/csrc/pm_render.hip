// Rendering (declared in include/pm_scene.h): the reference's photon-mapping
// and path-tracing integrators over every pixel sample.
//
// Photon mapping runs as a wavefront.  A batch of whole pixels starts as one
// camera segment per sample; every round, shade_kernel traces the round's
// segments and
//   * adds light hits and next-event estimates to their sample,
//   * stages gather queries (position, shading normal, wo, material) with
//     the segment's weight, one list per photon map,
//   * stages the next segments: specular children (both Fresnel branches
//     below depth 3, include/integrator.h:344-374) and final-gathering rays.
// Staging uses slots fixed by the segment's index; a prefix sum then compacts
// each list in segment order, so the lists' order, and with it every
// estimate, does not depend on scheduling, and no list needs an atomic
// counter.  The gather lists go through pm_gather_radiance, the library's
// photon-density estimate, and scatter_kernel adds weight * radiance to the
// samples.  Sample radiances accumulate as unsigned 32.32 fixed point, so
// their sums do not depend on the order in which segments finish.
//
// Random numbers are counter based: each segment carries a 64-bit stream id
// (the sample, re-hashed at every split) and a dimension counter.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <vector>

#include "pm_scene_dev.hpp"

namespace pm {
namespace {

using namespace scn;

constexpr double kFixScale = 4294967296.0;     // 2^32: 32.32 fixed point
constexpr float kFixLimit = 1073741824.0f;     // 2^30: larger sample terms are dropped
constexpr uint32_t kKindCamera = 0;
constexpr uint32_t kKindGather = 1;            // a final-gathering ray
constexpr int kSplitDepth = 3;                 // include/integrator.h:325, 344

// A segment's lineage (8 bits) names it among the segments of its sample:
// 1 for the camera ray, 2 l + b for branch b of a split, | 0x80 for the
// final-gathering ray of a diffuse hit.  Within one round no two segments of
// a sample share a lineage.  A query's weight carries (sample, lineage).
constexpr int kLineageBits = 8;

struct Segment {
  float o[3];
  float d[3];
  float w[3];
  uint32_t sample;  // sample index within the batch
  uint32_t dim;     // next RNG dimension of the stream
  uint32_t meta;    // depth | kind << 16 | lineage << 17
  uint64_t stream;
};

// A round's segments in memory, field-major: word f of segment i at
// w[f * stride + i] (o, d, w, sample, dim, meta, stream low, stream high), so
// a wave's loads and stores of one field are contiguous instead of 56 B apart.
constexpr int kSegWords = 14;
struct Segs {
  uint32_t* w;
  uint32_t stride;
};

__device__ __forceinline__ Segment seg_load(Segs s, uint32_t i) {
  const uint32_t* p = s.w + i;
  const uint32_t st = s.stride;
  Segment sg;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    sg.o[k] = __uint_as_float(p[k * st]);
    sg.d[k] = __uint_as_float(p[(3 + k) * st]);
    sg.w[k] = __uint_as_float(p[(6 + k) * st]);
  }
  sg.sample = p[9 * st];
  sg.dim = p[10 * st];
  sg.meta = p[11 * st];
  sg.stream = static_cast<uint64_t>(p[12 * st]) | (static_cast<uint64_t>(p[13 * st]) << 32);
  return sg;
}

__device__ __forceinline__ void seg_store(Segs s, uint32_t i, V3 o, V3 d, V3 w, uint32_t sample,
                                          uint32_t dim, uint32_t meta, uint64_t stream) {
  uint32_t* p = s.w + i;
  const uint32_t st = s.stride;
  p[0] = __float_as_uint(o.x); p[st] = __float_as_uint(o.y); p[2 * st] = __float_as_uint(o.z);
  p[3 * st] = __float_as_uint(d.x); p[4 * st] = __float_as_uint(d.y);
  p[5 * st] = __float_as_uint(d.z);
  p[6 * st] = __float_as_uint(w.x); p[7 * st] = __float_as_uint(w.y);
  p[8 * st] = __float_as_uint(w.z);
  p[9 * st] = sample;
  p[10 * st] = dim;
  p[11 * st] = meta;
  p[12 * st] = static_cast<uint32_t>(stream);
  p[13 * st] = static_cast<uint32_t>(stream >> 32);
}

struct Weighted {
  float w[3];
  uint32_t tag;     // sample << 8 | lineage
};

struct RenderDev {
  SceneDev S;
  int width, height, spp, max_depth, fg_depth;
  int caustics;  // 1: caustics gathers are wanted
  uint64_t seed;
  // given rays (6 floats each: origin, direction) in place of camera samples:
  // sample i starts with ray i (one sample per ray, width = rays, height 1)
  const float* rays;
};

// The first ray of sample sid: the given ray, or the camera's through the
// jittered sensor point of its pixel (examples/main.cpp:40-41,
// include/camera.h:41-47).
__device__ __forceinline__ void first_ray(const RenderDev& R, const pmx_camera& cam,
                                          uint64_t sid, V3& o, V3& d) {
  if (R.rays) {
    const float* r = R.rays + 6 * sid;
    o = v3(r[0], r[1], r[2]);
    d = v3(r[3], r[4], r[5]);
    return;
  }
  const int64_t pix = static_cast<int64_t>(sid / static_cast<uint64_t>(R.spp));
  const int row = static_cast<int>(pix / R.width);
  const int col = static_cast<int>(pix % R.width);
  const float su = (2.0f * (col + rnd(R.seed, sid, 0)) - R.width) / R.height;
  const float sv = (2.0f * (row + rnd(R.seed, sid, 1)) - R.height) / R.height;
  const V3 pos = ld3(cam.position), fwd = ld3(cam.forward), rgt = ld3(cam.right),
           up = ld3(cam.up);
  o = pos + rgt * su + up * sv;
  d = normalize(pos + fwd * cam.focal - o);
}

__device__ __forceinline__ V3 ldv(const float* p) { return v3(p[0], p[1], p[2]); }

// Adds one term of a sample's radiance; NaN, negative or huge terms mark the
// sample as dropped (examples/main.cpp:46-53 drops a NaN or negative sample).
__device__ __forceinline__ void add_term(unsigned long long* acc, uint32_t* bad, uint32_t s,
                                         float r, float g, float b) {
  if (!(r >= 0.0f && g >= 0.0f && b >= 0.0f && r < kFixLimit && g < kFixLimit &&
        b < kFixLimit)) {
    atomicOr(&bad[s], 1u);
    return;
  }
  if (r > 0.0f) atomicAdd(&acc[3 * s + 0], __double2ull_rn(static_cast<double>(r) * kFixScale));
  if (g > 0.0f) atomicAdd(&acc[3 * s + 1], __double2ull_rn(static_cast<double>(g) * kFixScale));
  if (b > 0.0f) atomicAdd(&acc[3 * s + 2], __double2ull_rn(static_cast<double>(b) * kFixScale));
}

// Adds v (< 2^kBits) of every lane active here to one of kRayShards 64-bit
// counters (chosen by workgroup), one atomic per wave: the sum is assembled
// bit plane by bit plane from ballots.  Spreading the atomics keeps them
// from queueing on one address; the host sums the shards.
constexpr int kRayShards = 256;
template <int kBits>
__device__ __forceinline__ void wave_count_add(unsigned long long* shards, uint32_t v) {
  const uint64_t active = __ballot(1);
  unsigned long long sum = 0;
#pragma unroll
  for (int b = 0; b < kBits; ++b)
    sum += static_cast<unsigned long long>(__popcll(__ballot((v >> b) & 1u))) << b;
  if (static_cast<int>(__lane_id()) == __builtin_ctzll(active))
    atomicAdd(&shards[blockIdx.x % kRayShards], sum);
}

// A segment's outputs go to slots fixed by its index i: next-round segments
// at 2 i and 2 i + 1 of the round's staging segments, its query of each map at
// i of that map's staging list.  Its flags (kSegCount bits, kFlagGlobal,
// kFlagCaustics) say which slots it filled; a prefix sum over the flags then
// compacts every list in segment order.  No atomics, and the order does not
// depend on scheduling.
constexpr uint32_t kSegCount = 3u;       // segments written (0, 1 or 2)
constexpr uint32_t kFlagGlobal = 4u;     // a global-map query
constexpr uint32_t kFlagCaustics = 8u;   // a caustics-map query

__device__ __forceinline__ void push_segment(Segs out, uint32_t i, uint32_t& flags, V3 o, V3 d,
                                             V3 w, uint32_t sample, uint32_t dim,
                                             uint32_t depth, uint32_t kind, uint32_t lineage,
                                             uint64_t stream) {
  const uint32_t at = 2 * i + (flags & kSegCount);
  ++flags;
  seg_store(out, at, o, d, w, sample, dim, depth | (kind << 16) | (lineage << 17), stream);
}

__device__ __forceinline__ void push_query(pm_query* q, Weighted* wq, uint32_t at, V3 x, V3 ns,
                                           V3 wo, int mat, V3 w, uint32_t sample,
                                           uint32_t lineage) {
  pm_query r;
  r.pos[0] = x.x; r.pos[1] = x.y; r.pos[2] = x.z;
  r.ns[0] = ns.x; r.ns[1] = ns.y; r.ns[2] = ns.z;
  r.wo[0] = wo.x; r.wo[1] = wo.y; r.wo[2] = wo.z;
  r.material = mat;
  q[at] = r;
  Weighted ww;
  ww.w[0] = w.x; ww.w[1] = w.y; ww.w[2] = w.z;
  ww.tag = (sample << kLineageBits) | lineage;
  wq[at] = ww;
}

__device__ __forceinline__ V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }

// Lambert evaluation (include/material.h:80-88): rho / pi when wo and wi are
// on the shading normal's side.
__device__ __forceinline__ float lambert_on(V3 wo, V3 wi, V3 ns) {
  return (dot(wo, ns) < 0.0f || dot(wi, ns) < 0.0f) ? 0.0f : 1.0f / kPiF;
}

// The child stream of a specular split.
__device__ __forceinline__ uint64_t child_stream(uint64_t stream, uint32_t branch) {
  return mix64(stream ^ (0xa24baed4963ee407ull * (branch + 1)));
}

__global__ __launch_bounds__(kThreads) void camera_kernel(RenderDev R, pmx_camera cam,
                                                          int64_t sample0, uint32_t n,
                                                          Segs out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t sid = static_cast<uint64_t>(sample0) + i;
  V3 sensor, d;
  first_ray(R, cam, sid, sensor, d);
  // camera pdf 1; depth 0, camera, lineage 1; the stream is the sample id
  seg_store(out, i, sensor, d, v3(1.0f, 1.0f, 1.0f), i, 2, 1u << 17, sid);
}

struct ShadeOut {
  Segs seg;        // staging: 2 slots per input segment
  pm_query* qg;    // staging: 1 slot per input segment, each map
  Weighted* wg;
  pm_query* qc;
  Weighted* wc;
  uint32_t* flags; // per input segment
  unsigned long long* acc;
  uint32_t* bad;
  unsigned long long* rays;  // kRayShards counters
};

// One round of PhotonMapping::integrateRecursive and
// computeIndirectIlluminationRecursive (include/integrator.h:215-252,
// 290-377) over the round's segments.
__device__ __forceinline__ uint32_t shade_one(const RenderDev& R, const Segs& in, uint32_t i,
                                              const ShadeOut& O, int* stk, uint32_t& n_rays) {
  const SceneDev& S = R.S;
  uint32_t flags = 0;
  const Segment sg = seg_load(in, i);
  V3 o = ldv(sg.o), d = ldv(sg.d), w = ldv(sg.w);
  uint32_t depth = sg.meta & 0xffffu;
  const uint32_t kind = (sg.meta >> 16) & 1u;
  const uint32_t lineage = sg.meta >> 17;
  uint32_t dim = sg.dim;
  const uint64_t stream = sg.stream;
  const uint64_t seed = R.seed;
  // A bounce with a single child (a sampled specular direction) continues
  // here rather than as a segment of the next round, with the same state the
  // segment would carry (so the same random numbers): chains of specular
  // bounces, such as total internal reflection in glass, end within one
  // round instead of keeping a few segments alive for many rounds.  depth
  // grows with every continuation and each needs depth + 1 < max_depth, so
  // the loop ends.
  for (;;) {
    // integrateRecursive returns 0 at maxDepth (include/integrator.h:291)
    if (kind == kKindCamera && static_cast<int>(depth) >= R.max_depth) return flags;
    ++n_rays;
    const Hit h = closest_hit(S, o, d, INFINITY, stk);
    if (h.tri >= 0) {
      const pmx_triangle& tri = S.tris[h.tri];
      const int m = material_of(S, tri);
      const int type = S.mats[m].type;
      const V3 rho = ldv(S.mats[m].rho);
      const V3 x = o + d * h.t;
      const V3 ns_raw = shading_normal(tri, h.u, h.v);
      const V3 ns = normalize(ns_raw);
      const V3 wo = d * -1.0f;
      const float eta = S.ior ? S.ior[m] : 1.5f;
      if (kind == kKindCamera) {
        if (tri.is_light) {
          // a light hit returns its emission (include/integrator.h:296-299)
          const V3 le = emission(S, tri);
          add_term(O.acc, O.bad, sg.sample, w.x * le.x, w.y * le.y, w.z * le.z);
        } else if (type == PM_MATERIAL_LAMBERT) {
          if (static_cast<int>(depth) >= R.fg_depth) {
            push_query(O.qg, O.wg, i, x, ns_raw, wo, tri.material, w, sg.sample, lineage);
            flags |= kFlagGlobal;
          } else {
            // direct illumination by light sampling (include/integrator.h:178-213)
            if (S.n_lights > 0) {
              int li = static_cast<int>(rnd(seed, stream, dim++) * S.n_lights);
              li = li < S.n_lights - 1 ? li : S.n_lights - 1;
              const pmx_triangle& lt = S.tris[S.lights[li]];
              const V3 l0 = ld3(lt.v[0]);
              const V3 lc = cross(ld3(lt.v[1]) - l0, ld3(lt.v[2]) - l0);
              const float area = 0.5f * sqrtf(dot(lc, lc));
              const float su0 = sqrtf(rnd(seed, stream, dim++));
              const float b0 = 1.0f - su0, b1 = rnd(seed, stream, dim++) * su0;
              const V3 lp = l0 * (1.0f - b0 - b1) + ld3(lt.v[1]) * b0 + ld3(lt.v[2]) * b1;
              const V3 lns = shading_normal(lt, b0, b1);
              const V3 to = lp - x;
              const float r = sqrtf(dot(to, to));
              const V3 wi = to * (1.0f / r);
              const float pdf_dir = (1.0f / area) * r * r / fabsf(dot(wi * -1.0f, lns));
              const bool blocked = occluded(S, x, wi, r - kRayEps, stk);
              ++n_rays;
              if (!blocked) {
                const float f = lambert_on(wo, wi, ns);
                const float c = f * fabsf(dot(wi, ns)) /
                                ((1.0f / static_cast<float>(S.n_lights)) * pdf_dir);
                const V3 le = emission(S, lt);
                add_term(O.acc, O.bad, sg.sample, w.x * rho.x * c * le.x, w.y * rho.y * c * le.y,
                         w.z * rho.z * c * le.z);
              }
            }
            // caustics by the caustics map (include/integrator.h:155-175)
            if (R.caustics)
              {
              push_query(O.qc, O.wc, i, x, ns_raw, wo, tri.material, w, sg.sample, lineage);
              flags |= kFlagCaustics;
            }
            // final gathering (include/integrator.h:215-252), recursion depth 0
            if (R.max_depth > 0) {
              float scale;
              const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
              const float c = fabsf(dot(ns, dir)) * scale;
              push_segment(O.seg, i, flags, x, dir, mul(w, rho) * c, sg.sample, dim, 0,
                           kKindGather, lineage | 0x80u, stream);
            }
          }
        } else if (static_cast<int>(depth) + 1 < R.max_depth) {
          // specular: one sampled direction from depth 3 on, every direction below
          const V3 ng = geometric_normal(tri);
          if (static_cast<int>(depth) >= kSplitDepth) {
            float scale;
            const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
            const float c = cos_term(wo, dir, ns, ng, false) * scale;
            o = x;
            d = dir;
            w = mul(w, rho) * c;
            ++depth;
            continue;
          } else {
            // sampleAllDirection (include/material.h:123-129, 203-243) in the
            // shading frame; each branch's weight is f |cos| with f = (its
            // share) * rho / |cos|
            V3 ft, fb;
            basis(ns, ft, fb);
            const V3 wol = v3(dot(wo, ft), dot(wo, ns), dot(wo, fb));
            V3 dirs[2];
            float share[2];
            int nb = 0;
            if (type == PM_MATERIAL_MIRROR) {
              dirs[0] = v3(-wol.x, wol.y, -wol.z);
              share[0] = 1.0f;
              nb = 1;
            } else {
              const bool outside = wol.y > 0.0f;
              const float eta_o = outside ? 1.0f : eta, eta_i = outside ? eta : 1.0f;
              const V3 n = outside ? v3(0.f, 1.f, 0.f) : v3(0.f, -1.f, 0.f);
              const float cos_o = dot(wol, n);
              const float fr = schlick(cos_o, eta_o, eta_i);
              dirs[0] = wol * -1.0f + n * (2.0f * cos_o);
              share[0] = fr;
              nb = 1;
              const V3 th = (wol - n * cos_o) * (-eta_o / eta_i);
              const float th2 = dot(th, th);
              if (th2 <= 1.0f) {
                dirs[1] = th + n * (-sqrtf(fmaxf(1.0f - th2, 0.0f)));
                share[1] = (1.0f - fr) * (eta_o * eta_o) / (eta_i * eta_i);
                nb = 2;
              } else {
                share[0] = 1.0f;  // total internal reflection
              }
            }
            // (both slots unrolled, so dirs and share stay in registers)
  #pragma unroll
            for (int b = 0; b < 2; ++b) {
              if (b >= nb) break;
              const V3 dl = dirs[b];
              const V3 dir = ft * dl.x + ns * dl.y + fb * dl.z;
              const float c = share[b] / fabsf(dl.y) * fabsf(dot(dir, ns));
              push_segment(O.seg, i, flags, x, dir, mul(w, rho) * c, sg.sample, 0, depth + 1,
                           kKindCamera, 2 * lineage + b, child_stream(stream, b));
            }
          }
        }
      } else {
        // a final-gathering ray: the photon map at a diffuse hit, or one more
        // sampled bounce at a specular hit
        if (type == PM_MATERIAL_LAMBERT) {
          push_query(O.qg, O.wg, i, x, ns_raw, wo, tri.material, w, sg.sample, lineage);
          flags |= kFlagGlobal;
        } else if (static_cast<int>(depth) + 1 < R.max_depth) {
          float scale;
          const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
          const float c = fabsf(dot(ns, dir)) * scale;
          o = x;
          d = dir;
          w = mul(w, rho) * c;
          ++depth;
          continue;
        }
      }
    }
    return flags;
  }
}

__global__ __launch_bounds__(kThreads) void shade_kernel(RenderDev R, Segs in, uint32_t n_in,
                                                         ShadeOut O) {
  extern __shared__ int stk_all[];  // SceneImpl::stack_bytes
  int* stk = stk_all + threadIdx.x;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n_in) return;
  uint32_t n_rays = 0;
  O.flags[i] = shade_one(R, in, i, O, stk, n_rays);
  wave_count_add<16>(O.rays, n_rays);
}

// ---- list compaction: one scan over the flag words for all three lists ----
// A chunk of kListChunk flags is taken by one block of kListThreads threads,
// wave w holding the consecutive elements [w * 1024, (w + 1) * 1024), 64 per
// round.  An element's three counts (segments written, a global query, a
// caustics query) are summed packed: 10-bit fields within a round of a wave,
// 21-bit fields within a chunk (a chunk's totals stay below 2^13).
constexpr int kListThreads = 256;
constexpr int kListItems = 16;
constexpr uint32_t kListChunk = kListThreads * kListItems;
constexpr int kListWaves = kListThreads / 64;

__device__ __forceinline__ uint32_t list_counts10(uint32_t f) {
  return (f & kSegCount) | (((f >> 2) & 1u) << 10) | (((f >> 3) & 1u) << 20);
}
__device__ __forceinline__ uint64_t widen21(uint32_t v) {
  return static_cast<uint64_t>(v & 1023u) | (static_cast<uint64_t>((v >> 10) & 1023u) << 21) |
         (static_cast<uint64_t>((v >> 20) & 1023u) << 42);
}
__device__ __forceinline__ uint32_t field21(uint64_t v, int c) {
  return static_cast<uint32_t>((v >> (21 * c)) & ((1u << 21) - 1u));
}
// inclusive prefix over the wave's lanes
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = static_cast<int>(__lane_id());
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

// Totals of each chunk's three counts: part[3 * chunk + c].
__global__ __launch_bounds__(kListThreads) void list_reduce_kernel(
    const uint32_t* __restrict__ flags, uint32_t n, uint32_t* __restrict__ part) {
  __shared__ uint64_t wt[kListWaves];
  const int lane = static_cast<int>(__lane_id());
  const int w = threadIdx.x >> 6;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kListChunk + w * 64 * kListItems;
  uint64_t sum = 0;
#pragma unroll
  for (int r = 0; r < kListItems; ++r) {
    const uint64_t e = base + r * 64 + lane;
    if (e < n) sum += widen21(list_counts10(flags[e]));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if (lane == 0) wt[w] = sum;
  __syncthreads();
  if (threadIdx.x < 3) {
    uint64_t t = 0;
    for (int i = 0; i < kListWaves; ++i) t += wt[i];
    part[3 * blockIdx.x + threadIdx.x] = field21(t, static_cast<int>(threadIdx.x));
  }
}

// Exclusive prefix of the chunk totals of each count, in place (one block),
// and the three list lengths into totals[0..2].
__global__ __launch_bounds__(kListThreads) void list_top_kernel(uint32_t* __restrict__ part,
                                                                uint32_t n_chunks,
                                                                uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[kListWaves];
  const int lane = static_cast<int>(__lane_id());
  const int w = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n_chunks; b += kListThreads) {
      const uint32_t j = b + threadIdx.x;
      const uint32_t v = j < n_chunks ? part[3 * j + c] : 0u;
      const uint32_t incl = wave_incl_scan(v);
      if (lane == 63) wsum[w] = incl;
      __syncthreads();
      uint32_t before = 0, all = 0;
      for (int i = 0; i < kListWaves; ++i) {
        if (i < w) before += wsum[i];
        all += wsum[i];
      }
      if (j < n_chunks) part[3 * j + c] = carry + before + incl - v;
      carry += all;
      __syncthreads();
    }
    if (threadIdx.x == 0) totals[c] = carry;
  }
}

// Moves the staged outputs of a shade round to their compacted positions:
// a chunk's element takes its list positions from the chunk's base (the
// scanned totals) plus the counts before it in the chunk.
__global__ __launch_bounds__(kListThreads) void compact_kernel(
    const uint32_t* __restrict__ flags, uint32_t n, const uint32_t* __restrict__ part,
    Segs stage, Segs next, const pm_query* __restrict__ qg_s, const Weighted* __restrict__ wg_s,
    pm_query* __restrict__ qg, Weighted* __restrict__ wg, const pm_query* __restrict__ qc_s,
    const Weighted* __restrict__ wc_s, pm_query* __restrict__ qc, Weighted* __restrict__ wc) {
  __shared__ uint64_t wt[kListWaves];
  const int lane = static_cast<int>(__lane_id());
  const int w = threadIdx.x >> 6;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kListChunk + w * 64 * kListItems;
  uint32_t f[kListItems];
  uint64_t before[kListItems];  // counts of the wave's earlier elements, 21-bit fields
  uint64_t run = 0;
#pragma unroll
  for (int r = 0; r < kListItems; ++r) {
    const uint64_t e = base + r * 64 + lane;
    f[r] = e < n ? flags[e] : 0u;
    const uint32_t v = list_counts10(f[r]);
    const uint32_t incl = wave_incl_scan(v);
    before[r] = run + widen21(incl - v);
    run += widen21(__shfl(incl, 63, 64));
  }
  if (lane == 0) wt[w] = run;
  __syncthreads();
  uint64_t wb = 0;
  for (int i = 0; i < w; ++i) wb += wt[i];
  const uint32_t b_seg = part[3 * blockIdx.x + 0] + field21(wb, 0);
  const uint32_t b_qg = part[3 * blockIdx.x + 1] + field21(wb, 1);
  const uint32_t b_qc = part[3 * blockIdx.x + 2] + field21(wb, 2);
#pragma unroll
  for (int r = 0; r < kListItems; ++r) {
    const uint64_t e64 = base + r * 64 + lane;
    if (e64 >= n) continue;
    const uint32_t i = static_cast<uint32_t>(e64);
    const uint32_t fl = f[r];
    const uint32_t ns = fl & kSegCount;
    const uint32_t at_seg = b_seg + field21(before[r], 0);
    {
      // a segment's kSegWords words (SoA, stride apart) through restrict pointers,
      // so the loads of later items need not wait for earlier stores
      const uint32_t* __restrict__ src = stage.w;
      uint32_t* __restrict__ dst = next.w;
      const uint32_t sst = stage.stride, dstr = next.stride;
      for (uint32_t j = 0; j < ns; ++j) {
        uint32_t v[kSegWords];
#pragma unroll
        for (int k = 0; k < kSegWords; ++k) v[k] = src[2 * i + j + k * sst];
#pragma unroll
        for (int k = 0; k < kSegWords; ++k) dst[at_seg + j + k * dstr] = v[k];
      }
    }
    if (fl & kFlagGlobal) {
      const uint32_t at = b_qg + field21(before[r], 1);
      qg[at] = qg_s[i];
      wg[at] = wg_s[i];
    }
    if (fl & kFlagCaustics) {
      const uint32_t at = b_qc + field21(before[r], 2);
      qc[at] = qc_s[i];
      wc[at] = wc_s[i];
    }
  }
}

// weight * estimated radiance into the samples
__global__ __launch_bounds__(kThreads) void scatter_kernel(const Weighted* __restrict__ wq,
                                                           const float* __restrict__ rgb,
                                                           uint32_t n,
                                                           unsigned long long* __restrict__ acc,
                                                           uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const Weighted ww = wq[i];
  add_term(acc, bad, ww.tag >> kLineageBits, ww.w[0] * rgb[3 * i + 0], ww.w[1] * rgb[3 * i + 1],
           ww.w[2] * rgb[3 * i + 2]);
}

// PathTracing::integrate (include/integrator.h:60-101), one sample per lane.
__global__ __launch_bounds__(kThreads) void path_trace_kernel(RenderDev R, pmx_camera cam,
                                                              int64_t sample0, uint32_t n,
                                                              unsigned long long* __restrict__ acc,
                                                              uint32_t* __restrict__ bad,
                                                              unsigned long long* rays) {
  extern __shared__ int stk_all[];  // SceneImpl::stack_bytes
  int* stk = stk_all + threadIdx.x;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const SceneDev& S = R.S;
  const uint64_t sid = static_cast<uint64_t>(sample0) + i;
  const uint64_t seed = R.seed, stream = sid;
  V3 o, d;
  first_ray(R, cam, sid, o, d);
  uint32_t dim = 2;
  V3 tp = v3(1.f, 1.f, 1.f);
  V3 L = v3(0.f, 0.f, 0.f);
  uint32_t n_rays = 0;
  for (int k = 0; k < R.max_depth; ++k) {
    const Hit h = closest_hit(S, o, d, INFINITY, stk);
    ++n_rays;
    if (h.tri < 0) break;
    if (k > 0) {
      const float p = fminf(fmaxf(tp.x, fmaxf(tp.y, tp.z)), 1.0f);
      if (rnd(seed, stream, dim++) >= p) break;
      tp = tp * (1.0f / p);
    }
    const pmx_triangle& tri = S.tris[h.tri];
    if (tri.is_light) L = L + mul(tp, emission(S, tri));
    const int m = material_of(S, tri);
    const V3 x = o + d * h.t;
    const V3 ns = normalize(shading_normal(tri, h.u, h.v));
    const V3 wo = d * -1.0f;
    float scale;
    const V3 dir = sample_bxdf(S.mats[m].type, S.ior ? S.ior[m] : 1.5f, wo, ns, false, seed,
                               stream, dim, scale);
    const float c = cos_term(wo, dir, ns, geometric_normal(tri), false) * scale;
    tp = mul(tp, ldv(S.mats[m].rho)) * c;
    o = x;
    d = dir;
  }
  // the sample's only term: written, not accumulated
  if (!(L.x >= 0.0f && L.y >= 0.0f && L.z >= 0.0f && L.x < kFixLimit && L.y < kFixLimit &&
        L.z < kFixLimit)) {
    bad[i] = 1u;
  } else {
    acc[3 * i + 0] = __double2ull_rn(static_cast<double>(L.x) * kFixScale);
    acc[3 * i + 1] = __double2ull_rn(static_cast<double>(L.y) * kFixScale);
    acc[3 * i + 2] = __double2ull_rn(static_cast<double>(L.z) * kFixScale);
  }
  wave_count_add<16>(rays, n_rays);
}

// Pixel means and variances of the means over the batch's samples; dropped
// samples count as zero (examples/main.cpp:46-53, 66).
__global__ __launch_bounds__(kThreads) void resolve_kernel(const unsigned long long* __restrict__ acc,
                                                           const uint32_t* __restrict__ bad,
                                                           int spp, uint32_t n_pix, int64_t pix0,
                                                           float* __restrict__ mean,
                                                           float* __restrict__ var,
                                                           unsigned long long* dropped,
                                                           int lanes) {
  // `lanes` lanes per pixel (a power of two <= 64, the least >= spp when spp
  // is below 64): lane l of a pixel takes its samples l, l + lanes, ..., so
  // a wave's loads cover consecutive samples, and the sums are then reduced
  // over the pixel's lanes in a fixed order.
  const int lane = static_cast<int>(threadIdx.x) & (lanes - 1);
  const uint32_t p = (blockIdx.x * kThreads + threadIdx.x) / static_cast<uint32_t>(lanes);
  if (p >= n_pix) return;  // (all lanes of the pixel)
  double s[3] = {0.0, 0.0, 0.0}, q[3] = {0.0, 0.0, 0.0};
  uint32_t nd = 0;
  for (int k = lane; k < spp; k += lanes) {
    const size_t si = static_cast<size_t>(p) * spp + k;
    if (bad[si]) {
      ++nd;
      continue;
    }
    for (int c = 0; c < 3; ++c) {
      const double x = static_cast<double>(acc[3 * si + c]) / kFixScale;
      s[c] += x;
      q[c] += x * x;
    }
  }
  // (a pixel's lanes are aligned within the wave, so the xor offsets below
  // stay inside them; the pixels of a wave end together)
  for (int o = lanes / 2; o >= 1; o >>= 1) {
    for (int c = 0; c < 3; ++c) {
      s[c] += __shfl_xor(s[c], o, 64);
      q[c] += __shfl_xor(q[c], o, 64);
    }
    nd += __shfl_xor(nd, o, 64);
  }
  if (lane != 0) return;
  const int64_t at = (pix0 + p) * 3;
  for (int c = 0; c < 3; ++c) {
    const double m = s[c] / spp;
    mean[at + c] = static_cast<float>(m);
    if (var) {
      const double v = spp > 1 ? (q[c] - spp * m * m) / (static_cast<double>(spp) * (spp - 1)) : 0.0;
      var[at + c] = static_cast<float>(v > 0.0 ? v : 0.0);
    }
  }
  if (nd) atomicAdd(dropped, static_cast<unsigned long long>(nd));
}

__global__ void zero_kernel(uint32_t* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

int zero_words(void* p, size_t bytes) {
  const size_t n = bytes / 4;
  if (n == 0) return PM_OK;
  hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     nullptr, static_cast<uint32_t*>(p), n);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

unsigned grid_of(uint64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }
// Lanes resolve_kernel gives a pixel of spp samples.
int resolve_lanes(int spp) {
  int l = 1;
  while (l < spp && l < 64) l <<= 1;
  return l;
}

int check_params(const pmx_scene* scene, const pmx_camera* cam, const pmx_render_params* p,
                 float* out_mean) {
  if (!scene || !cam || !p || !out_mean) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (p->out_mem != PM_MEM_DEVICE && p->out_mem != PM_MEM_HOST)
    return fail(PM_ERR_INVALID_ARG, "bad out_mem");
  if (p->width < 1 || p->height < 1 || p->spp < 1 || p->max_depth < 0 || p->max_depth > 65535)
    return fail(PM_ERR_INVALID_ARG, "bad image size, spp or depth (max_depth <= 65535)");
  if (static_cast<int64_t>(p->width) * p->height * p->spp > (int64_t{1} << 40))
    return fail(PM_ERR_INVALID_ARG, "too many samples");
  if (p->batch_pixels < 0) return fail(PM_ERR_INVALID_ARG, "negative batch size");
  return PM_OK;
}

// Device images for a render whose outputs are host memory; copies back.
struct ImageOut {
  DevBuf mean, var;
  float* d_mean = nullptr;
  float* d_var = nullptr;
  int begin(const pmx_render_params* p, float* out_mean, float* out_var) {
    if (p->out_mem == PM_MEM_DEVICE) {
      d_mean = out_mean;
      d_var = out_var;
      return PM_OK;
    }
    const size_t bytes = static_cast<size_t>(p->width) * p->height * 3 * sizeof(float);
    if (int rc = mean.reserve(bytes)) return rc;
    d_mean = mean.as<float>();
    if (out_var) {
      if (int rc = var.reserve(bytes)) return rc;
      d_var = var.as<float>();
    }
    return PM_OK;
  }
  int end(const pmx_render_params* p, float* out_mean, float* out_var) {
    if (p->out_mem == PM_MEM_DEVICE) return PM_OK;
    const size_t bytes = static_cast<size_t>(p->width) * p->height * 3 * sizeof(float);
    PM_HIP(hipMemcpy(out_mean, d_mean, bytes, hipMemcpyDeviceToHost));
    if (out_var) PM_HIP(hipMemcpy(out_var, d_var, bytes, hipMemcpyDeviceToHost));
    return PM_OK;
  }
};

// pixels per batch: whole pixels, at most 2^24 samples (or the caller's
// choice).  Large batches matter: the gather's cost per query falls as the
// queries of a launch get denser.
int64_t batch_pixels_of(const pmx_render_params* p) {
  if (p->batch_pixels > 0) return p->batch_pixels;
  return std::max<int64_t>(1, (int64_t{1} << 24) / p->spp);
}

}  // namespace
}  // namespace pm

namespace pm {
namespace {

// Grows buf to at least `need` bytes (by half again at least), keeping its
// first `keep` bytes.  The copy runs on the null stream, which the render
// uses throughout, and completes before the old storage is freed.
int grow_keep(DevBuf& buf, size_t need, size_t keep) {
  if (need <= buf.bytes) return PM_OK;
  DevBuf nb;
  if (int rc = nb.reserve(std::max(need, buf.bytes + buf.bytes / 2))) return rc;
  if (keep > 0) {
    PM_HIP(hipMemcpyAsync(nb.ptr, buf.ptr, keep, hipMemcpyDeviceToDevice, nullptr));
    PM_HIP(hipStreamSynchronize(nullptr));
  }
  std::swap(buf.ptr, nb.ptr);
  std::swap(buf.bytes, nb.bytes);
  return PM_OK;  // nb frees the old storage
}

int render_pm(const pmx_scene* scene, const pmx_camera* cam, pm_photonmap* global_map,
              pm_photonmap* caustics_map, const pmx_render_params* p, const float* in_rays,
              float* out_mean, float* out_var, pmx_render_stats* stats) {
  if (int rc = check_params(scene, cam, p, out_mean)) return rc;
  if (!global_map) return fail(PM_ERR_INVALID_ARG, "the global photon map is required");
  if (p->final_gathering_depth > 0 && !caustics_map)
    return fail(PM_ERR_INVALID_ARG, "final gathering needs the caustics photon map");
  if (p->k < 1) return fail(PM_ERR_INVALID_ARG, "k must be positive");
  const auto t_start = std::chrono::steady_clock::now();
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  ImageOut img;
  if (int rc = img.begin(p, out_mean, out_var)) return rc;
  // An empty caustics map estimates 0 everywhere (computeCausticsWithPhotonMap
  // over no photons, include/integrator.h:155-175), so its queries are not
  // made at all.
  const bool caustics_lookups =
      p->final_gathering_depth > 0 && reinterpret_cast<PhotonMapImpl*>(caustics_map)->n > 0;
  RenderDev R{s->dev(0), p->width, p->height, p->spp, p->max_depth, p->final_gathering_depth,
              caustics_lookups ? 1 : 0, p->seed, in_rays};
  const int64_t n_pix = static_cast<int64_t>(p->width) * p->height;
  const int64_t bp = std::min(batch_pixels_of(p), n_pix);
  const int64_t bsamples = bp * p->spp;
  if (bsamples > (int64_t{1} << (32 - kLineageBits)))
    return fail(PM_ERR_INVALID_ARG, "batch too large (at most 2^24 samples)");
  // Buffers grow to each round's needs: a segment spawns at most two
  // segments and one query per photon map.
  DevBuf seg[2], stage, qg_s, wg_s, qc_s, wc_s, qg, wg, qc, wc, rgb, r2, acc, bad, counters;
  DevBuf flags, part;
  int rc = acc.reserve(static_cast<size_t>(bsamples) * 3 * sizeof(unsigned long long));
  if (!rc) rc = bad.reserve(static_cast<size_t>(bsamples) * sizeof(uint32_t));
  if (!rc) rc = counters.reserve((8 + kRayShards + 8) * sizeof(unsigned long long));
  if (!rc) rc = seg[0].reserve(static_cast<size_t>(bsamples) * kSegWords * sizeof(uint32_t));
  if (rc) return rc;
  // A round of n_in segments stages at most 2 n_in segments and n_in queries
  // per map (fixed slots), then compacts them: the segments into the next
  // round's list, the queries and their weights onto the end of the batch's
  // list of each map (used[map] entries so far).  Each map's list is
  // gathered once, when the batch's rounds are done: one sort and one
  // persistent launch per map and batch instead of one per round (the late
  // rounds hold few queries, and a launch of few groups is mostly its slowest
  // waves).  The accumulators are fixed point, so the order in which the
  // terms land does not matter.
  size_t used[2] = {0, 0};
  const auto reserve_round = [&](size_t n_in, DevBuf& next) -> int {
    int e = stage.reserve(2 * n_in * kSegWords * sizeof(uint32_t));
    if (!e) e = next.reserve(2 * n_in * kSegWords * sizeof(uint32_t));
    if (!e) e = qg_s.reserve(n_in * sizeof(pm_query));
    if (!e) e = wg_s.reserve(n_in * sizeof(Weighted));
    // the batch's query lists grow by at most n_in entries each
    if (!e) e = grow_keep(qg, (used[0] + n_in) * sizeof(pm_query), used[0] * sizeof(pm_query));
    if (!e) e = grow_keep(wg, (used[0] + n_in) * sizeof(Weighted), used[0] * sizeof(Weighted));
    if (!e && R.caustics) e = qc_s.reserve(n_in * sizeof(pm_query));
    if (!e && R.caustics) e = wc_s.reserve(n_in * sizeof(Weighted));
    if (!e && R.caustics)
      e = grow_keep(qc, (used[1] + n_in) * sizeof(pm_query), used[1] * sizeof(pm_query));
    if (!e && R.caustics)
      e = grow_keep(wc, (used[1] + n_in) * sizeof(Weighted), used[1] * sizeof(Weighted));
    if (!e) e = flags.reserve(n_in * sizeof(uint32_t));
    if (!e) e = part.reserve(3 * ((n_in + kListChunk - 1) / kListChunk + 1) * sizeof(uint32_t));
    return e;
  };
  // counter words (uint32): 0-2 the lists' lengths after a round; uint64 at
  // words 8 (dropped samples) and from 16 the kRayShards ray counters
  auto* c32 = counters.as<uint32_t>();
  auto* dropped = reinterpret_cast<unsigned long long*>(c32 + 8);
  auto* rays = reinterpret_cast<unsigned long long*>(c32 + 16);
  if (int e = zero_words(counters.ptr, (8 + kRayShards + 8) * sizeof(unsigned long long)))
    return e;
  pmx_render_stats st{};
  for (int64_t pix0 = 0; pix0 < n_pix; pix0 += bp) {
    const int64_t np = std::min(bp, n_pix - pix0);
    const uint32_t ns = static_cast<uint32_t>(np * p->spp);
    if (int e = zero_words(acc.ptr, static_cast<size_t>(ns) * 3 * sizeof(unsigned long long)))
      return e;
    if (int e = zero_words(bad.ptr, static_cast<size_t>(ns) * sizeof(uint32_t))) return e;
    int cur_i = 0;
    // each segment buffer's field stride: the records it was reserved for
    uint32_t stride[2] = {ns, 0};
    if (int e = seg[cur_i].reserve(static_cast<size_t>(ns) * kSegWords * sizeof(uint32_t)))
      return e;
    hipLaunchKernelGGL(camera_kernel, dim3(grid_of(ns)), dim3(kThreads), 0, nullptr, R, *cam,
                       pix0 * p->spp, ns, Segs{seg[cur_i].as<uint32_t>(), stride[cur_i]});
    PM_HIP(hipGetLastError());
    uint32_t n_cur = ns;
    ++st.batches;
    while (n_cur > 0) {
      if (2 * static_cast<uint64_t>(n_cur) > 0xffffffffull)
        return fail(PM_ERR_UNSUPPORTED, "render round too large");
      if (int e = reserve_round(n_cur, seg[cur_i ^ 1])) return e;
      stride[cur_i ^ 1] = 2 * n_cur;
      const Segs cur{seg[cur_i].as<uint32_t>(), stride[cur_i]};
      const Segs nxt{seg[cur_i ^ 1].as<uint32_t>(), stride[cur_i ^ 1]};
      const Segs stg{stage.as<uint32_t>(), 2 * n_cur};
      ShadeOut O{stg, qg_s.as<pm_query>(), wg_s.as<Weighted>(), qc_s.as<pm_query>(),
                 wc_s.as<Weighted>(), flags.as<uint32_t>(), acc.as<unsigned long long>(),
                 bad.as<uint32_t>(), rays};
      const dim3 grid(grid_of(n_cur));
      hipLaunchKernelGGL(shade_kernel, grid, dim3(kThreads), s->stack_bytes(0), nullptr, R, cur,
                         n_cur, O);
      // the three lists' positions: chunk totals, their prefix (and the
      // lists' lengths into c32[0..2]), then the compaction itself
      const uint32_t n_chunks = static_cast<uint32_t>((n_cur + kListChunk - 1) / kListChunk);
      hipLaunchKernelGGL(list_reduce_kernel, dim3(n_chunks), dim3(kListThreads), 0, nullptr,
                         flags.as<uint32_t>(), n_cur, part.as<uint32_t>());
      hipLaunchKernelGGL(list_top_kernel, dim3(1), dim3(kListThreads), 0, nullptr,
                         part.as<uint32_t>(), n_chunks, c32);
      hipLaunchKernelGGL(compact_kernel, dim3(n_chunks), dim3(kListThreads), 0, nullptr,
                         flags.as<uint32_t>(), n_cur, part.as<uint32_t>(),
                         stg, nxt, qg_s.as<pm_query>(), wg_s.as<Weighted>(),
                         qg.as<pm_query>() + used[0], wg.as<Weighted>() + used[0],
                         qc_s.as<pm_query>(), wc_s.as<Weighted>(),
                         R.caustics ? qc.as<pm_query>() + used[1] : nullptr,
                         R.caustics ? wc.as<Weighted>() + used[1] : nullptr);
      PM_HIP(hipGetLastError());
      uint32_t cnt[3];
      PM_HIP(hipMemcpy(cnt, c32, sizeof(cnt), hipMemcpyDeviceToHost));
      ++st.iterations;
      used[0] += cnt[1];
      used[1] += cnt[2];
      cur_i ^= 1;
      n_cur = cnt[0];
    }
    // The lists are in round order and, within a round, in segment order,
    // which follows from the camera order through every compaction, so the
    // estimates do not depend on scheduling.
    for (int map_i = 0; map_i < 2; ++map_i) {
      const size_t nq = used[map_i];
      used[map_i] = 0;
      if (nq == 0) continue;
      if (nq > 0xffffffffull) return fail(PM_ERR_UNSUPPORTED, "render batch gathers too large");
      pm_photonmap* map = map_i == 0 ? global_map : caustics_map;
      const double n_emit = map_i == 0 ? p->n_emit_global : p->n_emit_caustics;
      const pm_query* q = map_i == 0 ? qg.as<pm_query>() : qc.as<pm_query>();
      const Weighted* wq = map_i == 0 ? wg.as<Weighted>() : wc.as<Weighted>();
      if (int e = rgb.reserve(nq * 3 * sizeof(float))) return e;
      if (int e = r2.reserve(nq * sizeof(float))) return e;
      if (int e = pm_gather_radiance(map, static_cast<int64_t>(nq), q, s->mats.as<pm_material>(),
                                     s->n_mat, p->k, n_emit, rgb.as<float>(), r2.as<float>(),
                                     nullptr, PM_MEM_DEVICE, nullptr))
        return e;
      hipLaunchKernelGGL(scatter_kernel, dim3(grid_of(nq)), dim3(kThreads), 0, nullptr, wq,
                         rgb.as<float>(), static_cast<uint32_t>(nq), acc.as<unsigned long long>(),
                         bad.as<uint32_t>());
      PM_HIP(hipGetLastError());
      (map_i == 0 ? st.gathers_global : st.gathers_caustics) += static_cast<int64_t>(nq);
    }
    hipLaunchKernelGGL(resolve_kernel,
                       dim3(grid_of(static_cast<uint64_t>(np) * resolve_lanes(p->spp))),
                       dim3(kThreads), 0, nullptr, acc.as<unsigned long long>(),
                       bad.as<uint32_t>(), p->spp, static_cast<uint32_t>(np), pix0, img.d_mean,
                       img.d_var, dropped, resolve_lanes(p->spp));
    PM_HIP(hipGetLastError());
  }
  unsigned long long tail[1 + kRayShards];
  PM_HIP(hipMemcpy(tail, dropped, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  PM_HIP(hipMemcpy(tail + 1, rays, kRayShards * sizeof(unsigned long long),
                   hipMemcpyDeviceToHost));
  PM_HIP(hipDeviceSynchronize());
  unsigned long long n_rays_all = 0;
  for (int k = 0; k < kRayShards; ++k) n_rays_all += tail[1 + k];
  st.samples = n_pix * p->spp;
  st.rays = static_cast<int64_t>(n_rays_all);
  st.dropped_samples = static_cast<int64_t>(tail[0]);
  st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
              .count();
  if (int rc = img.end(p, out_mean, out_var)) return rc;
  if (stats) *stats = st;
  return PM_OK;
}

int render_pt(const pmx_scene* scene, const pmx_camera* cam, const pmx_render_params* p,
              const float* in_rays, float* out_mean, float* out_var, pmx_render_stats* stats) {
  if (int rc = check_params(scene, cam, p, out_mean)) return rc;
  const auto t_start = std::chrono::steady_clock::now();
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  ImageOut img;
  if (int rc = img.begin(p, out_mean, out_var)) return rc;
  RenderDev R{s->dev(0), p->width, p->height, p->spp, p->max_depth, 0, 0, p->seed, in_rays};
  const int64_t n_pix = static_cast<int64_t>(p->width) * p->height;
  const int64_t bp = std::min(batch_pixels_of(p), n_pix);
  const int64_t bsamples = bp * p->spp;
  if (bsamples > (int64_t{1} << 28)) return fail(PM_ERR_INVALID_ARG, "batch too large");
  DevBuf acc, bad, counters;
  int rc = acc.reserve(static_cast<size_t>(bsamples) * 3 * sizeof(unsigned long long));
  if (!rc) rc = bad.reserve(static_cast<size_t>(bsamples) * sizeof(uint32_t));
  if (!rc) rc = counters.reserve((1 + kRayShards) * sizeof(unsigned long long));
  if (rc) return rc;
  // the dropped-sample count, then the kRayShards ray counters
  auto* dropped = counters.as<unsigned long long>();
  auto* rays = dropped + 1;
  if (int e = zero_words(counters.ptr, (1 + kRayShards) * sizeof(unsigned long long))) return e;
  pmx_render_stats st{};
  for (int64_t pix0 = 0; pix0 < n_pix; pix0 += bp) {
    const int64_t np = std::min(bp, n_pix - pix0);
    const uint32_t ns = static_cast<uint32_t>(np * p->spp);
    if (int e = zero_words(bad.ptr, static_cast<size_t>(ns) * sizeof(uint32_t))) return e;
    hipLaunchKernelGGL(path_trace_kernel, dim3(grid_of(ns)), dim3(kThreads), s->stack_bytes(0),
                       nullptr, R, *cam,
                       pix0 * p->spp, ns, acc.as<unsigned long long>(), bad.as<uint32_t>(), rays);
    PM_HIP(hipGetLastError());
    hipLaunchKernelGGL(resolve_kernel,
                       dim3(grid_of(static_cast<uint64_t>(np) * resolve_lanes(p->spp))),
                       dim3(kThreads), 0, nullptr, acc.as<unsigned long long>(),
                       bad.as<uint32_t>(), p->spp, static_cast<uint32_t>(np), pix0, img.d_mean,
                       img.d_var, dropped, resolve_lanes(p->spp));
    PM_HIP(hipGetLastError());
    ++st.batches;
    ++st.iterations;
  }
  unsigned long long tail[1 + kRayShards];
  PM_HIP(hipMemcpy(tail, dropped, sizeof(tail), hipMemcpyDeviceToHost));
  PM_HIP(hipDeviceSynchronize());
  unsigned long long n_rays_all = 0;
  for (int k = 0; k < kRayShards; ++k) n_rays_all += tail[1 + k];
  st.samples = n_pix * p->spp;
  st.rays = static_cast<int64_t>(n_rays_all);
  st.dropped_samples = static_cast<int64_t>(tail[0]);
  st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
              .count();
  if (int rc = img.end(p, out_mean, out_var)) return rc;
  if (stats) *stats = st;
  return PM_OK;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

int pmx_render_photon_mapping(const pmx_scene* scene, const pmx_camera* cam,
                              pm_photonmap* global_map, pm_photonmap* caustics_map,
                              const pmx_render_params* p, float* out_mean, float* out_var,
                              pmx_render_stats* stats) {
  return render_pm(scene, cam, global_map, caustics_map, p, nullptr, out_mean, out_var, stats);
}

int pmx_render_path_tracing(const pmx_scene* scene, const pmx_camera* cam,
                            const pmx_render_params* p, float* out_mean, float* out_var,
                            pmx_render_stats* stats) {
  return render_pt(scene, cam, p, nullptr, out_mean, out_var, stats);
}

int pmx_integrate_rays(const pmx_scene* scene, int integrator, pm_photonmap* global_map,
                       pm_photonmap* caustics_map, const pmx_render_params* params, int64_t n,
                       const float* rays, float* out_rgb) {
  if (!scene || !params || (n > 0 && (!rays || !out_rgb)))
    return fail(PM_ERR_INVALID_ARG, "null argument");
  if (integrator != PMX_INTEGRATOR_PATH_TRACING && integrator != PMX_INTEGRATOR_PHOTON_MAPPING)
    return fail(PM_ERR_INVALID_ARG, "unknown integrator");
  if (n < 0 || n > (int64_t{1} << 31) - 1) return fail(PM_ERR_INVALID_ARG, "bad ray count");
  if (n == 0) return PM_OK;
  if (params->out_mem != PM_MEM_DEVICE && params->out_mem != PM_MEM_HOST)
    return fail(PM_ERR_INVALID_ARG, "bad out_mem");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  // one sample per ray: a width x 1 image of one sample per pixel
  pmx_render_params p = *params;
  p.width = static_cast<int32_t>(n);
  p.height = 1;
  p.spp = 1;
  DevBuf staged;
  const float* d_rays = rays;
  if (p.out_mem == PM_MEM_HOST) {
    if (int rc = staged.reserve(static_cast<size_t>(n) * 6 * sizeof(float))) return rc;
    PM_HIP(hipMemcpy(staged.ptr, rays, static_cast<size_t>(n) * 6 * sizeof(float),
                     hipMemcpyHostToDevice));
    d_rays = staged.as<float>();
  }
  const pmx_camera unused{};
  return integrator == PMX_INTEGRATOR_PATH_TRACING
             ? render_pt(scene, &unused, &p, d_rays, out_rgb, nullptr, nullptr)
             : render_pm(scene, &unused, global_map, caustics_map, &p, d_rays, out_rgb,
                         nullptr, nullptr);
}

}  // extern "C"
