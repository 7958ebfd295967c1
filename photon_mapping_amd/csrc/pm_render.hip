// Rendering (declared in include/pm_scene.h): the reference's photon-mapping
// and path-tracing integrators over every pixel sample.
//
// Photon mapping runs as a wavefront.  A batch of whole pixels starts as one
// camera segment per sample; every round, shade_kernel traces the round's
// segments and
//   * adds light hits and next-event estimates to their sample,
//   * appends gather queries (position, shading normal, wo, material) with
//     the segment's weight, one list per photon map,
//   * appends the next segments: specular children (both Fresnel branches
//     below depth 3, include/integrator.h:344-374) and final-gathering rays.
// The round's gather lists are sorted by (sample, lineage), go through
// pm_gather_radiance, the library's photon-density estimate, and
// scatter_kernel adds weight * radiance to the samples.  Sample radiances accumulate as unsigned 32.32 fixed point, so
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
// a sample share a lineage, so (sample, lineage) orders a round's gather
// queries the same way however the segments were scheduled.
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

struct Weighted {
  float w[3];
  uint32_t tag;     // sample << 8 | lineage: the query's sort key
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

// Slot for each lane active here in an append-only list: one atomic per wave
// (the lowest active lane adds the active count), ranks by lane order.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter) {
  const uint64_t active = __ballot(1);
  const int leader = __builtin_ctzll(active);
  const int lane = static_cast<int>(__lane_id());
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(active)));
  base = __shfl(base, leader);
  return base + static_cast<uint32_t>(__popcll(active & ((1ull << lane) - 1ull)));
}

// Adds v (< 2^kBits) of every lane active here to a 64-bit counter, one
// atomic per wave: the sum is assembled bit plane by bit plane from ballots.
template <int kBits>
__device__ __forceinline__ void wave_count_add(unsigned long long* counter, uint32_t v) {
  const uint64_t active = __ballot(1);
  unsigned long long sum = 0;
#pragma unroll
  for (int b = 0; b < kBits; ++b)
    sum += static_cast<unsigned long long>(__popcll(__ballot((v >> b) & 1u))) << b;
  if (static_cast<int>(__lane_id()) == __builtin_ctzll(active)) atomicAdd(counter, sum);
}

__device__ __forceinline__ void push_segment(Segment* out, uint32_t* n_out, uint32_t cap, V3 o,
                                             V3 d, V3 w, uint32_t sample, uint32_t dim,
                                             uint32_t depth, uint32_t kind, uint32_t lineage,
                                             uint64_t stream, uint32_t* overflow) {
  const uint32_t at = wave_append(n_out);
  if (at >= cap) {
    atomicOr(overflow, 1u);
    return;
  }
  Segment& sg = out[at];
  sg.o[0] = o.x; sg.o[1] = o.y; sg.o[2] = o.z;
  sg.d[0] = d.x; sg.d[1] = d.y; sg.d[2] = d.z;
  sg.w[0] = w.x; sg.w[1] = w.y; sg.w[2] = w.z;
  sg.sample = sample;
  sg.dim = dim;
  sg.meta = depth | (kind << 16) | (lineage << 17);
  sg.stream = stream;
}

__device__ __forceinline__ void push_query(pm_query* q, Weighted* wq, uint32_t* n_q, uint32_t cap,
                                           V3 x, V3 ns, V3 wo, int mat, V3 w, uint32_t sample,
                                           uint32_t lineage, uint32_t* overflow) {
  const uint32_t at = wave_append(n_q);
  if (at >= cap) {
    atomicOr(overflow, 1u);
    return;
  }
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
                                                          Segment* __restrict__ out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint64_t sid = static_cast<uint64_t>(sample0) + i;
  const uint64_t stream = sid;
  V3 sensor, d;
  first_ray(R, cam, sid, sensor, d);
  Segment& sg = out[i];
  sg.o[0] = sensor.x; sg.o[1] = sensor.y; sg.o[2] = sensor.z;
  sg.d[0] = d.x; sg.d[1] = d.y; sg.d[2] = d.z;
  sg.w[0] = 1.0f; sg.w[1] = 1.0f; sg.w[2] = 1.0f;  // camera pdf 1
  sg.sample = i;
  sg.dim = 2;
  sg.meta = 1u << 17;  // depth 0, camera, lineage 1
  sg.stream = stream;
}

struct ShadeOut {
  Segment* seg;
  uint32_t* n_seg;
  uint32_t seg_cap;
  pm_query* qg;
  Weighted* wg;
  uint32_t* n_qg;
  pm_query* qc;
  Weighted* wc;
  uint32_t* n_qc;
  uint32_t q_cap;
  unsigned long long* acc;
  uint32_t* bad;
  uint32_t* overflow;
  unsigned long long* rays;
};

// One round of PhotonMapping::integrateRecursive and
// computeIndirectIlluminationRecursive (include/integrator.h:215-252,
// 290-377) over the round's segments.
__global__ __launch_bounds__(kThreads) void shade_kernel(RenderDev R,
                                                         const Segment* __restrict__ in,
                                                         uint32_t n_in, ShadeOut O) {
  extern __shared__ int stk_all[];  // SceneImpl::stack_bytes
  int* stk = stk_all + threadIdx.x;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n_in) return;
  const SceneDev& S = R.S;
  const Segment sg = in[i];
  const V3 o = ldv(sg.o), d = ldv(sg.d), w = ldv(sg.w);
  const uint32_t depth = sg.meta & 0xffffu, kind = (sg.meta >> 16) & 1u;
  const uint32_t lineage = sg.meta >> 17;
  uint32_t dim = sg.dim;
  const uint64_t stream = sg.stream;
  const uint64_t seed = R.seed;
  // integrateRecursive returns 0 at maxDepth (include/integrator.h:291)
  if (kind == kKindCamera && static_cast<int>(depth) >= R.max_depth) return;
  uint32_t n_rays = 1;
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
          push_query(O.qg, O.wg, O.n_qg, O.q_cap, x, ns_raw, wo, tri.material, w, sg.sample,
                     lineage, O.overflow);
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
            push_query(O.qc, O.wc, O.n_qc, O.q_cap, x, ns_raw, wo, tri.material, w, sg.sample,
                       lineage, O.overflow);
          // final gathering (include/integrator.h:215-252), recursion depth 0
          if (R.max_depth > 0) {
            float scale;
            const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
            const float c = fabsf(dot(ns, dir)) * scale;
            push_segment(O.seg, O.n_seg, O.seg_cap, x, dir, mul(w, rho) * c, sg.sample, dim,
                         0, kKindGather, lineage | 0x80u, stream, O.overflow);
          }
        }
      } else if (static_cast<int>(depth) + 1 < R.max_depth) {
        // specular: one sampled direction from depth 3 on, every direction below
        const V3 ng = geometric_normal(tri);
        if (static_cast<int>(depth) >= kSplitDepth) {
          float scale;
          const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
          const float c = cos_term(wo, dir, ns, ng, false) * scale;
          push_segment(O.seg, O.n_seg, O.seg_cap, x, dir, mul(w, rho) * c, sg.sample, dim,
                       depth + 1, kKindCamera, lineage, stream, O.overflow);
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
            push_segment(O.seg, O.n_seg, O.seg_cap, x, dir, mul(w, rho) * c, sg.sample, 0,
                         depth + 1, kKindCamera, 2 * lineage + b, child_stream(stream, b),
                         O.overflow);
          }
        }
      }
    } else {
      // a final-gathering ray: the photon map at a diffuse hit, or one more
      // sampled bounce at a specular hit
      if (type == PM_MATERIAL_LAMBERT) {
        push_query(O.qg, O.wg, O.n_qg, O.q_cap, x, ns_raw, wo, tri.material, w, sg.sample,
                   lineage, O.overflow);
      } else if (static_cast<int>(depth) + 1 < R.max_depth) {
        float scale;
        const V3 dir = sample_bxdf(type, eta, wo, ns, false, seed, stream, dim, scale);
        const float c = fabsf(dot(ns, dir)) * scale;
        push_segment(O.seg, O.n_seg, O.seg_cap, x, dir, mul(w, rho) * c, sg.sample, dim,
                     depth + 1, kKindGather, lineage, stream, O.overflow);
      }
    }
  }
  wave_count_add<2>(O.rays, n_rays);
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

__global__ __launch_bounds__(kThreads) void tag_kernel(const Weighted* __restrict__ wq, uint32_t n,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  keys[i] = wq[i].tag;
  vals[i] = i;
}

// The round's queries in (sample, lineage) order.
__global__ __launch_bounds__(kThreads) void permute_kernel(const uint32_t* __restrict__ perm,
                                                           uint32_t n,
                                                           const pm_query* __restrict__ q,
                                                           const Weighted* __restrict__ wq,
                                                           pm_query* __restrict__ q_out,
                                                           Weighted* __restrict__ w_out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  q_out[i] = q[j];
  w_out[i] = wq[j];
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
                                                           unsigned long long* dropped) {
  const uint32_t p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n_pix) return;
  double s[3] = {0.0, 0.0, 0.0}, q[3] = {0.0, 0.0, 0.0};
  uint32_t nd = 0;
  for (int k = 0; k < spp; ++k) {
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
  int sample_bits = 1;
  while ((int64_t{1} << sample_bits) < bsamples) ++sample_bits;
  const int key_bits = sample_bits + kLineageBits;
  // Buffers grow to each round's needs: a segment spawns at most two
  // segments and one query per photon map.
  DevBuf seg[2], qg, wg, qc, wc, rgb, r2, acc, bad, counters;
  DevBuf keys_a, keys_b, vals_a, vals_b, q_sorted, w_sorted;
  Workspace ws;
  int rc = acc.reserve(static_cast<size_t>(bsamples) * 3 * sizeof(unsigned long long));
  if (!rc) rc = bad.reserve(static_cast<size_t>(bsamples) * sizeof(uint32_t));
  if (!rc) rc = counters.reserve(16 * sizeof(unsigned long long));
  if (!rc) rc = seg[0].reserve(static_cast<size_t>(bsamples) * sizeof(Segment));
  if (rc) return rc;
  const auto reserve_round = [&](size_t n_in, DevBuf& out) -> int {
    const size_t q = n_in;
    int e = out.reserve(2 * n_in * sizeof(Segment));
    if (!e) e = qg.reserve(q * sizeof(pm_query));
    if (!e) e = wg.reserve(q * sizeof(Weighted));
    if (!e && R.caustics) e = qc.reserve(q * sizeof(pm_query));
    if (!e && R.caustics) e = wc.reserve(q * sizeof(Weighted));
    if (!e) e = rgb.reserve(q * 3 * sizeof(float));
    if (!e) e = r2.reserve(q * sizeof(float));
    if (!e) e = keys_a.reserve(q * sizeof(uint32_t));
    if (!e) e = keys_b.reserve(q * sizeof(uint32_t));
    if (!e) e = vals_a.reserve(q * sizeof(uint32_t));
    if (!e) e = vals_b.reserve(q * sizeof(uint32_t));
    if (!e) e = q_sorted.reserve(q * sizeof(pm_query));
    if (!e) e = w_sorted.reserve(q * sizeof(Weighted));
    if (!e) e = sort_reserve(&ws, static_cast<int64_t>(q));
    return e;
  };
  // counter words (uint32): 0 segments out, 1 global queries, 2 caustics
  // queries, 3 overflow flag; uint64 at words 8 (rays) and 10 (dropped samples)
  auto* c32 = counters.as<uint32_t>();
  auto* rays = reinterpret_cast<unsigned long long*>(c32 + 8);
  auto* dropped = reinterpret_cast<unsigned long long*>(c32 + 10);
  if (int e = zero_words(counters.ptr, 16 * sizeof(unsigned long long))) return e;
  pmx_render_stats st{};
  for (int64_t pix0 = 0; pix0 < n_pix; pix0 += bp) {
    const int64_t np = std::min(bp, n_pix - pix0);
    const uint32_t ns = static_cast<uint32_t>(np * p->spp);
    if (int e = zero_words(acc.ptr, static_cast<size_t>(ns) * 3 * sizeof(unsigned long long)))
      return e;
    if (int e = zero_words(bad.ptr, static_cast<size_t>(ns) * sizeof(uint32_t))) return e;
    int cur_i = 0;
    if (int e = seg[cur_i].reserve(static_cast<size_t>(ns) * sizeof(Segment))) return e;
    hipLaunchKernelGGL(camera_kernel, dim3(grid_of(ns)), dim3(kThreads), 0, nullptr, R, *cam,
                       pix0 * p->spp, ns, seg[cur_i].as<Segment>());
    PM_HIP(hipGetLastError());
    uint32_t n_cur = ns;
    ++st.batches;
    while (n_cur > 0) {
      if (2 * static_cast<uint64_t>(n_cur) > 0xffffffffull)
        return fail(PM_ERR_UNSUPPORTED, "render round too large");
      if (int e = reserve_round(n_cur, seg[cur_i ^ 1])) return e;
      Segment* cur = seg[cur_i].as<Segment>();
      Segment* nxt = seg[cur_i ^ 1].as<Segment>();
      if (int e = zero_words(c32, 4 * sizeof(uint32_t))) return e;
      ShadeOut O{nxt, c32 + 0, 2 * n_cur, qg.as<pm_query>(), wg.as<Weighted>(), c32 + 1,
                 qc.as<pm_query>(), wc.as<Weighted>(), c32 + 2, n_cur,
                 acc.as<unsigned long long>(), bad.as<uint32_t>(), c32 + 3, rays};
      hipLaunchKernelGGL(shade_kernel, dim3(grid_of(n_cur)), dim3(kThreads), s->stack_bytes(0),
                         nullptr, R, cur,
                         n_cur, O);
      PM_HIP(hipGetLastError());
      uint32_t cnt[4];
      PM_HIP(hipMemcpy(cnt, c32, sizeof(cnt), hipMemcpyDeviceToHost));
      // the capacities bound what a round can produce, so this is a guard
      if (cnt[3]) return fail(PM_ERR_UNSUPPORTED, "render segment or query list overflow");
      ++st.iterations;
      for (int map_i = 0; map_i < 2; ++map_i) {
        const uint32_t nq = cnt[1 + map_i];
        if (nq == 0) continue;
        pm_photonmap* map = map_i == 0 ? global_map : caustics_map;
        const double n_emit = map_i == 0 ? p->n_emit_global : p->n_emit_caustics;
        const pm_query* q = map_i == 0 ? qg.as<pm_query>() : qc.as<pm_query>();
        const Weighted* wq = map_i == 0 ? wg.as<Weighted>() : wc.as<Weighted>();
        // (sample, lineage) order, so the estimates do not depend on scheduling
        hipLaunchKernelGGL(tag_kernel, dim3(grid_of(nq)), dim3(kThreads), 0, nullptr, wq, nq,
                           keys_a.as<uint32_t>(), vals_a.as<uint32_t>());
        PM_HIP(hipGetLastError());
        int which = 0;
        if (int e = radix_sort_u32(&ws, keys_a.as<uint32_t>(), vals_a.as<uint32_t>(),
                                   keys_b.as<uint32_t>(), vals_b.as<uint32_t>(), nq, key_bits,
                                   nullptr, &which))
          return e;
        hipLaunchKernelGGL(permute_kernel, dim3(grid_of(nq)), dim3(kThreads), 0, nullptr,
                           which ? vals_b.as<uint32_t>() : vals_a.as<uint32_t>(), nq, q, wq,
                           q_sorted.as<pm_query>(), w_sorted.as<Weighted>());
        PM_HIP(hipGetLastError());
        if (int e = pm_gather_radiance(map, nq, q_sorted.as<pm_query>(),
                                       s->mats.as<pm_material>(), s->n_mat, p->k, n_emit,
                                       rgb.as<float>(), r2.as<float>(), nullptr, PM_MEM_DEVICE,
                                       nullptr))
          return e;
        hipLaunchKernelGGL(scatter_kernel, dim3(grid_of(nq)), dim3(kThreads), 0, nullptr,
                           w_sorted.as<Weighted>(), rgb.as<float>(), nq,
                           acc.as<unsigned long long>(), bad.as<uint32_t>());
        PM_HIP(hipGetLastError());
        (map_i == 0 ? st.gathers_global : st.gathers_caustics) += nq;
      }
      cur_i ^= 1;
      n_cur = cnt[0];
    }
    hipLaunchKernelGGL(resolve_kernel, dim3(grid_of(np)), dim3(kThreads), 0, nullptr,
                       acc.as<unsigned long long>(), bad.as<uint32_t>(), p->spp,
                       static_cast<uint32_t>(np), pix0, img.d_mean, img.d_var, dropped);
    PM_HIP(hipGetLastError());
  }
  unsigned long long tail[2];
  PM_HIP(hipMemcpy(tail, rays, sizeof(tail), hipMemcpyDeviceToHost));
  PM_HIP(hipDeviceSynchronize());
  st.samples = n_pix * p->spp;
  st.rays = static_cast<int64_t>(tail[0]);
  st.dropped_samples = static_cast<int64_t>(tail[1]);
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
  if (!rc) rc = counters.reserve(4 * sizeof(unsigned long long));
  if (rc) return rc;
  auto* rays = counters.as<unsigned long long>();
  if (int e = zero_words(counters.ptr, 4 * sizeof(unsigned long long))) return e;
  pmx_render_stats st{};
  for (int64_t pix0 = 0; pix0 < n_pix; pix0 += bp) {
    const int64_t np = std::min(bp, n_pix - pix0);
    const uint32_t ns = static_cast<uint32_t>(np * p->spp);
    if (int e = zero_words(bad.ptr, static_cast<size_t>(ns) * sizeof(uint32_t))) return e;
    hipLaunchKernelGGL(path_trace_kernel, dim3(grid_of(ns)), dim3(kThreads), s->stack_bytes(0),
                       nullptr, R, *cam,
                       pix0 * p->spp, ns, acc.as<unsigned long long>(), bad.as<uint32_t>(), rays);
    PM_HIP(hipGetLastError());
    hipLaunchKernelGGL(resolve_kernel, dim3(grid_of(np)), dim3(kThreads), 0, nullptr,
                       acc.as<unsigned long long>(), bad.as<uint32_t>(), p->spp,
                       static_cast<uint32_t>(np), pix0, img.d_mean, img.d_var, rays + 1);
    PM_HIP(hipGetLastError());
    ++st.batches;
    ++st.iterations;
  }
  unsigned long long tail[2];
  PM_HIP(hipMemcpy(tail, rays, sizeof(tail), hipMemcpyDeviceToHost));
  PM_HIP(hipDeviceSynchronize());
  st.samples = n_pix * p->spp;
  st.rays = static_cast<int64_t>(tail[0]);
  st.dropped_samples = static_cast<int64_t>(tail[1]);
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
