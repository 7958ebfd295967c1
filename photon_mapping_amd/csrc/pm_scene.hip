// Triangle scenes on the device (declared in include/pm_scene.h): a scene
// handle with a host-built BVH (pm_bvh.hip) traversed here per ray, batched
// closest-hit queries, photon tracing (global and caustic maps), camera /
// final-gathering query generation, and a surrogate photon array.
//
// Ray kernels run one ray or path per lane in workgroups of 128 lanes; each
// lane keeps its BVH traversal stack in LDS, laid out entry-major so that the
// 64 lanes of a wave touch 64 consecutive words.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "pm_bvh.hpp"
#include "pm_internal.hpp"
#include "pm_scene.h"

namespace pm {
namespace {

constexpr int kThreads = 128;
constexpr float kPiF = 3.14159265359f;  // include/core.h:7
constexpr float kRayEps = 1e-5f;        // include/core.h:19
constexpr int kSpecularChain = 16;      // specular bounces a final-gathering ray may follow

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// Counter-based uniform [0, 1): SplitMix64 finaliser of (seed, id, dim).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float rnd(uint64_t seed, uint64_t id, uint32_t dim) {
  const uint64_t h = mix64(seed ^ mix64(id * 0x9e3779b97f4a7c15ull + dim));
  return static_cast<float>(h >> 40) * (1.0f / 16777216.0f);
}

// Frame around a normal (include/core.h:224-231).
__device__ __forceinline__ void basis(V3 n, V3& t, V3& b) {
  if (fabsf(n.y) < 0.9f) t = normalize(cross(n, v3(0.f, 1.f, 0.f)));
  else t = normalize(cross(n, v3(0.f, 0.f, -1.f)));
  b = normalize(cross(t, n));
}

// Cosine-weighted direction in the frame (t, n, b) (include/sampler.h:84-91).
__device__ __forceinline__ V3 cosine_dir(float u, float v, V3 t, V3 n, V3 b) {
  const float c = fminf(fmaxf(1.0f - 2.0f * u, -1.0f), 1.0f);
  const float theta = 0.5f * acosf(c);
  const float phi = 2.0f * kPiF * v;
  const float lx = cosf(phi) * sinf(theta), ly = cosf(theta), lz = sinf(phi) * sinf(theta);
  return t * lx + n * ly + b * lz;
}

// ---- scene and closest hit ----------------------------------------------------

struct SceneDev {
  const pmx_triangle* tris;  // original order: normals, material, light flag
  const float4* tri4;        // leaf order: hit-test operands (pm_bvh.hpp)
  const BvhNode* nodes;
  int n_tri;
  int brute;                 // 1: test every triangle in index order
  const pm_material* mats;
  const float* ior;          // per material
  int n_mat;
  const int32_t* lights;     // light triangle indices
  int n_lights;
  float le[3];
};

struct Hit {
  float t;
  int tri;
  float u, v;
};

// Two-sided Moller-Trumbore test; true when the ray meets the triangle at
// t > kRayEps.  Both traversal methods call exactly this arithmetic on the
// same operands (v0, v1 - v0, v2 - v0), so they report identical t, u, v.
__device__ __forceinline__ bool mt_test(V3 p0, V3 e1, V3 e2, V3 o, V3 d, float& t, float& u,
                                        float& v) {
  const V3 pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (fabsf(det) < 1e-12f) return false;
  const float inv = 1.0f / det;
  const V3 tv = o - p0;
  u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return false;
  const V3 qv = cross(tv, e1);
  v = dot(d, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = dot(e2, qv) * inv;
  return t > kRayEps;
}

// Keeps the hit with the smaller (t, triangle index).
__device__ __forceinline__ void consider(Hit& h, float t, float u, float v, int tri) {
  if (t < h.t || (t == h.t && tri < h.tri)) h = Hit{t, tri, u, v};
}

__device__ __forceinline__ void leaf_hits(const SceneDev& S, int code, V3 o, V3 d, Hit& h) {
  const int first = code >> 4, count = code & 15;
  for (int j = 0; j < count; ++j) {
    const float4* q = S.tri4 + 3 * (first + j);
    const float4 a = q[0], b = q[1], c = q[2];
    float t, u, v;
    if (mt_test(xyz(a), xyz(b), xyz(c), o, d, t, u, v)) consider(h, t, u, v, __float_as_int(a.w));
  }
}

// Entry distance of the ray into one child box, +inf when missed or beyond tmax.
__device__ __forceinline__ float box_entry(float lx, float hx, float ly, float hy, float lz,
                                           float hz, V3 o, V3 inv, float tmax) {
  const float x0 = (lx - o.x) * inv.x, x1 = (hx - o.x) * inv.x;
  const float y0 = (ly - o.y) * inv.y, y1 = (hy - o.y) * inv.y;
  const float z0 = (lz - o.z) * inv.z, z1 = (hz - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
  const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
  return (tn <= tf && tf >= 0.0f && tn <= tmax) ? tn : INFINITY;
}

// Closest hit in (kRayEps, tmax), both faces (include/scene.h:307-348 with
// Embree replaced).  stk: this lane's LDS stack, entries kThreads apart.
__device__ Hit closest_hit(const SceneDev& S, V3 o, V3 d, float tmax, int* stk) {
  Hit h{tmax, -1, 0.f, 0.f};
  if (S.brute) {
    for (int i = 0; i < S.n_tri; ++i) {
      const V3 p0 = ld3(S.tris[i].v[0]);
      float t, u, v;
      if (mt_test(p0, ld3(S.tris[i].v[1]) - p0, ld3(S.tris[i].v[2]) - p0, o, d, t, u, v))
        consider(h, t, u, v, i);
    }
    return h;
  }
  const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int node = 0, sp = 0;
  for (;;) {
    const BvhNode nd = S.nodes[node];
    float tl = box_entry(nd.bx.x, nd.bx.y, nd.by.x, nd.by.y, nd.bz.x, nd.bz.y, o, inv, h.t);
    float tr = box_entry(nd.bx.z, nd.bx.w, nd.by.z, nd.by.w, nd.bz.z, nd.bz.w, o, inv, h.t);
    if (tl != INFINITY && nd.ref.x < 0) {
      leaf_hits(S, ~nd.ref.x, o, d, h);
      tl = INFINITY;
    }
    if (tr != INFINITY && nd.ref.y < 0) {
      leaf_hits(S, ~nd.ref.y, o, d, h);
      tr = INFINITY;
    }
    // a leaf may have shortened the ray: drop children now beyond it
    if (tl > h.t) tl = INFINITY;
    if (tr > h.t) tr = INFINITY;
    if (tl != INFINITY && tr != INFINITY) {
      const bool left_first = tl <= tr;
      stk[sp * kThreads] = left_first ? nd.ref.y : nd.ref.x;
      ++sp;
      node = left_first ? nd.ref.x : nd.ref.y;
    } else if (tl != INFINITY) {
      node = nd.ref.x;
    } else if (tr != INFINITY) {
      node = nd.ref.y;
    } else {
      if (sp == 0) break;
      --sp;
      node = stk[sp * kThreads];
    }
  }
  return h;
}

__device__ __forceinline__ V3 shading_normal(const pmx_triangle& tr, float u, float v) {
  return ld3(tr.n[0]) * (1.0f - u - v) + ld3(tr.n[1]) * u + ld3(tr.n[2]) * v;
}

__device__ __forceinline__ V3 geometric_normal(const pmx_triangle& tr) {
  const V3 p0 = ld3(tr.v[0]);
  return normalize(cross(ld3(tr.v[1]) - p0, ld3(tr.v[2]) - p0));
}

__device__ __forceinline__ int material_of(const SceneDev& S, const pmx_triangle& tr) {
  return (tr.material >= 0 && tr.material < S.n_mat) ? tr.material : 0;
}

// Cosine term of a bounce (include/integrator.h:25-46): zero when the
// shading and geometric normals disagree about a side.
__device__ __forceinline__ float cos_term(V3 wo, V3 wi, V3 ns, V3 ng, bool from_light) {
  const float wi_ns = dot(wi, ns), wi_ng = dot(wi, ng);
  const float wo_ns = dot(wo, ns), wo_ng = dot(wo, ng);
  if (wi_ng * wi_ns <= 0.0f || wo_ng * wo_ns <= 0.0f) return 0.0f;
  return from_light ? fabsf(wo_ns) * fabsf(wi_ng) / fabsf(wo_ng) : fabsf(wi_ns);
}

// Schlick Fresnel reflectance (include/material.h:43-48).
__device__ __forceinline__ float schlick(float cos_i, float eta_i, float eta_t) {
  const float f0 = (eta_i - eta_t) * (eta_i - eta_t) / ((eta_i + eta_t) * (eta_i + eta_t));
  const float c = fmaxf(1.0f - fabsf(cos_i), 0.0f);
  return f0 + (1.0f - f0) * c * c * c * c * c;
}

// Samples the material's direction (include/material.h:73-245) for the
// outgoing world direction wo at a surface with shading normal ns.  Writes
// the world direction and returns f / (pdf * rho): the factor the BxDF
// contributes besides its albedo and the cosine term.  Draws from dim on.
__device__ V3 sample_bxdf(int type, float eta, V3 wo, V3 ns, bool from_light, uint64_t seed,
                          uint64_t id, uint32_t& dim, float& scale) {
  V3 ft, fb;
  basis(ns, ft, fb);
  const V3 wol = v3(dot(wo, ft), dot(wo, ns), dot(wo, fb));
  V3 wil;
  if (type == PM_MATERIAL_LAMBERT) {
    wil = cosine_dir(rnd(seed, id, dim), rnd(seed, id, dim + 1), v3(1.f, 0.f, 0.f),
                     v3(0.f, 1.f, 0.f), v3(0.f, 0.f, 1.f));
    dim += 2;
    // f = rho / pi on the upper hemisphere for both directions, pdf = cos / pi
    scale = (wol.y < 0.0f || wil.y <= 0.0f) ? 0.0f : 1.0f / wil.y;
  } else if (type == PM_MATERIAL_MIRROR) {
    wil = v3(-wol.x, wol.y, -wol.z);
    scale = 1.0f / fabsf(wil.y);
  } else {
    const bool outside = wol.y > 0.0f;
    const float eta_o = outside ? 1.0f : eta, eta_i = outside ? eta : 1.0f;
    const V3 n = outside ? v3(0.f, 1.f, 0.f) : v3(0.f, -1.f, 0.f);
    const float cos_o = dot(wol, n);
    const float fr = schlick(cos_o, eta_o, eta_i);
    wil = wol * -1.0f + n * (2.0f * cos_o);
    float s = 1.0f;
    if (rnd(seed, id, dim++) >= fr) {
      const V3 th = (wol - n * cos_o) * (-eta_o / eta_i);
      const float th2 = dot(th, th);
      if (th2 <= 1.0f) {
        wil = th + n * (-sqrtf(fmaxf(1.0f - th2, 0.0f)));
        // radiance scaling for camera paths (include/material.h:176-183)
        if (!from_light) s = (eta_o * eta_o) / (eta_i * eta_i);
      }
    }
    scale = s / fabsf(wil.y);
  }
  return ft * wil.x + ns * wil.y + fb * wil.z;
}

// ---- kernels --------------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void intersect_kernel(
    SceneDev S, int64_t n, const float* __restrict__ org, const float* __restrict__ dir,
    int32_t* __restrict__ out_tri, float* __restrict__ out_t, float* __restrict__ out_u,
    float* __restrict__ out_v) {
  __shared__ int stk[kBvhStack * kThreads];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const V3 o = v3(org[i], org[n + i], org[2 * n + i]);
  const V3 d = v3(dir[i], dir[n + i], dir[2 * n + i]);
  const Hit h = closest_hit(S, o, d, INFINITY, stk + threadIdx.x);
  out_tri[i] = h.tri;
  out_t[i] = h.t;
  if (out_u) out_u[i] = h.u;
  if (out_v) out_v[i] = h.v;
}

__global__ __launch_bounds__(kThreads) void fg_queries_kernel(
    SceneDev S, pmx_camera cam, int width, int height, int spp, int depth, uint64_t seed,
    pm_query* __restrict__ out, int32_t* __restrict__ fg_mat) {
  __shared__ int stk_all[kBvhStack * kThreads];
  int* stk = stk_all + threadIdx.x;
  const int64_t id = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t total = static_cast<int64_t>(width) * height * spp;
  if (id >= total) return;
  const int64_t pix = id / spp;
  const int i = static_cast<int>(pix / width);
  const int j = static_cast<int>(pix % width);
  const V3 pos = ld3(cam.position), fwd = ld3(cam.forward), rgt = ld3(cam.right),
           up = ld3(cam.up);
  const V3 pinhole = pos + fwd * cam.focal;
  V3 qp = v3(0.f, 0.f, 0.f), qn = v3(0.f, 1.f, 0.f), qwo = v3(0.f, 1.f, 0.f);
  int qmat = 0, pmat = -1;
  uint32_t dim = 0;
  for (int attempt = 0; attempt < 64; ++attempt) {
    // camera sample (examples/main.cpp:40-41, include/camera.h:41-47)
    const float su = (2.0f * (j + rnd(seed, id, dim++)) - width) / height;
    const float sv = (2.0f * (i + rnd(seed, id, dim++)) - height) / height;
    const V3 sensor = pos + rgt * su + up * sv;
    const V3 d0 = normalize(pinhole - sensor);
    const Hit h0 = closest_hit(S, sensor, d0, INFINITY, stk);
    if (h0.tri < 0) continue;
    const pmx_triangle& t0 = S.tris[h0.tri];
    if (t0.is_light || S.mats[material_of(S, t0)].type != PM_MATERIAL_LAMBERT) continue;
    const V3 p0 = sensor + d0 * h0.t;
    const V3 n0 = shading_normal(t0, h0.u, h0.v);
    if (depth == 0) {
      qp = p0; qn = n0; qwo = d0 * -1.0f; qmat = t0.material;
      break;
    }
    V3 t, b;
    basis(n0, t, b);
    bool found = false;
    for (int bounce = 0; bounce < 16 && !found; ++bounce) {
      // final-gathering ray; specular hits are followed (include/integrator.h:215-252)
      V3 o = p0;
      V3 d = cosine_dir(rnd(seed, id, dim), rnd(seed, id, dim + 1), t, n0, b);
      dim += 2;
      for (int chain = 0; chain <= kSpecularChain; ++chain) {
        const Hit h1 = closest_hit(S, o, d, INFINITY, stk);
        if (h1.tri < 0) break;
        const pmx_triangle& t1 = S.tris[h1.tri];
        const int m1 = material_of(S, t1);
        const V3 x = o + d * h1.t;
        const V3 ns = shading_normal(t1, h1.u, h1.v);
        if (S.mats[m1].type == PM_MATERIAL_LAMBERT) {
          qp = x; qn = ns; qwo = d * -1.0f; qmat = t1.material;
          pmat = t0.material;
          found = true;
          break;
        }
        float scale;
        d = sample_bxdf(S.mats[m1].type, S.ior ? S.ior[m1] : 1.5f, d * -1.0f, normalize(ns),
                        false, seed, id, dim, scale);
        o = x;
      }
    }
    if (found) break;
  }
  pm_query q;
  q.pos[0] = qp.x; q.pos[1] = qp.y; q.pos[2] = qp.z;
  q.ns[0] = qn.x; q.ns[1] = qn.y; q.ns[2] = qn.z;
  q.wo[0] = qwo.x; q.wo[1] = qwo.y; q.wo[2] = qwo.z;
  q.material = qmat;
  out[id] = q;
  if (fg_mat) fg_mat[id] = pmat;
}

__global__ __launch_bounds__(kThreads) void surrogate_photons_kernel(
    const pmx_triangle* __restrict__ tris, const float* __restrict__ cdf, int n_tri,
    const float* __restrict__ albedo, int n_mat, float pr, float pg, float pb, int64_t n,
    uint64_t seed, float* __restrict__ pos, float* __restrict__ wi, float* __restrict__ flux) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const float ut = rnd(seed, i, 0);
  int lo = 0, hi = n_tri - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > ut) hi = mid;
    else lo = mid + 1;
  }
  const pmx_triangle& tr = tris[lo];
  // uniform point (include/triangle.h:84-114)
  const float su0 = sqrtf(rnd(seed, i, 1));
  const float b0 = 1.0f - su0, b1 = rnd(seed, i, 2) * su0;
  const V3 p = ld3(tr.v[0]) * (1.0f - b0 - b1) + ld3(tr.v[1]) * b0 + ld3(tr.v[2]) * b1;
  const V3 nn = normalize(shading_normal(tr, b0, b1));
  V3 t, b;
  basis(nn, t, b);
  const V3 w = cosine_dir(rnd(seed, i, 3), rnd(seed, i, 4), t, nn, b);
  const float s = 0.5f + 0.5f * rnd(seed, i, 5);
  const int m = (tr.material >= 0 && tr.material < n_mat) ? tr.material : 0;
  pos[i] = p.x; pos[n + i] = p.y; pos[2 * n + i] = p.z;
  wi[i] = w.x; wi[n + i] = w.y; wi[2 * n + i] = w.z;
  flux[i] = pr * albedo[3 * m + 0] * s;
  flux[n + i] = pg * albedo[3 * m + 1] * s;
  flux[2 * n + i] = pb * albedo[3 * m + 2] * s;
}

// ---- photon tracing (include/integrator.h:393-563) -------------------------------

// One photon path per lane.  COUNT: photons stored per path into counts;
// otherwise the photons are written at offsets (original photon order =
// (path, bounce) order), those at or past n_store dropped.  CAUSTICS: the
// path ends at its first Lambert hit, stored only after specular hits.
template <bool COUNT, bool CAUSTICS>
__global__ __launch_bounds__(kThreads) void trace_kernel(
    SceneDev S, int64_t path0, int64_t n_paths, uint64_t seed, int max_depth,
    uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets, int64_t n_store,
    float* __restrict__ pos, float* __restrict__ wi, float* __restrict__ flux) {
  __shared__ int stk_all[kBvhStack * kThreads];
  int* stk = stk_all + threadIdx.x;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n_paths) return;
  const uint64_t path = static_cast<uint64_t>(path0 + i);
  uint32_t dim = 0;
  // light, point and direction (include/integrator.h:265-288, include/light.h:27-44)
  int li = static_cast<int>(rnd(seed, path, dim++) * S.n_lights);
  li = li < S.n_lights - 1 ? li : S.n_lights - 1;
  const pmx_triangle& lt = S.tris[S.lights[li]];
  const V3 l0 = ld3(lt.v[0]);
  const V3 le1 = ld3(lt.v[1]) - l0, le2 = ld3(lt.v[2]) - l0;
  const V3 lc = cross(le1, le2);
  const float area = 0.5f * sqrtf(dot(lc, lc));
  const float su0 = sqrtf(rnd(seed, path, dim++));
  const float b0 = 1.0f - su0, b1 = rnd(seed, path, dim++) * su0;
  V3 o = l0 * (1.0f - b0 - b1) + ld3(lt.v[1]) * b0 + ld3(lt.v[2]) * b1;
  const V3 ln = normalize(shading_normal(lt, b0, b1));
  V3 t, b;
  basis(ln, t, b);
  V3 d = cosine_dir(rnd(seed, path, dim), rnd(seed, path, dim + 1), t, ln, b);
  dim += 2;
  // Le / (choose pdf * point pdf * direction pdf) * cos = Le * n_lights * area * pi
  const float w0 = static_cast<float>(S.n_lights) * area * kPiF;
  float tr_r = S.le[0] * w0, tr_g = S.le[1] * w0, tr_b = S.le[2] * w0;
  uint32_t stored = 0;
  for (int k = 0; k < max_depth; ++k) {
    if (!(tr_r >= 0.0f && tr_g >= 0.0f && tr_b >= 0.0f)) break;  // NaN or negative
    const Hit h = closest_hit(S, o, d, INFINITY, stk);
    if (h.tri < 0) break;
    const pmx_triangle& tri = S.tris[h.tri];
    const V3 x = o + d * h.t;
    const int m = material_of(S, tri);
    const int type = S.mats[m].type;
    const V3 wo = d * -1.0f;
    if (type == PM_MATERIAL_LAMBERT) {
      // a caustic path stores at its first Lambert hit only after a specular
      // one (k > 0), then ends either way
      if (!CAUSTICS || k > 0) {
        if (!COUNT) {
          const int64_t at = offsets[i] + static_cast<int64_t>(stored);
          if (at < n_store) {
            pos[at] = x.x; pos[n_store + at] = x.y; pos[2 * n_store + at] = x.z;
            wi[at] = wo.x; wi[n_store + at] = wo.y; wi[2 * n_store + at] = wo.z;
            flux[at] = tr_r; flux[n_store + at] = tr_g; flux[2 * n_store + at] = tr_b;
          }
        }
        ++stored;
      }
      if (CAUSTICS) break;
    }
    // Russian roulette from the second bounce
    if (k > 0) {
      const float p_rr = fminf(fmaxf(tr_r, fmaxf(tr_g, tr_b)), 1.0f);
      if (rnd(seed, path, dim++) >= p_rr) break;
      tr_r /= p_rr; tr_g /= p_rr; tr_b /= p_rr;
    }
    // next direction (include/material.h:73-245)
    const V3 ns = normalize(shading_normal(tri, h.u, h.v));
    float scale;
    const V3 wiw = sample_bxdf(type, S.ior ? S.ior[m] : 1.5f, wo, ns, true, seed, path, dim,
                               scale);
    const float c = cos_term(wo, wiw, ns, geometric_normal(tri), true) * scale;
    const pm_material& mt = S.mats[m];
    tr_r *= mt.rho[0] * c;
    tr_g *= mt.rho[1] * c;
    tr_b *= mt.rho[2] * c;
    o = x;
    d = wiw;
  }
  if (COUNT) counts[i] = stored;
}

}  // namespace

// ---- host side --------------------------------------------------------------------

struct SceneImpl {
  Context* ctx = nullptr;
  int n_tri = 0, n_mat = 0, n_lights = 0;
  int bvh_nodes = 0, bvh_leaves = 0, bvh_depth = 0;
  double sah_cost = 0.0, light_area = 0.0;
  float le[3] = {0.f, 0.f, 0.f};
  bool has_ior = false;
  DevBuf tris, tri4, nodes, mats, ior, lights;

  SceneDev dev(int brute) const {
    return SceneDev{tris.as<pmx_triangle>(), tri4.as<float4>(), nodes.as<BvhNode>(),
                    n_tri, brute, mats.as<pm_material>(),
                    has_ior ? ior.as<float>() : nullptr, n_mat, lights.as<int32_t>(), n_lights,
                    {le[0], le[1], le[2]}};
  }
};

namespace {

uint64_t kind_seed(uint64_t seed, int kind) {
  return kind == PMX_PHOTONS_CAUSTICS ? mix64(seed ^ 0xc0057105ca05715ull) : seed;
}

int launch_trace_count(const SceneImpl* s, int kind, int64_t path0, int64_t n, uint64_t seed,
                       int max_depth, uint32_t* counts) {
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads)), block(kThreads);
  const SceneDev S = s->dev(0);
  if (kind == PMX_PHOTONS_CAUSTICS)
    hipLaunchKernelGGL((trace_kernel<true, true>), grid, block, 0, nullptr, S, path0, n, seed,
                       max_depth, counts, nullptr, int64_t{0}, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((trace_kernel<true, false>), grid, block, 0, nullptr, S, path0, n, seed,
                       max_depth, counts, nullptr, int64_t{0}, nullptr, nullptr, nullptr);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

int launch_trace_write(const SceneImpl* s, int kind, int64_t n_paths, uint64_t seed,
                       int max_depth, const int64_t* offsets, int64_t n_store, float* pos,
                       float* wi, float* flux) {
  const dim3 grid(static_cast<unsigned>((n_paths + kThreads - 1) / kThreads)), block(kThreads);
  const SceneDev S = s->dev(0);
  if (kind == PMX_PHOTONS_CAUSTICS)
    hipLaunchKernelGGL((trace_kernel<false, true>), grid, block, 0, nullptr, S, int64_t{0},
                       n_paths, seed, max_depth, nullptr, offsets, n_store, pos, wi, flux);
  else
    hipLaunchKernelGGL((trace_kernel<false, false>), grid, block, 0, nullptr, S, int64_t{0},
                       n_paths, seed, max_depth, nullptr, offsets, n_store, pos, wi, flux);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

constexpr int64_t kMinChunk = int64_t{1} << 16;
constexpr int64_t kMaxChunk = int64_t{1} << 22;

// Count pass over paths [0, ...): appends per-path exclusive offsets until
// either `stop_store` photons are stored (stop_store >= 0) or `stop_paths`
// paths are counted (stop_paths >= 0).  Returns the photons stored in *acc.
int count_paths(const SceneImpl* s, int kind, uint64_t seed, int max_depth, int64_t stop_store,
                int64_t stop_paths, std::vector<int64_t>& offs, int64_t* acc_out) {
  DevBuf dcount;
  if (int rc = dcount.reserve(kMaxChunk * sizeof(uint32_t))) return rc;
  std::vector<uint32_t> cnt(kMaxChunk);
  int64_t acc = 0, paths = 0;
  int64_t chunk = kMinChunk;
  for (;;) {
    if (stop_store >= 0 && acc >= stop_store) break;
    if (stop_paths >= 0) {
      if (paths >= stop_paths) break;
      chunk = std::min(kMaxChunk, stop_paths - paths);
    }
    if (paths > (int64_t{1} << 40)) return fail(PM_ERR_INVALID_ARG, "no photon is ever stored");
    if (int rc = launch_trace_count(s, kind, paths, chunk, seed, max_depth,
                                    dcount.as<uint32_t>()))
      return rc;
    PM_HIP(hipMemcpy(cnt.data(), dcount.ptr, chunk * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < chunk; ++j) {
      if (stop_store >= 0 && acc >= stop_store) break;
      offs.push_back(acc);
      acc += cnt[j];
      ++paths;
    }
    if (stop_store >= 0 && acc < stop_store) {
      // size the next chunk from the photons per path seen so far
      const double per_path = std::max(static_cast<double>(acc) / static_cast<double>(paths), 1e-3);
      const double want = 1.1 * static_cast<double>(stop_store - acc) / per_path + 1024.0;
      chunk = std::min<int64_t>(kMaxChunk, std::max<int64_t>(kMinChunk, static_cast<int64_t>(want)));
    }
  }
  *acc_out = acc;
  return PM_OK;
}

bool valid_kind(int kind) { return kind == PMX_PHOTONS_GLOBAL || kind == PMX_PHOTONS_CAUSTICS; }

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

int pmx_scene_create(pm_context* h, const pmx_triangle* tris, int n_tri, const pm_material* mats,
                     const float* mat_ior, int n_mat, const float* light_le,
                     pmx_scene** out_scene) {
  if (!h || !tris || !mats || !light_le || !out_scene)
    return fail(PM_ERR_INVALID_ARG, "null argument");
  *out_scene = nullptr;
  if (n_tri < 1 || n_mat < 1) return fail(PM_ERR_INVALID_ARG, "a scene needs triangles and materials");
  for (int m = 0; m < n_mat; ++m) {
    const int t = mats[m].type;
    if (t != PM_MATERIAL_LAMBERT && t != PM_MATERIAL_MIRROR && t != PM_MATERIAL_GLASS)
      return fail(PM_ERR_INVALID_ARG, "unknown material type");
    if (mat_ior && !(mat_ior[m] > 0.0f)) return fail(PM_ERR_INVALID_ARG, "ior must be positive");
  }
  auto* ctx = reinterpret_cast<Context*>(h);
  PM_HIP(hipSetDevice(ctx->device));
  BvhHost bvh;
  if (int rc = build_bvh(tris, n_tri, &bvh)) return rc;
  std::vector<int32_t> lights;
  double light_area = 0.0;
  for (int i = 0; i < n_tri; ++i) {
    if (!tris[i].is_light) continue;
    lights.push_back(i);
    const float* a = tris[i].v[0];
    const float* b = tris[i].v[1];
    const float* c = tris[i].v[2];
    const double e1[3] = {double(b[0]) - a[0], double(b[1]) - a[1], double(b[2]) - a[2]};
    const double e2[3] = {double(c[0]) - a[0], double(c[1]) - a[1], double(c[2]) - a[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1];
    const double cy = e1[2] * e2[0] - e1[0] * e2[2];
    const double cz = e1[0] * e2[1] - e1[1] * e2[0];
    light_area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
  }
  auto* s = new (std::nothrow) SceneImpl();
  if (!s) return fail(PM_ERR_OOM, "host allocation failed");
  s->ctx = ctx;
  s->n_tri = n_tri;
  s->n_mat = n_mat;
  s->n_lights = static_cast<int>(lights.size());
  s->bvh_nodes = static_cast<int>(bvh.nodes.size());
  s->bvh_leaves = bvh.leaves;
  s->bvh_depth = bvh.depth;
  s->sah_cost = bvh.sah_cost;
  s->light_area = light_area;
  for (int a = 0; a < 3; ++a) s->le[a] = light_le[a];
  s->has_ior = mat_ior != nullptr;
  const auto upload = [](DevBuf& buf, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PM_OK;
    if (int rc = buf.reserve(bytes)) return rc;
    PM_HIP(hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice));
    return PM_OK;
  };
  int rc = upload(s->tris, tris, n_tri * sizeof(pmx_triangle));
  if (!rc) rc = upload(s->tri4, bvh.tri4.data(), bvh.tri4.size() * sizeof(float4));
  if (!rc) rc = upload(s->nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(BvhNode));
  if (!rc) rc = upload(s->mats, mats, n_mat * sizeof(pm_material));
  if (!rc && mat_ior) rc = upload(s->ior, mat_ior, n_mat * sizeof(float));
  if (!rc) rc = upload(s->lights, lights.data(), lights.size() * sizeof(int32_t));
  if (rc) {
    delete s;
    return rc;
  }
  *out_scene = reinterpret_cast<pmx_scene*>(s);
  return PM_OK;
}

int pmx_scene_destroy(pmx_scene* scene) {
  delete reinterpret_cast<SceneImpl*>(scene);
  return PM_OK;
}

int pmx_scene_get_info(const pmx_scene* scene, pmx_scene_info* out) {
  if (!scene || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  *out = pmx_scene_info{s->n_tri, s->n_lights, s->bvh_nodes, s->bvh_leaves, s->bvh_depth, 0,
                        s->sah_cost, s->light_area};
  return PM_OK;
}

int pmx_scene_intersect(const pmx_scene* scene, int64_t n, const float* org_soa,
                        const float* dir_soa, int method, int32_t* out_tri, float* out_t,
                        float* out_u, float* out_v, void* stream) {
  if (!scene) return fail(PM_ERR_INVALID_ARG, "null scene");
  if (n < 0) return fail(PM_ERR_INVALID_ARG, "negative ray count");
  if (method != PMX_HIT_BVH && method != PMX_HIT_BRUTE_FORCE)
    return fail(PM_ERR_INVALID_ARG, "unknown hit method");
  if (n == 0) return PM_OK;
  if (!org_soa || !dir_soa || !out_tri || !out_t) return fail(PM_ERR_INVALID_ARG, "null array");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  hipLaunchKernelGGL(intersect_kernel, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, static_cast<hipStream_t>(stream), s->dev(method), n,
                     org_soa, dir_soa, out_tri, out_t, out_u, out_v);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

int pmx_fg_queries(const pmx_scene* scene, const pmx_camera* cam, int width, int height, int spp,
                   int depth, uint64_t seed, pm_query* out, int32_t* fg_mat_opt) {
  if (!scene || !cam || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (width < 1 || height < 1 || spp < 1 || (depth != 0 && depth != 1))
    return fail(PM_ERR_INVALID_ARG, "bad image size, spp or depth");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  const int64_t total = static_cast<int64_t>(width) * height * spp;
  hipLaunchKernelGGL(fg_queries_kernel, dim3(static_cast<unsigned>((total + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, nullptr, s->dev(0), *cam, width, height, spp, depth, seed,
                     out, fg_mat_opt);
  PM_HIP(hipGetLastError());
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

int pmx_trace_photons(const pmx_scene* scene, int kind, int64_t n_store, int max_depth,
                      uint64_t seed, float* pos_soa, float* wi_soa, float* flux_soa,
                      int64_t* n_emit) {
  if (!scene || !pos_soa || !wi_soa || !flux_soa || !n_emit)
    return fail(PM_ERR_INVALID_ARG, "null argument");
  if (!valid_kind(kind)) return fail(PM_ERR_INVALID_ARG, "unknown photon kind");
  if (n_store < 1 || max_depth < 1) return fail(PM_ERR_INVALID_ARG, "bad sizes");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  if (s->n_lights == 0) return fail(PM_ERR_INVALID_ARG, "the scene has no light triangle");
  PM_HIP(hipSetDevice(s->ctx->device));
  const uint64_t kseed = kind_seed(seed, kind);
  // pass 1: photons stored per path until n_store are reached
  std::vector<int64_t> offs;
  int64_t acc = 0;
  if (int rc = count_paths(s, kind, kseed, max_depth, n_store, -1, offs, &acc)) return rc;
  // the path holding photon n_store - 1 is the last one emitted
  const int64_t paths = static_cast<int64_t>(offs.size());
  *n_emit = paths;
  // pass 2: the same paths again, writing their photons in (path, bounce) order
  DevBuf doff;
  if (int rc = doff.reserve(offs.size() * sizeof(int64_t))) return rc;
  PM_HIP(hipMemcpy(doff.ptr, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  if (int rc = launch_trace_write(s, kind, paths, kseed, max_depth, doff.as<int64_t>(), n_store,
                                  pos_soa, wi_soa, flux_soa))
    return rc;
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

int pmx_trace_photon_map(const pmx_scene* scene, int kind, int64_t n_paths, int max_depth,
                         uint64_t seed, pm_photonmap** out_map, int64_t* n_stored) {
  if (!scene || !out_map || !n_stored) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out_map = nullptr;
  if (!valid_kind(kind)) return fail(PM_ERR_INVALID_ARG, "unknown photon kind");
  if (n_paths < 0 || max_depth < 1) return fail(PM_ERR_INVALID_ARG, "bad sizes");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  if (s->n_lights == 0 && n_paths > 0)
    return fail(PM_ERR_INVALID_ARG, "the scene has no light triangle");
  PM_HIP(hipSetDevice(s->ctx->device));
  const uint64_t kseed = kind_seed(seed, kind);
  std::vector<int64_t> offs;
  int64_t total = 0;
  if (n_paths > 0)
    if (int rc = count_paths(s, kind, kseed, max_depth, -1, n_paths, offs, &total)) return rc;
  if (total > (int64_t{1} << 31) - 1) return fail(PM_ERR_INVALID_ARG, "more than 2^31 photons stored");
  *n_stored = total;
  DevBuf doff, dpos, dwi, dflux;
  if (total > 0) {
    if (int rc = doff.reserve(offs.size() * sizeof(int64_t))) return rc;
    const size_t bytes = 3 * static_cast<size_t>(total) * sizeof(float);
    if (int rc = dpos.reserve(bytes)) return rc;
    if (int rc = dwi.reserve(bytes)) return rc;
    if (int rc = dflux.reserve(bytes)) return rc;
    PM_HIP(hipMemcpy(doff.ptr, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (int rc = launch_trace_write(s, kind, n_paths, kseed, max_depth, doff.as<int64_t>(), total,
                                    dpos.as<float>(), dwi.as<float>(), dflux.as<float>()))
      return rc;
    PM_HIP(hipDeviceSynchronize());
  }
  return pm_photonmap_create(reinterpret_cast<pm_context*>(s->ctx), total, dpos.as<float>(),
                             dwi.as<float>(), dflux.as<float>(), PM_MEM_DEVICE, out_map);
}

int pmx_surrogate_photons(pm_context* h, const pmx_triangle* tris, int n_tri,
                          const float* mat_albedo, int n_mat, const float* power, int64_t n,
                          uint64_t seed, float* pos_soa, float* wi_soa, float* flux_soa) {
  if (!h || !tris || !mat_albedo || !power || !pos_soa || !wi_soa || !flux_soa)
    return fail(PM_ERR_INVALID_ARG, "null argument");
  if (n_tri < 1 || n_mat < 1 || n < 1) return fail(PM_ERR_INVALID_ARG, "bad sizes");
  auto* ctx = reinterpret_cast<Context*>(h);
  PM_HIP(hipSetDevice(ctx->device));
  // area CDF on the host
  std::vector<float> cdf(n_tri);
  double acc = 0.0;
  for (int i = 0; i < n_tri; ++i) {
    const float* a = tris[i].v[0];
    const float* b = tris[i].v[1];
    const float* c = tris[i].v[2];
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1];
    const double cy = e1[2] * e2[0] - e1[0] * e2[2];
    const double cz = e1[0] * e2[1] - e1[1] * e2[0];
    acc += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    cdf[i] = static_cast<float>(acc);
  }
  for (auto& v : cdf) v = static_cast<float>(v / acc);
  cdf[n_tri - 1] = 1.0f;
  DevBuf dt, dc, da;
  if (int rc = dt.reserve(n_tri * sizeof(pmx_triangle))) return rc;
  if (int rc = dc.reserve(n_tri * sizeof(float))) return rc;
  if (int rc = da.reserve(3 * n_mat * sizeof(float))) return rc;
  PM_HIP(hipMemcpy(dt.ptr, tris, n_tri * sizeof(pmx_triangle), hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(dc.ptr, cdf.data(), n_tri * sizeof(float), hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(da.ptr, mat_albedo, 3 * n_mat * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(surrogate_photons_kernel, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, nullptr, dt.as<pmx_triangle>(), dc.as<float>(), n_tri,
                     da.as<float>(), n_mat, power[0], power[1], power[2], n, seed, pos_soa,
                     wi_soa, flux_soa);
  PM_HIP(hipGetLastError());
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

}  // extern "C"
