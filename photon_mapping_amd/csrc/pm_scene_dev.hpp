// Device-side scene code shared by the ray kernels (pm_scene.hip,
// pm_render.hip): vector helpers, the counter-based RNG, the closest-hit
// query over the BVH (or every triangle), material sampling, and the host
// record of a scene handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pm_bvh.hpp"
#include "pm_internal.hpp"
#include "pm_scene.h"

namespace pm {
namespace scn {

constexpr int kThreads = 128;  // lanes per workgroup of every ray kernel
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
  const float* le;           // emission table, 3 floats per slot
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
__device__ inline Hit closest_hit(const SceneDev& S, V3 o, V3 d, float tmax, int* stk) {
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

// Whether some triangle meets the ray at kRayEps < t < tmax: the shadow test
// of include/integrator.h:198-204 reads only whether scene.intersect found a
// hit.  It stops at the first such triangle.  Its answer equals
// closest_hit(...).tri >= 0: both run mt_test on the same operands, and every
// leaf that closest_hit reaches (boxes entered before its shrinking limit)
// is reachable here under the fixed limit tmax.
__device__ inline bool occluded(const SceneDev& S, V3 o, V3 d, float tmax, int* stk) {
  if (S.brute) {
    for (int i = 0; i < S.n_tri; ++i) {
      const V3 p0 = ld3(S.tris[i].v[0]);
      float t, u, v;
      if (mt_test(p0, ld3(S.tris[i].v[1]) - p0, ld3(S.tris[i].v[2]) - p0, o, d, t, u, v) &&
          t < tmax)
        return true;
    }
    return false;
  }
  auto leaf_any = [&](int code) {
    const int first = code >> 4, count = code & 15;
    for (int j = 0; j < count; ++j) {
      const float4* q = S.tri4 + 3 * (first + j);
      const float4 a = q[0], b = q[1], c = q[2];
      float t, u, v;
      if (mt_test(xyz(a), xyz(b), xyz(c), o, d, t, u, v) && t < tmax) return true;
    }
    return false;
  };
  const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int node = 0, sp = 0;
  for (;;) {
    const BvhNode nd = S.nodes[node];
    float tl = box_entry(nd.bx.x, nd.bx.y, nd.by.x, nd.by.y, nd.bz.x, nd.bz.y, o, inv, tmax);
    float tr = box_entry(nd.bx.z, nd.bx.w, nd.by.z, nd.by.w, nd.bz.z, nd.bz.w, o, inv, tmax);
    if (tl != INFINITY && nd.ref.x < 0) {
      if (leaf_any(~nd.ref.x)) return true;
      tl = INFINITY;
    }
    if (tr != INFINITY && nd.ref.y < 0) {
      if (leaf_any(~nd.ref.y)) return true;
      tr = INFINITY;
    }
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
  return false;
}

__device__ __forceinline__ V3 shading_normal(const pmx_triangle& tr, float u, float v) {
  return ld3(tr.n[0]) * (1.0f - u - v) + ld3(tr.n[1]) * u + ld3(tr.n[2]) * v;
}

__device__ __forceinline__ V3 geometric_normal(const pmx_triangle& tr) {
  const V3 p0 = ld3(tr.v[0]);
  return normalize(cross(ld3(tr.v[1]) - p0, ld3(tr.v[2]) - p0));
}

// Emission of a light triangle: the device copy's is_light holds its slot in
// the emission table plus one (0: not a light).
__device__ __forceinline__ V3 emission(const SceneDev& S, const pmx_triangle& tr) {
  const float* e = S.le + 3 * (tr.is_light - 1);
  return v3(e[0], e[1], e[2]);
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
__device__ inline V3 sample_bxdf(int type, float eta, V3 wo, V3 ns, bool from_light, uint64_t seed,
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


}  // namespace scn

// Host record behind a pmx_scene handle.
struct SceneImpl {
  Context* ctx = nullptr;
  int n_tri = 0, n_mat = 0, n_lights = 0;
  int bvh_nodes = 0, bvh_leaves = 0, bvh_depth = 0;
  double sah_cost = 0.0, light_area = 0.0;
  bool has_ior = false;
  DevBuf tris, tri4, nodes, mats, ior, lights, le;
  // Photon tracing's count pass keeps each path's count and first photons,
  // and records of the photons past them, here (grown as needed, reused by
  // later calls on the scene; calls on one context are externally
  // synchronised).
  mutable DevBuf trace_counts, trace_cache, trace_ovf;

  // LDS of a traversal stack for a workgroup of ray kernels: a ray pushes at
  // most one entry per internal level (bvh_depth + 1 of them; the builder
  // keeps bvh_depth <= kBvhStack - 2), entry-major across the kThreads lanes.
  // Sized to the scene, not to kBvhStack, so shallow BVHs leave LDS for
  // more resident waves.  Brute force uses none.
  size_t stack_bytes(int brute) const {
    return brute ? 0 : static_cast<size_t>(bvh_depth + 2) * scn::kThreads * sizeof(int);
  }

  scn::SceneDev dev(int brute) const {
    return scn::SceneDev{tris.as<pmx_triangle>(), tri4.as<float4>(), nodes.as<BvhNode>(),
                    n_tri, brute, mats.as<pm_material>(),
                    has_ior ? ior.as<float>() : nullptr, n_mat, lights.as<int32_t>(), n_lights,
                    le.as<float>()};
  }
};

}  // namespace pm
