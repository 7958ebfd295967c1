// Workload producers over a small triangle scene (declared in
// include/pm_scene.h): camera/final-gathering query generation, photon
// tracing, and a surrogate photon array.  Brute-force closest hit over all
// triangles: the scenes these serve have a few dozen triangles.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "pm_internal.hpp"
#include "pm_scene.h"

namespace pm {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxSceneTris = 4096;
constexpr float kPiF = 3.14159265359f;  // include/core.h:7
constexpr float kRayEps = 1e-5f;        // include/core.h:19

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

// Counter-based uniform [0, 1): SplitMix64 finaliser of (seed, id, dim).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
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

struct Hit {
  float t;
  int tri;
  float u, v;
};

// Closest hit in (kRayEps, tmax), both faces (Moller-Trumbore).
__device__ Hit closest_hit(const pmx_triangle* __restrict__ tris, int n_tri, V3 o, V3 d,
                           float tmax) {
  Hit h{tmax, -1, 0.f, 0.f};
  for (int i = 0; i < n_tri; ++i) {
    const V3 p0 = ld3(tris[i].v[0]);
    const V3 e1 = ld3(tris[i].v[1]) - p0;
    const V3 e2 = ld3(tris[i].v[2]) - p0;
    const V3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (fabsf(det) < 1e-12f) continue;
    const float inv = 1.0f / det;
    const V3 tv = o - p0;
    const float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) continue;
    const V3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) continue;
    const float t = dot(e2, qv) * inv;
    if (t > kRayEps && t < h.t) {
      h.t = t;
      h.tri = i;
      h.u = u;
      h.v = v;
    }
  }
  return h;
}

__device__ __forceinline__ V3 shading_normal(const pmx_triangle& tr, float u, float v) {
  return ld3(tr.n[0]) * (1.0f - u - v) + ld3(tr.n[1]) * u + ld3(tr.n[2]) * v;
}

__global__ __launch_bounds__(kThreads) void fg_queries_kernel(
    const pmx_triangle* __restrict__ tris, int n_tri, pmx_camera cam, int width, int height,
    int spp, int depth, uint64_t seed, pm_query* __restrict__ out,
    int32_t* __restrict__ fg_mat) {
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
    const Hit h0 = closest_hit(tris, n_tri, sensor, d0, INFINITY);
    if (h0.tri < 0 || tris[h0.tri].is_light) continue;
    const V3 p0 = sensor + d0 * h0.t;
    const V3 n0 = shading_normal(tris[h0.tri], h0.u, h0.v);
    if (depth == 0) {
      qp = p0; qn = n0; qwo = d0 * -1.0f; qmat = tris[h0.tri].material;
      break;
    }
    V3 t, b;
    basis(n0, t, b);
    bool found = false;
    for (int bounce = 0; bounce < 16 && !found; ++bounce) {
      const V3 d1 = cosine_dir(rnd(seed, id, dim), rnd(seed, id, dim + 1), t, n0, b);
      dim += 2;
      const Hit h1 = closest_hit(tris, n_tri, p0, d1, INFINITY);
      if (h1.tri < 0) continue;
      qp = p0 + d1 * h1.t;
      qn = shading_normal(tris[h1.tri], h1.u, h1.v);
      qwo = d1 * -1.0f;
      qmat = tris[h1.tri].material;
      pmat = tris[h0.tri].material;
      found = true;
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

// ---- photon tracing (include/integrator.h:393-469) -----------------------------

struct TraceScene {
  const pmx_triangle* tris;
  int n_tri;
  const pm_material* mats;
  const float* ior;  // per material (dielectric), may be null (1.5)
  int n_mat;
  const int32_t* lights;  // light triangle indices
  int n_lights;
  float le[3];
};

__device__ __forceinline__ V3 geometric_normal(const pmx_triangle& tr) {
  const V3 p0 = ld3(tr.v[0]);
  return normalize(cross(ld3(tr.v[1]) - p0, ld3(tr.v[2]) - p0));
}

// Cosine term of a photon bounce (include/integrator.h:25-46, FROM_LIGHT):
// zero when the shading and geometric normals disagree about a side.
__device__ __forceinline__ float cos_term_light(V3 wo, V3 wi, V3 ns, V3 ng) {
  const float wi_ns = dot(wi, ns), wi_ng = dot(wi, ng);
  const float wo_ns = dot(wo, ns), wo_ng = dot(wo, ng);
  if (wi_ng * wi_ns <= 0.0f || wo_ng * wo_ns <= 0.0f) return 0.0f;
  return fabsf(wo_ns) * fabsf(wi_ng) / fabsf(wo_ng);
}

// Schlick Fresnel reflectance (include/material.h:43-48).
__device__ __forceinline__ float schlick(float cos_i, float eta_i, float eta_t) {
  const float f0 = (eta_i - eta_t) * (eta_i - eta_t) / ((eta_i + eta_t) * (eta_i + eta_t));
  const float c = fmaxf(1.0f - fabsf(cos_i), 0.0f);
  return f0 + (1.0f - f0) * c * c * c * c * c;
}

// One photon path per thread.  COUNT: photons stored per path into counts;
// otherwise the photons are written at offsets (original photon order =
// (path, bounce) order), those at or past n_store dropped.
template <bool COUNT>
__global__ __launch_bounds__(kThreads) void trace_kernel(
    TraceScene S, int64_t path0, int64_t n_paths, uint64_t seed, int max_depth,
    uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets, int64_t n_store,
    float* __restrict__ pos, float* __restrict__ wi, float* __restrict__ flux) {
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
    const Hit h = closest_hit(S.tris, S.n_tri, o, d, INFINITY);
    if (h.tri < 0) break;
    const pmx_triangle& tri = S.tris[h.tri];
    const V3 x = o + d * h.t;
    const int m = (tri.material >= 0 && tri.material < S.n_mat) ? tri.material : 0;
    const int type = S.mats[m].type;
    const V3 wo = d * -1.0f;
    if (type == PM_MATERIAL_LAMBERT) {
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
    // Russian roulette from the second bounce
    if (k > 0) {
      const float p_rr = fminf(fmaxf(tr_r, fmaxf(tr_g, tr_b)), 1.0f);
      if (rnd(seed, path, dim++) >= p_rr) break;
      tr_r /= p_rr; tr_g /= p_rr; tr_b /= p_rr;
    }
    // next direction in the shading frame (include/material.h:73-245)
    const V3 ns = normalize(shading_normal(tri, h.u, h.v));
    const V3 ng = geometric_normal(tri);
    V3 ft, fb;
    basis(ns, ft, fb);
    const V3 wol = v3(dot(wo, ft), dot(wo, ns), dot(wo, fb));
    V3 wil;
    float scale = 0.0f;  // f / pdf (the albedo is applied below)
    if (type == PM_MATERIAL_LAMBERT) {
      const V3 cd = cosine_dir(rnd(seed, path, dim), rnd(seed, path, dim + 1), v3(1.f, 0.f, 0.f),
                               v3(0.f, 1.f, 0.f), v3(0.f, 0.f, 1.f));
      dim += 2;
      wil = cd;
      scale = (wol.y < 0.0f || wil.y < 0.0f || wil.y <= 0.0f) ? 0.0f : 1.0f / wil.y;
    } else if (type == PM_MATERIAL_MIRROR) {
      wil = v3(-wol.x, wol.y, -wol.z);
      scale = 1.0f / fabsf(wil.y);
    } else {
      const float eta = S.ior ? S.ior[m] : 1.5f;
      const bool outside = wol.y > 0.0f;
      const float eta_o = outside ? 1.0f : eta, eta_i = outside ? eta : 1.0f;
      const V3 n = outside ? v3(0.f, 1.f, 0.f) : v3(0.f, -1.f, 0.f);
      const float cos_o = dot(wol, n);
      const float fr = schlick(cos_o, eta_o, eta_i);
      const V3 refl = wol * -1.0f + n * (2.0f * cos_o);
      wil = refl;
      if (rnd(seed, path, dim++) >= fr) {
        const V3 th = (wol - n * cos_o) * (-eta_o / eta_i);
        const float th2 = dot(th, th);
        if (th2 <= 1.0f) wil = th + n * (-sqrtf(fmaxf(1.0f - th2, 0.0f)));
      }
      scale = 1.0f / fabsf(wil.y);
    }
    const V3 wiw = ft * wil.x + ns * wil.y + fb * wil.z;
    // Lambert: f = rho / pi, pdf = cos / pi; specular: f = rho / |cos|, pdf = 1
    const float c = cos_term_light(wo, wiw, ns, ng) * scale;
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
}  // namespace pm

using namespace pm;

extern "C" {

int pmx_trace_photons(pm_context* h, const pmx_triangle* tris, int n_tri,
                      const pm_material* mats, const float* mat_ior, int n_mat,
                      const float* light_le, int64_t n_store, int max_depth, uint64_t seed,
                      float* pos_soa, float* wi_soa, float* flux_soa, int64_t* n_emit) {
  if (!h || !tris || !mats || !light_le || !pos_soa || !wi_soa || !flux_soa || !n_emit)
    return fail(PM_ERR_INVALID_ARG, "null argument");
  if (n_tri < 1 || n_tri > kMaxSceneTris || n_mat < 1 || n_store < 1 || max_depth < 1)
    return fail(PM_ERR_INVALID_ARG, "bad sizes");
  auto* ctx = reinterpret_cast<Context*>(h);
  PM_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> lights;
  for (int i = 0; i < n_tri; ++i)
    if (tris[i].is_light) lights.push_back(i);
  if (lights.empty()) return fail(PM_ERR_INVALID_ARG, "the scene has no light triangle");
  DevBuf dt, dm, di, dl, dcount, doff;
  if (int rc = dt.reserve(n_tri * sizeof(pmx_triangle))) return rc;
  if (int rc = dm.reserve(n_mat * sizeof(pm_material))) return rc;
  if (int rc = dl.reserve(lights.size() * sizeof(int32_t))) return rc;
  PM_HIP(hipMemcpy(dt.ptr, tris, n_tri * sizeof(pmx_triangle), hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(dm.ptr, mats, n_mat * sizeof(pm_material), hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(dl.ptr, lights.data(), lights.size() * sizeof(int32_t),
                   hipMemcpyHostToDevice));
  if (mat_ior) {
    if (int rc = di.reserve(n_mat * sizeof(float))) return rc;
    PM_HIP(hipMemcpy(di.ptr, mat_ior, n_mat * sizeof(float), hipMemcpyHostToDevice));
  }
  TraceScene S{dt.as<pmx_triangle>(), n_tri, dm.as<pm_material>(),
               mat_ior ? di.as<float>() : nullptr, n_mat, dl.as<int32_t>(),
               static_cast<int>(lights.size()), {light_le[0], light_le[1], light_le[2]}};
  // pass 1: photons stored per path, chunk by chunk until n_store are
  // reached; after a pilot chunk, each chunk is sized from the photons per
  // path seen so far
  constexpr int64_t kMinChunk = int64_t{1} << 16;
  constexpr int64_t kMaxChunk = int64_t{1} << 22;
  if (int rc = dcount.reserve(kMaxChunk * sizeof(uint32_t))) return rc;
  std::vector<int64_t> offs;  // exclusive prefix per path
  int64_t acc = 0, paths = 0;
  int64_t chunk = kMinChunk;
  std::vector<uint32_t> cnt(kMaxChunk);
  while (acc < n_store) {
    if (paths > (int64_t{1} << 40)) return fail(PM_ERR_INVALID_ARG, "no photon is ever stored");
    hipLaunchKernelGGL(trace_kernel<true>, dim3((chunk + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, nullptr, S, paths, chunk, seed, max_depth,
                       dcount.as<uint32_t>(), nullptr, int64_t{0}, nullptr, nullptr, nullptr);
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpy(cnt.data(), dcount.ptr, chunk * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < chunk && acc < n_store; ++j) {
      offs.push_back(acc);
      acc += cnt[j];
      ++paths;
    }
    if (acc < n_store) {
      const double per_path = std::max(static_cast<double>(acc) / static_cast<double>(paths), 1e-3);
      const double want = 1.1 * static_cast<double>(n_store - acc) / per_path + 1024.0;
      chunk = std::min<int64_t>(kMaxChunk, std::max<int64_t>(kMinChunk, static_cast<int64_t>(want)));
    }
  }
  // the path holding photon n_store - 1 is the last one emitted
  *n_emit = paths;
  // pass 2: the same paths again, writing their photons in (path, bounce) order
  if (int rc = doff.reserve(offs.size() * sizeof(int64_t))) return rc;
  PM_HIP(hipMemcpy(doff.ptr, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(trace_kernel<false>, dim3((paths + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, nullptr, S, int64_t{0}, paths, seed, max_depth, nullptr,
                     doff.as<int64_t>(), n_store, pos_soa, wi_soa, flux_soa);
  PM_HIP(hipGetLastError());
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

int pmx_fg_queries(pm_context* h, const pmx_triangle* tris, int n_tri, const pmx_camera* cam,
                   int width, int height, int spp, int depth, uint64_t seed, pm_query* out,
                   int32_t* fg_mat_opt) {
  if (!h || !tris || !cam || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (n_tri < 1 || n_tri > kMaxSceneTris)
    return fail(PM_ERR_INVALID_ARG, "triangle count must be in 1..4096");
  if (width < 1 || height < 1 || spp < 1 || (depth != 0 && depth != 1))
    return fail(PM_ERR_INVALID_ARG, "bad image size, spp or depth");
  auto* ctx = reinterpret_cast<Context*>(h);
  PM_HIP(hipSetDevice(ctx->device));
  DevBuf dt;
  if (int rc = dt.reserve(n_tri * sizeof(pmx_triangle))) return rc;
  PM_HIP(hipMemcpy(dt.ptr, tris, n_tri * sizeof(pmx_triangle), hipMemcpyHostToDevice));
  const int64_t total = static_cast<int64_t>(width) * height * spp;
  hipLaunchKernelGGL(fg_queries_kernel, dim3((total + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, nullptr, dt.as<pmx_triangle>(), n_tri, *cam, width,
                     height, spp, depth, seed, out, fg_mat_opt);
  PM_HIP(hipGetLastError());
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

int pmx_surrogate_photons(pm_context* h, const pmx_triangle* tris, int n_tri,
                          const float* mat_albedo, int n_mat, const float* power, int64_t n,
                          uint64_t seed, float* pos_soa, float* wi_soa, float* flux_soa) {
  if (!h || !tris || !mat_albedo || !power || !pos_soa || !wi_soa || !flux_soa)
    return fail(PM_ERR_INVALID_ARG, "null argument");
  if (n_tri < 1 || n_tri > kMaxSceneTris || n_mat < 1 || n < 1)
    return fail(PM_ERR_INVALID_ARG, "bad sizes");
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
  hipLaunchKernelGGL(surrogate_photons_kernel, dim3((n + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, nullptr, dt.as<pmx_triangle>(), dc.as<float>(), n_tri,
                     da.as<float>(), n_mat, power[0], power[1], power[2], n, seed, pos_soa,
                     wi_soa, flux_soa);
  PM_HIP(hipGetLastError());
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

}  // extern "C"
