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
#include <deque>
#include <vector>

#include "pm_scene_dev.hpp"

namespace pm {
namespace {

using namespace scn;

// ---- kernels --------------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void intersect_kernel(
    SceneDev S, int64_t n, const float* __restrict__ org, const float* __restrict__ dir,
    int32_t* __restrict__ out_tri, float* __restrict__ out_t, float* __restrict__ out_u,
    float* __restrict__ out_v) {
  extern __shared__ int stk[];  // SceneImpl::stack_bytes
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
  extern __shared__ int stk_all[];  // SceneImpl::stack_bytes
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
  // a sample that finds no query point in 64 draws keeps this placeholder:
  // material -1, whose estimate is zero
  int qmat = -1, pmat = -1;
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

// A photon in the count pass's cache: position, wi, flux.
constexpr int kCacheFloats = 9;
// an overflow record: the photon's 9 floats, its place in its path, and the
// path (two words)
constexpr int kOvfFloats = 12;

// One photon path per lane.  COUNT: photons stored per path into counts, and
// (with a cache) the path's first cache_cap photons into its cache slots;
// otherwise the photons are written at offsets (original photon order =
// (path, bounce) order), those at or past n_store dropped.  The paths are
// path0 + i, or path_list[i] when a list is given.  CAUSTICS: the path ends at
// its first Lambert hit, stored only after specular hits.
template <bool COUNT, bool CAUSTICS>
__global__ __launch_bounds__(kThreads) void trace_kernel(
    SceneDev S, int64_t path0, int64_t n_paths, uint64_t seed, int max_depth,
    uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets, int64_t n_store,
    float* __restrict__ pos, float* __restrict__ wi, float* __restrict__ flux,
    float* __restrict__ cache, int cache_cap, const int64_t* __restrict__ path_list,
    float* __restrict__ ovf, uint32_t* __restrict__ ovf_n, uint32_t ovf_cap) {
  extern __shared__ int stk_all[];  // SceneImpl::stack_bytes
  int* stk = stk_all + threadIdx.x;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n_paths) return;
  const uint64_t path = static_cast<uint64_t>(path_list ? path_list[i] : path0 + i);
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
  const V3 le = emission(S, lt);
  float tr_r = le.x * w0, tr_g = le.y * w0, tr_b = le.z * w0;
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
        } else if (cache && stored < static_cast<uint32_t>(cache_cap)) {
          float* c = cache + (i * cache_cap + stored) * kCacheFloats;
          c[0] = x.x; c[1] = x.y; c[2] = x.z;
          c[3] = wo.x; c[4] = wo.y; c[5] = wo.z;
          c[6] = tr_r; c[7] = tr_g; c[8] = tr_b;
        } else if (ovf) {
          // past the cache: a record of the photon, its path and its place
          // in the path, appended to the overflow area
          const uint32_t r = atomicAdd(ovf_n, 1u);
          if (r < ovf_cap) {
            float* c = ovf + static_cast<int64_t>(r) * kOvfFloats;
            c[0] = x.x; c[1] = x.y; c[2] = x.z;
            c[3] = wo.x; c[4] = wo.y; c[5] = wo.z;
            c[6] = tr_r; c[7] = tr_g; c[8] = tr_b;
            c[9] = __uint_as_float(stored);
            c[10] = __uint_as_float(static_cast<uint32_t>(path));
            c[11] = __uint_as_float(static_cast<uint32_t>(path >> 32));
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

// The count pass's cached photons to their places: thread t takes slot
// t % cap of path t / cap of a chunk (paths [0, n) of it, their counts and
// offsets), and writes it at offsets[path] + slot when the path stored that
// many and the place is below n_store.
__global__ __launch_bounds__(kThreads) void cache_scatter_kernel(
    const float* __restrict__ cache, int cap, int64_t n, const uint32_t* __restrict__ counts,
    const int64_t* __restrict__ offsets, int64_t n_store, float* __restrict__ pos,
    float* __restrict__ wi, float* __restrict__ flux) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t pl = t / cap;
  const int j = static_cast<int>(t - pl * cap);
  if (pl >= n) return;
  if (static_cast<uint32_t>(j) >= counts[pl]) return;
  const int64_t at = offsets[pl] + j;
  if (at >= n_store) return;
  const float* c = cache + t * kCacheFloats;
  pos[at] = c[0]; pos[n_store + at] = c[1]; pos[2 * n_store + at] = c[2];
  wi[at] = c[3]; wi[n_store + at] = c[4]; wi[2 * n_store + at] = c[5];
  flux[at] = c[6]; flux[n_store + at] = c[7]; flux[2 * n_store + at] = c[8];
}

// The overflow records to their places: record r holds photon `place` of
// its path; it goes to offsets[path] + place when the path is one of the
// n_paths emitted and the place is below n_store.
__global__ __launch_bounds__(kThreads) void ovf_scatter_kernel(
    const float* __restrict__ ovf, int64_t n_rec, const int64_t* __restrict__ offsets,
    int64_t n_paths, int64_t n_store, float* __restrict__ pos, float* __restrict__ wi,
    float* __restrict__ flux) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= n_rec) return;
  const float* c = ovf + r * kOvfFloats;
  const uint64_t path = static_cast<uint64_t>(__float_as_uint(c[10])) |
                        (static_cast<uint64_t>(__float_as_uint(c[11])) << 32);
  if (path >= static_cast<uint64_t>(n_paths)) return;
  const int64_t at = offsets[path] + static_cast<int64_t>(__float_as_uint(c[9]));
  if (at >= n_store) return;
  pos[at] = c[0]; pos[n_store + at] = c[1]; pos[2 * n_store + at] = c[2];
  wi[at] = c[3]; wi[n_store + at] = c[4]; wi[2 * n_store + at] = c[5];
  flux[at] = c[6]; flux[n_store + at] = c[7]; flux[2 * n_store + at] = c[8];
}

}  // namespace

// ---- host side --------------------------------------------------------------------

namespace {

using namespace scn;

uint64_t kind_seed(uint64_t seed, int kind) {
  return kind == PMX_PHOTONS_CAUSTICS ? mix64(seed ^ 0xc0057105ca05715ull) : seed;
}

int launch_trace_count(const SceneImpl* s, int kind, int64_t path0, int64_t n, uint64_t seed,
                       int max_depth, uint32_t* counts, float* cache, int cap, float* ovf,
                       uint32_t* ovf_n, uint32_t ovf_cap) {
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads)), block(kThreads);
  const SceneDev S = s->dev(0);
  if (kind == PMX_PHOTONS_CAUSTICS)
    hipLaunchKernelGGL((trace_kernel<true, true>), grid, block, s->stack_bytes(0), nullptr, S, path0, n, seed,
                       max_depth, counts, nullptr, int64_t{0}, nullptr, nullptr, nullptr, cache,
                       cap, nullptr, ovf, ovf_n, ovf_cap);
  else
    hipLaunchKernelGGL((trace_kernel<true, false>), grid, block, s->stack_bytes(0), nullptr, S, path0, n, seed,
                       max_depth, counts, nullptr, int64_t{0}, nullptr, nullptr, nullptr, cache,
                       cap, nullptr, ovf, ovf_n, ovf_cap);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

// The write pass over paths [0, n_paths), or over the n_paths paths of
// path_list (offsets then holds their offsets, in list order).
int launch_trace_write(const SceneImpl* s, int kind, int64_t n_paths, uint64_t seed,
                       int max_depth, const int64_t* offsets, int64_t n_store, float* pos,
                       float* wi, float* flux, const int64_t* path_list = nullptr) {
  const dim3 grid(static_cast<unsigned>((n_paths + kThreads - 1) / kThreads)), block(kThreads);
  const SceneDev S = s->dev(0);
  if (kind == PMX_PHOTONS_CAUSTICS)
    hipLaunchKernelGGL((trace_kernel<false, true>), grid, block, s->stack_bytes(0), nullptr, S, int64_t{0},
                       n_paths, seed, max_depth, nullptr, offsets, n_store, pos, wi, flux, nullptr,
                       0, path_list, nullptr, nullptr, 0u);
  else
    hipLaunchKernelGGL((trace_kernel<false, false>), grid, block, s->stack_bytes(0), nullptr, S, int64_t{0},
                       n_paths, seed, max_depth, nullptr, offsets, n_store, pos, wi, flux, nullptr,
                       0, path_list, nullptr, nullptr, 0u);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

constexpr int64_t kMinChunk = int64_t{1} << 16;
constexpr int64_t kMaxChunk = int64_t{1} << 22;
// Photons a path keeps in the count pass's cache: a caustic path stores at
// most one; a global path in the box scenes stores about four on average.
int cache_cap_of(int kind) { return kind == PMX_PHOTONS_CAUSTICS ? 1 : 16; }
// Most bytes of cached photons one tracing call keeps before it falls back to
// tracing every path twice.
constexpr int64_t kCacheBytesMax = int64_t{2} << 30;

// A chunk of the count pass whose counts and first photons stay on the
// device, in the scene's trace_counts and trace_cache from these offsets.
struct CountChunk {
  int64_t path0 = 0;      // first path of the chunk
  int64_t n = 0;          // paths of the chunk that are emitted
  int64_t count_at = 0;   // uint32 words into trace_counts
  int64_t cache_at = 0;   // floats into trace_cache
};

// Grows buf to at least `need` bytes (by half again at least), keeping its
// first `keep` bytes.
int grow_keep(DevBuf& buf, size_t need, size_t keep) {
  if (need <= buf.bytes) return PM_OK;
  DevBuf nb;
  if (int rc = nb.reserve(std::max(need, buf.bytes + buf.bytes / 2))) return rc;
  if (keep > 0) PM_HIP(hipMemcpy(nb.ptr, buf.ptr, keep, hipMemcpyDeviceToDevice));
  std::swap(buf.ptr, nb.ptr);
  std::swap(buf.bytes, nb.bytes);
  return PM_OK;  // nb frees the old storage
}

// Count pass over paths [0, ...): appends per-path exclusive offsets until
// either `stop_store` photons are stored (stop_store >= 0) or `stop_paths`
// paths are counted (stop_paths >= 0).  Returns the photons stored in *acc.
// With chunks, the pass also keeps every chunk's counts and each path's first
// cache_cap_of(kind) photons on the device, and lists in *overflow the paths
// that stored more; when that cache would pass kCacheBytesMax it is dropped
// (chunks left empty, *cached false) and the caller traces every path again.
int count_paths(const SceneImpl* s, int kind, uint64_t seed, int max_depth, int64_t stop_store,
                int64_t stop_paths, std::vector<int64_t>& offs, int64_t* acc_out,
                std::deque<CountChunk>* chunks = nullptr,
                std::vector<int64_t>* overflow = nullptr, bool* cached = nullptr,
                int64_t* ovf_recs = nullptr) {
  DevBuf dcount;              // counts of a chunk that is not cached
  std::vector<uint32_t> cnt;  // grown to the largest chunk
  const int cap = cache_cap_of(kind);
  bool caching = chunks != nullptr && overflow != nullptr;
  // the overflow area: a counter word, then the records (a caustic path
  // never stores past its cache of one)
  if (ovf_recs) *ovf_recs = -1;
  float* ovf = nullptr;
  uint32_t* ovf_n = nullptr;
  // (the context's PM_OPT_TRACE_RECORDS records at most; when more are made
  // the paths that overflowed their cache are traced again instead)
  const uint32_t ovf_cap = static_cast<uint32_t>(s->ctx->trace_records);
  if (caching && ovf_recs && kind != PMX_PHOTONS_CAUSTICS && ovf_cap > 0) {
    if (int rc = s->trace_ovf.reserve(256 + static_cast<size_t>(ovf_cap) * kOvfFloats *
                                                sizeof(float)))
      return rc;
    ovf_n = s->trace_ovf.as<uint32_t>();
    ovf = reinterpret_cast<float*>(static_cast<char*>(s->trace_ovf.ptr) + 256);
    if (int rc = launch_fill_u32(ovf_n, 0u, 1, nullptr)) return rc;
  }
  int64_t counts_used = 0, cache_used = 0;  // words / floats of the scene's buffers in use
  int64_t acc = 0, paths = 0;
  int64_t chunk = kMinChunk;
  for (;;) {
    if (stop_store >= 0 && acc >= stop_store) break;
    if (stop_paths >= 0) {
      if (paths >= stop_paths) break;
      chunk = std::min(kMaxChunk, stop_paths - paths);
    }
    // a scene that stores (almost) nothing would loop for ever
    if (stop_store >= 0 && paths >= (int64_t{1} << 26) &&
        static_cast<double>(acc) * 1e4 < static_cast<double>(stop_store))
      return fail(PM_ERR_INVALID_ARG, "photon tracing stores too few photons to reach the count");
    uint32_t* dc = nullptr;
    float* dcache = nullptr;
    CountChunk* cc = nullptr;
    if (caching) {
      const int64_t cf = chunk * cap * kCacheFloats;  // floats of this chunk's cache
      if ((cache_used + cf) * static_cast<int64_t>(sizeof(float)) > kCacheBytesMax) {
        caching = false;
        chunks->clear();
        overflow->clear();
      } else {
        if (int rc = grow_keep(s->trace_counts, (counts_used + chunk) * sizeof(uint32_t),
                               counts_used * sizeof(uint32_t)))
          return rc;
        if (int rc = grow_keep(s->trace_cache, (cache_used + cf) * sizeof(float),
                               cache_used * sizeof(float)))
          return rc;
        cc = &chunks->emplace_back();
        cc->count_at = counts_used;
        cc->cache_at = cache_used;
        dc = s->trace_counts.as<uint32_t>() + counts_used;
        dcache = s->trace_cache.as<float>() + cache_used;
        counts_used += chunk;
        cache_used += cf;
      }
    }
    if (!dc) {
      if (int rc = dcount.reserve(chunk * sizeof(uint32_t))) return rc;
      dc = dcount.as<uint32_t>();
    }
    if (int rc = launch_trace_count(s, kind, paths, chunk, seed, max_depth, dc, dcache, cap,
                                    dcache ? ovf : nullptr, ovf_n, ovf_cap))
      return rc;
    if (cnt.size() < static_cast<size_t>(chunk)) cnt.resize(static_cast<size_t>(chunk));
    PM_HIP(hipMemcpy(cnt.data(), dc, chunk * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (cc) cc->path0 = paths;
    for (int64_t j = 0; j < chunk; ++j) {
      if (stop_store >= 0 && acc >= stop_store) break;
      offs.push_back(acc);
      acc += cnt[j];
      if (cc && cnt[j] > static_cast<uint32_t>(cap)) overflow->push_back(paths);
      ++paths;
    }
    if (cc) cc->n = paths - cc->path0;
    if (stop_store >= 0 && acc < stop_store) {
      // size the next chunk from the photons per path seen so far
      const double per_path = std::max(static_cast<double>(acc) / static_cast<double>(paths), 1e-3);
      const double want = 1.1 * static_cast<double>(stop_store - acc) / per_path + 1024.0;
      chunk = std::min<int64_t>(kMaxChunk, std::max<int64_t>(kMinChunk, static_cast<int64_t>(want)));
    }
  }
  *acc_out = acc;
  if (cached) *cached = caching;
  if (caching && ovf_n) {
    // every photon past a cache has its record: no path is traced again
    uint32_t n_rec = 0;
    PM_HIP(hipMemcpy(&n_rec, ovf_n, sizeof(n_rec), hipMemcpyDeviceToHost));
    if (n_rec <= ovf_cap) {
      *ovf_recs = n_rec;
      overflow->clear();
    }
  }
  return PM_OK;
}

// The write pass after a count pass: the cached photons of every chunk to
// their places, then the count pass's overflow records (ovf_recs >= 0), or
// the paths that stored more than the cache holds traced again (all their
// photons written, the cached ones to the same values); or, without a
// cache, every path traced again.
int write_paths(const SceneImpl* s, int kind, uint64_t seed, int max_depth,
                const std::vector<int64_t>& offs, const int64_t* doff, bool cached,
                const std::deque<CountChunk>& chunks, const std::vector<int64_t>& overflow,
                int64_t n_store, float* pos, float* wi, float* flux, int64_t ovf_recs) {
  const int64_t paths = static_cast<int64_t>(offs.size());
  if (!cached)
    return launch_trace_write(s, kind, paths, seed, max_depth, doff, n_store, pos, wi, flux);
  const int cap = cache_cap_of(kind);
  for (const CountChunk& cc : chunks) {
    if (cc.n <= 0) continue;
    const int64_t threads = cc.n * cap;
    hipLaunchKernelGGL(cache_scatter_kernel, dim3(static_cast<unsigned>((threads + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, nullptr, s->trace_cache.as<float>() + cc.cache_at, cap,
                       cc.n, s->trace_counts.as<uint32_t>() + cc.count_at, doff + cc.path0,
                       n_store, pos, wi, flux);
    PM_HIP(hipGetLastError());
  }
  if (ovf_recs > 0) {
    hipLaunchKernelGGL(ovf_scatter_kernel,
                       dim3(static_cast<unsigned>((ovf_recs + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, nullptr,
                       reinterpret_cast<const float*>(static_cast<const char*>(s->trace_ovf.ptr) +
                                                      256),
                       ovf_recs, doff, paths, n_store, pos, wi, flux);
    PM_HIP(hipGetLastError());
  }
  if (!overflow.empty()) {
    const size_t no = overflow.size();
    std::vector<int64_t> ooff(no);
    for (size_t i = 0; i < no; ++i) ooff[i] = offs[static_cast<size_t>(overflow[i])];
    DevBuf dlist;
    if (int rc = dlist.reserve(2 * no * sizeof(int64_t))) return rc;
    int64_t* dl = dlist.as<int64_t>();
    PM_HIP(hipMemcpy(dl, overflow.data(), no * sizeof(int64_t), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(dl + no, ooff.data(), no * sizeof(int64_t), hipMemcpyHostToDevice));
    if (int rc = launch_trace_write(s, kind, static_cast<int64_t>(no), seed, max_depth, dl + no,
                                    n_store, pos, wi, flux, dl))
      return rc;
    PM_HIP(hipDeviceSynchronize());  // dlist is freed on return
  }
  return PM_OK;
}

bool valid_kind(int kind) { return kind == PMX_PHOTONS_GLOBAL || kind == PMX_PHOTONS_CAUSTICS; }

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

int pmx_scene_create(pm_context* h, const pmx_triangle* tris, int n_tri, const pm_material* mats,
                     const float* mat_ior, int n_mat, const float* light_le,
                     const float* mat_le, pmx_scene** out_scene) {
  if (!h || !tris || !mats || !out_scene) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (!light_le && !mat_le) return fail(PM_ERR_INVALID_ARG, "no emission given (light_le or mat_le)");
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
  // emission table: slot 0 = light_le, slot 1 + m = material m's emission;
  // the device copy of each triangle carries its slot + 1 in is_light
  std::vector<float> le_table(3 * (1 + static_cast<size_t>(n_mat)), 0.0f);
  if (light_le)
    for (int a = 0; a < 3; ++a) le_table[a] = light_le[a];
  std::vector<pmx_triangle> dev_tris(tris, tris + n_tri);
  std::vector<int32_t> lights;
  double light_area = 0.0;
  for (int i = 0; i < n_tri; ++i) {
    int slot = -1;
    const int m = (tris[i].material >= 0 && tris[i].material < n_mat) ? tris[i].material : 0;
    if (mat_le && (mat_le[3 * m] > 0.0f || mat_le[3 * m + 1] > 0.0f || mat_le[3 * m + 2] > 0.0f)) {
      slot = 1 + m;
      for (int a = 0; a < 3; ++a) le_table[3 * slot + a] = mat_le[3 * m + a];
    } else if (light_le && tris[i].is_light) {
      slot = 0;
    }
    dev_tris[i].is_light = slot + 1;
    if (slot < 0) continue;
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
  s->has_ior = mat_ior != nullptr;
  const auto upload = [](DevBuf& buf, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PM_OK;
    if (int rc = buf.reserve(bytes)) return rc;
    PM_HIP(hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice));
    return PM_OK;
  };
  int rc = upload(s->tris, dev_tris.data(), n_tri * sizeof(pmx_triangle));
  if (!rc) rc = upload(s->le, le_table.data(), le_table.size() * sizeof(float));
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
                        float* out_u, float* out_v, int mem, void* stream) {
  if (!scene) return fail(PM_ERR_INVALID_ARG, "null scene");
  if (n < 0) return fail(PM_ERR_INVALID_ARG, "negative ray count");
  if (method != PMX_HIT_BVH && method != PMX_HIT_BRUTE_FORCE)
    return fail(PM_ERR_INVALID_ARG, "unknown hit method");
  if (mem != PM_MEM_DEVICE && mem != PM_MEM_HOST) return fail(PM_ERR_INVALID_ARG, "bad mem");
  if (n == 0) return PM_OK;
  if (!org_soa || !dir_soa || !out_tri || !out_t) return fail(PM_ERR_INVALID_ARG, "null array");
  const auto* s = reinterpret_cast<const SceneImpl*>(scene);
  PM_HIP(hipSetDevice(s->ctx->device));
  const auto hs = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads));
  if (mem == PM_MEM_DEVICE) {
    hipLaunchKernelGGL(intersect_kernel, grid, dim3(kThreads), s->stack_bytes(method), hs, s->dev(method), n,
                       org_soa, dir_soa, out_tri, out_t, out_u, out_v);
    PM_HIP(hipGetLastError());
    return PM_OK;
  }
  // host arrays: stage through device memory and synchronise
  const size_t v3 = 3 * static_cast<size_t>(n) * sizeof(float);
  const size_t v1 = static_cast<size_t>(n) * sizeof(float);
  DevBuf d_org, d_dir, d_tri, d_t, d_u, d_v;
  int rc = d_org.reserve(v3);
  if (!rc) rc = d_dir.reserve(v3);
  if (!rc) rc = d_tri.reserve(v1);
  if (!rc) rc = d_t.reserve(v1);
  if (!rc && out_u) rc = d_u.reserve(v1);
  if (!rc && out_v) rc = d_v.reserve(v1);
  if (rc) return rc;
  PM_HIP(hipMemcpyAsync(d_org.ptr, org_soa, v3, hipMemcpyHostToDevice, hs));
  PM_HIP(hipMemcpyAsync(d_dir.ptr, dir_soa, v3, hipMemcpyHostToDevice, hs));
  hipLaunchKernelGGL(intersect_kernel, grid, dim3(kThreads), s->stack_bytes(method), hs, s->dev(method), n,
                     d_org.as<float>(), d_dir.as<float>(), d_tri.as<int32_t>(), d_t.as<float>(),
                     out_u ? d_u.as<float>() : nullptr, out_v ? d_v.as<float>() : nullptr);
  PM_HIP(hipGetLastError());
  PM_HIP(hipMemcpyAsync(out_tri, d_tri.ptr, v1, hipMemcpyDeviceToHost, hs));
  PM_HIP(hipMemcpyAsync(out_t, d_t.ptr, v1, hipMemcpyDeviceToHost, hs));
  if (out_u) PM_HIP(hipMemcpyAsync(out_u, d_u.ptr, v1, hipMemcpyDeviceToHost, hs));
  if (out_v) PM_HIP(hipMemcpyAsync(out_v, d_v.ptr, v1, hipMemcpyDeviceToHost, hs));
  PM_HIP(hipStreamSynchronize(hs));
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
                     dim3(kThreads), s->stack_bytes(0), nullptr, s->dev(0), *cam, width, height, spp, depth, seed,
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
  // pass 1: photons stored per path until n_store are reached (the first
  // photons of each path kept in a cache)
  std::vector<int64_t> offs;
  std::deque<CountChunk> chunks;
  std::vector<int64_t> over;
  bool cached = false;
  int64_t acc = 0;
  int64_t ovf_recs = -1;
  if (int rc = count_paths(s, kind, kseed, max_depth, n_store, -1, offs, &acc, &chunks, &over,
                           &cached, &ovf_recs))
    return rc;
  // the path holding photon n_store - 1 is the last one emitted
  const int64_t paths = static_cast<int64_t>(offs.size());
  *n_emit = paths;
  // pass 2: the photons in (path, bounce) order
  DevBuf doff;
  if (int rc = doff.reserve(offs.size() * sizeof(int64_t))) return rc;
  PM_HIP(hipMemcpy(doff.ptr, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  if (int rc = write_paths(s, kind, kseed, max_depth, offs, doff.as<int64_t>(), cached, chunks,
                           over, n_store, pos_soa, wi_soa, flux_soa, ovf_recs))
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
  std::deque<CountChunk> chunks;
  std::vector<int64_t> over;
  bool cached = false;
  int64_t total = 0;
  int64_t ovf_recs = -1;
  if (n_paths > 0)
    if (int rc = count_paths(s, kind, kseed, max_depth, -1, n_paths, offs, &total, &chunks, &over,
                             &cached, &ovf_recs))
      return rc;
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
    if (int rc = write_paths(s, kind, kseed, max_depth, offs, doff.as<int64_t>(), cached, chunks,
                             over, total, dpos.as<float>(), dwi.as<float>(), dflux.as<float>(),
                             ovf_recs))
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
