// CPU restatement of the reference's two integrators and render loop, against
// which the device renderer's images are tested.
//
// TEST INFRASTRUCTURE ONLY: built into libpm_oracle.so and called by tests/
// alone; the product library never links or calls it.
//
// What it restates (paths relative to the reference checkout):
//   * the render loop, sample jitter, NaN / negative drops and the division
//     by spp                                     examples/main.cpp:33-70
//   * Camera::sampleRay                          include/camera.h:41-47
//   * PathTracing::integrate                     include/integrator.h:60-101
//   * PhotonMapping::integrateRecursive, computeDirectIllumination,
//     computeIndirectIlluminationRecursive,
//     computeRadianceWithPhotonMap / Caustics    include/integrator.h:133-377
//   * cosTerm                                    include/integrator.h:25-46
//   * Lambert, Mirror, Glass (sampleDirection, sampleAllDirection, Schlick)
//                                                include/material.h:16-245
//   * Triangle::samplePoint, computeShadingNormal include/triangle.h:64-114
//   * orthonormalBasis, sampleCosineHemisphere   include/core.h:224-247,
//                                                include/sampler.h:84-91
//   * Scene::sampleLight                         include/scene.h:350-355
//   * the gather query points of a camera pass: the primary diffuse hit
//     (include/integrator.h:305-308) or the hit of one cosine-sampled
//     final-gathering ray followed through specular surfaces (:215-252),
//     with the benchmark producer's redraw rules (orc_fg_queries)
// Ray casting is a two-sided Moller-Trumbore test (the reference calls
// Embree; its hit points are unpinned), with hits in (1e-5, tmax), the
// smallest t winning and ties going to the lower triangle index; a
// median-split BVH returns exactly the brute-force loop's hits
// (orc_closest_hits checks one against the other).  The photon-map estimates use the reference kd-tree
// and FP32 estimate of oracle_core.hpp over the photon arrays the caller
// passes.  Random numbers come from this file's own stream per pixel
// (SplitMix64), independent of the device renderer's, so the two images are
// independent Monte Carlo estimates of the same expectation.

#include "oracle_core.hpp"

namespace {

using namespace orc;

constexpr float kRayEps = 1e-5f;  // include/core.h:19

struct V {
  float x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator*(float s, V a) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator/(V a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V mul(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V normalize(V a) { return a / std::sqrt(dot(a, a)); }
inline V ld(const float* p) { return {p[0], p[1], p[2]}; }

// Triangle record, 80 bytes (the library's pmx_triangle).
struct Tri {
  float v[3][3];
  float n[3][3];
  int32_t material;
  int32_t is_light;
};
static_assert(sizeof(Tri) == 80, "triangle layout");

struct Bvh;

struct Scene {
  const Tri* tris;
  int nt;
  const MaterialRec* mats;
  const float* ior;
  int nm;
  std::vector<V> tri_le;   // emission per triangle (zero: not a light)
  std::vector<int> lights;
  const Bvh* bvh = nullptr;  // when set, closest() traverses it (same results)
  bool is_light(int t) const { return tri_le[t].x > 0 || tri_le[t].y > 0 || tri_le[t].z > 0; }
  int mat_of(int t) const {
    const int m = tris[t].material;
    return (m >= 0 && m < nm) ? m : 0;
  }
};

struct Rng {
  uint64_t s;
  float next() {
    s += 0x9e3779b97f4a7c15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return static_cast<float>(z >> 40) * (1.0f / 16777216.0f);
  }
};

struct Hit {
  float t;
  int tri;
  float u, v;
};

// The two-sided Moller-Trumbore test of one triangle: t (or -1 on a miss).
inline float tri_hit(const Tri& tr, V o, V d, float& u_out, float& v_out) {
  const V p0 = ld(tr.v[0]);
  const V e1 = ld(tr.v[1]) - p0, e2 = ld(tr.v[2]) - p0;
  const V pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (std::fabs(det) < 1e-12f) return -1.0f;
  const float inv = 1.0f / det;
  const V tv = o - p0;
  const float u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return -1.0f;
  const V qv = cross(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  u_out = u;
  v_out = v;
  return dot(e2, qv) * inv;
}

// Brute force: every triangle in index order, the smallest t in
// (1e-5, tmax), ties to the lower index.
Hit closest_brute(const Scene& S, V o, V d, float tmax) {
  Hit h{tmax, -1, 0.f, 0.f};
  for (int i = 0; i < S.nt; ++i) {
    float u, v;
    const float t = tri_hit(S.tris[i], o, d, u, v);
    if (t > kRayEps && t < h.t) h = Hit{t, i, u, v};
  }
  return h;
}

// A median-split bounding-volume hierarchy over triangle centroids (leaves
// of at most 4 triangles), boxes padded by 1e-4 of the scene extent so that
// they contain every hit the triangle test reports.  Traversal keeps the
// smallest (t, index) over the triangles of every box the ray meets no later
// than the best t so far, so it returns the brute-force result exactly.
struct Bvh {
  struct Node {
    float lo[3], hi[3];
    int left = -1, right = -1;  // children (inner node)
    int first = 0, count = 0;   // triangle range of idx (leaf when count > 0)
  };
  std::vector<Node> nodes;
  std::vector<int> idx;

  void build(const Tri* tris, int n) {
    idx.resize(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<V> c(n);
    float slo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                    std::numeric_limits<float>::max()};
    float shi[3] = {-slo[0], -slo[1], -slo[2]};
    for (int i = 0; i < n; ++i) {
      V a = ld(tris[i].v[0]), b = ld(tris[i].v[1]), e = ld(tris[i].v[2]);
      c[i] = (a + b + e) / 3.0f;
      for (const V& p : {a, b, e}) {
        const float q[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) {
          slo[k] = std::min(slo[k], q[k]);
          shi[k] = std::max(shi[k], q[k]);
        }
      }
    }
    const float ext = std::max({shi[0] - slo[0], shi[1] - slo[1], shi[2] - slo[2], 1e-6f});
    pad = 1e-4f * ext + 1e-6f;
    nodes.reserve(2 * n / 3 + 2);
    if (n > 0) make(tris, c, 0, n);
  }

  float pad = 0.0f;

  int make(const Tri* tris, const std::vector<V>& c, int first, int count) {
    const int id = static_cast<int>(nodes.size());
    nodes.emplace_back();
    Node nd;
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = std::numeric_limits<float>::max();
      nd.hi[k] = -std::numeric_limits<float>::max();
    }
    for (int j = first; j < first + count; ++j)
      for (int vtx = 0; vtx < 3; ++vtx)
        for (int k = 0; k < 3; ++k) {
          nd.lo[k] = std::min(nd.lo[k], tris[idx[j]].v[vtx][k]);
          nd.hi[k] = std::max(nd.hi[k], tris[idx[j]].v[vtx][k]);
        }
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] -= pad;
      nd.hi[k] += pad;
    }
    if (count <= 4) {
      nd.first = first;
      nd.count = count;
      nodes[id] = nd;
      return id;
    }
    int ax = 0;
    float best = -1.0f;
    for (int k = 0; k < 3; ++k)
      if (nd.hi[k] - nd.lo[k] > best) {
        best = nd.hi[k] - nd.lo[k];
        ax = k;
      }
    const int mid = first + count / 2;
    auto key = [&](int t) { return ax == 0 ? c[t].x : ax == 1 ? c[t].y : c[t].z; };
    std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + first + count,
                     [&](int a, int b) { return key(a) < key(b) || (key(a) == key(b) && a < b); });
    nodes[id] = nd;
    const int l = make(tris, c, first, mid - first);
    const int r = make(tris, c, mid, first + count - mid);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }

  // Entry distance of the ray into the node's box, or +inf when it misses
  // the box within [0, tmax].
  static float enter(const Node& nd, V o, V d, float tmax) {
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float t0 = 0.0f, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
      if (dd[k] == 0.0f) {
        if (oo[k] < nd.lo[k] || oo[k] > nd.hi[k]) return std::numeric_limits<float>::infinity();
        continue;
      }
      float a = (nd.lo[k] - oo[k]) / dd[k], b = (nd.hi[k] - oo[k]) / dd[k];
      if (a > b) std::swap(a, b);
      t0 = std::max(t0, a);
      t1 = std::min(t1, b);
    }
    // a relative slack keeps the entry at or before any hit inside the box
    t0 = t0 - 1e-5f * std::fabs(t0);
    return t0 <= t1 * (1.0f + 1e-5f) ? t0 : std::numeric_limits<float>::infinity();
  }

  Hit closest(const Tri* tris, V o, V d, float tmax) const {
    Hit h{tmax, -1, 0.f, 0.f};
    if (nodes.empty()) return h;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const Node& nd = nodes[stack[--sp]];
      if (enter(nd, o, d, tmax) > h.t) continue;
      if (nd.count > 0) {
        for (int j = nd.first; j < nd.first + nd.count; ++j) {
          const int i = idx[j];
          float u, v;
          const float t = tri_hit(tris[i], o, d, u, v);
          if (t > kRayEps && t < tmax && (t < h.t || (t == h.t && h.tri >= 0 && i < h.tri)))
            h = Hit{t, i, u, v};
        }
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
    return h;
  }
};

Hit closest(const Scene& S, V o, V d, float tmax) {
  return S.bvh ? S.bvh->closest(S.tris, o, d, tmax) : closest_brute(S, o, d, tmax);
}

struct Surf {
  V pos, ns, ng, t, b;  // shading normal unnormalised, as the reference's
  int tri, mat;
};

void basis(V n, V& t, V& b) {
  if (std::fabs(n.y) < 0.9f) t = normalize(cross(n, V{0, 1, 0}));
  else t = normalize(cross(n, V{0, 0, -1}));
  b = normalize(cross(t, n));
}

Surf surface(const Scene& S, V o, V d, const Hit& h) {
  const Tri& tr = S.tris[h.tri];
  Surf s;
  s.pos = o + d * h.t;
  s.ns = ld(tr.n[0]) * (1.0f - h.u - h.v) + ld(tr.n[1]) * h.u + ld(tr.n[2]) * h.v;
  const V p0 = ld(tr.v[0]);
  s.ng = normalize(cross(ld(tr.v[1]) - p0, ld(tr.v[2]) - p0));
  basis(s.ns, s.t, s.b);
  s.tri = h.tri;
  s.mat = S.mat_of(h.tri);
  return s;
}

V to_local(V w, const Surf& s) { return {dot(w, s.t), dot(w, s.ns), dot(w, s.b)}; }
V to_world(V w, const Surf& s) {
  return {w.x * s.t.x + w.y * s.ns.x + w.z * s.b.x, w.x * s.t.y + w.y * s.ns.y + w.z * s.b.y,
          w.x * s.t.z + w.y * s.ns.z + w.z * s.b.z};
}

V cosine_hemisphere(float u, float v, float& pdf) {
  const float theta = 0.5f * std::acos(std::clamp(1.0f - 2.0f * u, -1.0f, 1.0f));
  const float phi = 2.0f * kPiF * v;
  pdf = std::cos(theta) / kPiF;
  return {std::cos(phi) * std::sin(theta), std::cos(theta), std::sin(phi) * std::sin(theta)};
}

float schlick(float c, float ei, float et) {
  const float f0 = (ei - et) * (ei - et) / ((ei + et) * (ei + et));
  const float x = std::max(1.0f - std::fabs(c), 0.0f);
  return f0 + (1.0f - f0) * x * x * x * x * x;
}

enum Transport { FROM_LIGHT, FROM_CAMERA };

float cos_term(V wo, V wi, const Surf& s, Transport td) {
  const float wi_ns = dot(wi, s.ns), wi_ng = dot(wi, s.ng);
  const float wo_ns = dot(wo, s.ns), wo_ng = dot(wo, s.ng);
  if (wi_ng * wi_ns <= 0 || wo_ng * wo_ns <= 0) return 0.0f;
  if (td == FROM_CAMERA) return std::fabs(wi_ns);
  return std::fabs(wo_ns) * std::fabs(wi_ng) / std::fabs(wo_ng);
}

// Primitive::sampleBxDF: f (world) and direction and pdf
V sample_bxdf(const Scene& S, const Surf& s, V wo, Transport td, Rng& rng, V& wi, float& pdf) {
  const MaterialRec& m = S.mats[s.mat];
  const V rho = ld(m.rho);
  const V wol = to_local(wo, s);
  V wil;
  V f{0, 0, 0};
  if (m.type == 0) {
    const float u = rng.next(), v = rng.next();
    wil = cosine_hemisphere(u, v, pdf);
    if (!(wol.y < 0 || wil.y < 0)) f = rho / kPiF;
  } else if (m.type == 1) {
    wil = {-wol.x, wol.y, -wol.z};
    pdf = 1.0f;
    f = rho / std::fabs(wil.y);
  } else {
    const float eta = S.ior ? S.ior[s.mat] : 1.5f;
    const bool outside = wol.y > 0;
    const float eo = outside ? 1.0f : eta, ei = outside ? eta : 1.0f;
    const V n = outside ? V{0, 1, 0} : V{0, -1, 0};
    const float c = dot(wol, n);
    const float fr = schlick(c, eo, ei);
    const V refl = wol * -1.0f + n * (2.0f * c);
    pdf = 1.0f;
    if (rng.next() < fr) {
      wil = refl;
      f = rho / std::fabs(wil.y);
    } else {
      const V th = (wol - n * c) * (-eo / ei);
      if (dot(th, th) > 1.0f) {
        wil = refl;
        f = rho / std::fabs(wil.y);
      } else {
        wil = th + n * (-std::sqrt(std::max(1.0f - dot(th, th), 0.0f)));
        const float sc = td == FROM_CAMERA ? (eo * eo) / (ei * ei) : 1.0f;
        f = rho * (sc / std::fabs(wil.y));
      }
    }
  }
  wi = to_world(wil, s);
  return f;
}

struct Map {
  const KdTreeOracle* tree;
  const float* wi_soa;
  const float* flux_soa;
  int64_t n;
  float n_emit;
};

V estimate(const Scene& S, const Map& M, const Surf& s, V wo, int k) {
  if (M.n == 0) return {0, 0, 0};
  std::vector<int32_t> idx(k);
  float r2 = 0.0f;
  const int cnt = M.tree->query(s.pos.x, s.pos.y, s.pos.z, k, idx.data(), &r2);
  QueryRec q;
  q.pos[0] = s.pos.x; q.pos[1] = s.pos.y; q.pos[2] = s.pos.z;
  q.ns[0] = s.ns.x; q.ns[1] = s.ns.y; q.ns[2] = s.ns.z;
  q.wo[0] = wo.x; q.wo[1] = wo.y; q.wo[2] = wo.z;
  q.material = S.tris[s.tri].material;
  float out[3];
  estimate_ref32(M.wi_soa, M.flux_soa, M.n, q, S.mats, S.nm, idx.data(), cnt, r2, M.n_emit,
                 out);
  return {out[0], out[1], out[2]};
}

struct PM {
  const Scene& S;
  Map global, caustics;
  int k, fg_depth, max_depth;

  V direct(V wo, const Surf& s, Rng& rng) const {
    const int nl = static_cast<int>(S.lights.size());
    unsigned li = static_cast<unsigned>(nl * rng.next());
    if (li == static_cast<unsigned>(nl)) --li;
    const float pdf_choose = 1.0f / nl;
    const Tri& lt = S.tris[S.lights[li]];
    const V p1 = ld(lt.v[0]), p2 = ld(lt.v[1]), p3 = ld(lt.v[2]);
    const float u0 = rng.next(), u1 = rng.next();
    const float su0 = std::sqrt(u0);
    const float b0 = 1.0f - su0, b1 = u1 * su0;
    const V lp = (1.0f - b0 - b1) * p1 + b0 * p2 + b1 * p3;
    const V lns = ld(lt.n[0]) * (1.0f - b0 - b1) + ld(lt.n[1]) * b0 + ld(lt.n[2]) * b1;
    const float area = 0.5f * std::sqrt(dot(cross(p2 - p1, p3 - p1), cross(p2 - p1, p3 - p1)));
    const float pdf_pos = 1.0f / area;
    const V wi = normalize(lp - s.pos);
    const float r = std::sqrt(dot(lp - s.pos, lp - s.pos));
    const float pdf_dir = pdf_pos * r * r / std::fabs(dot(wi * -1.0f, lns));
    if (closest(S, s.pos, wi, r - kRayEps).tri >= 0) return {0, 0, 0};
    const V le = S.tri_le[S.lights[li]];
    const MaterialRec& m = S.mats[s.mat];
    const V wol = to_local(wo, s), wil = to_local(wi, s);
    V f{0, 0, 0};
    if (m.type == 0 && !(wol.y < 0 || wil.y < 0)) f = ld(m.rho) / kPiF;
    const float c = std::fabs(dot(wi, s.ns));
    return mul(f * c, le) / (pdf_choose * pdf_dir);
  }

  V indirect(V wo, const Surf& s, Rng& rng, int depth) const {
    if (depth >= max_depth) return {0, 0, 0};
    V dir;
    float pdf;
    const V f = sample_bxdf(S, s, wo, FROM_CAMERA, rng, dir, pdf);
    const float c = std::fabs(dot(s.ns, dir));
    const Hit h = closest(S, s.pos, dir, std::numeric_limits<float>::max());
    if (h.tri < 0) return {0, 0, 0};
    const Surf s2 = surface(S, s.pos, dir, h);
    const V wo2 = dir * -1.0f;
    if (S.mats[s2.mat].type == 0) return mul(f * c, estimate(S, global, s2, wo2, k)) / pdf;
    return mul(f * c, indirect(wo2, s2, rng, depth + 1)) / pdf;
  }

  V integrate(V o, V d, Rng& rng, int depth) const {
    if (depth >= max_depth) return {0, 0, 0};
    const Hit h = closest(S, o, d, std::numeric_limits<float>::max());
    if (h.tri < 0) return {0, 0, 0};
    if (S.is_light(h.tri)) return S.tri_le[h.tri];
    const Surf s = surface(S, o, d, h);
    const V wo = d * -1.0f;
    const MaterialRec& m = S.mats[s.mat];
    if (m.type == 0) {
      if (depth >= fg_depth) return estimate(S, global, s, wo, k);
      const V ld_ = direct(wo, s, rng);
      const V lc = estimate(S, caustics, s, wo, k);
      const V li = indirect(wo, s, rng, 0);
      return ld_ + lc + li;
    }
    if (depth >= 3) {
      V dir;
      float pdf;
      const V f = sample_bxdf(S, s, wo, FROM_CAMERA, rng, dir, pdf);
      const V tp = f * cos_term(wo, dir, s, FROM_CAMERA) / pdf;
      return mul(tp, integrate(s.pos, dir, rng, depth + 1));
    }
    // sampleAllDirection
    const V wol = to_local(wo, s);
    const V rho = ld(m.rho);
    V dirs[2];
    V fs[2];
    int nb = 0;
    if (m.type == 1) {
      const V wi = {-wol.x, wol.y, -wol.z};
      dirs[0] = wi;
      fs[0] = rho / std::fabs(wi.y);
      nb = 1;
    } else {
      const float eta = S.ior ? S.ior[s.mat] : 1.5f;
      const bool outside = wol.y > 0;
      const float eo = outside ? 1.0f : eta, ei = outside ? eta : 1.0f;
      const V n = outside ? V{0, 1, 0} : V{0, -1, 0};
      const float c = dot(wol, n);
      const float fr = schlick(c, eo, ei);
      const V wr = wol * -1.0f + n * (2.0f * c);
      dirs[0] = wr;
      fs[0] = fr * rho / std::fabs(wr.y);
      nb = 1;
      const V th = (wol - n * c) * (-eo / ei);
      if (dot(th, th) > 1.0f) {
        fs[0] = rho / std::fabs(wr.y);
      } else {
        const V tr = th + n * (-std::sqrt(std::max(1.0f - dot(th, th), 0.0f)));
        dirs[1] = tr;
        fs[1] = (1.0f - fr) * ((eo * eo) / (ei * ei)) * rho / std::fabs(tr.y);
        nb = 2;
      }
    }
    V lo{0, 0, 0};
    for (int i = 0; i < nb; ++i) {
      const V dir = to_world(dirs[i], s);
      const V tp = fs[i] * std::fabs(dot(dir, s.ns));
      lo = lo + mul(tp, integrate(s.pos, dir, rng, depth + 1));
    }
    return lo;
  }
};

V path_trace(const Scene& S, V o, V d, Rng& rng, int max_depth) {
  V L{0, 0, 0}, tp{1, 1, 1};
  for (int k = 0; k < max_depth; ++k) {
    const Hit h = closest(S, o, d, std::numeric_limits<float>::max());
    if (h.tri < 0) break;
    if (k > 0) {
      const float p = std::min(std::max(tp.x, std::max(tp.y, tp.z)), 1.0f);
      if (rng.next() >= p) break;
      tp = tp / p;
    }
    if (S.is_light(h.tri)) L = L + mul(tp, S.tri_le[h.tri]);
    const Surf s = surface(S, o, d, h);
    V dir;
    float pdf;
    const V f = sample_bxdf(S, s, d * -1.0f, FROM_CAMERA, rng, dir, pdf);
    tp = mul(tp, f * cos_term(d * -1.0f, dir, s, FROM_CAMERA) / pdf);
    o = s.pos;
    d = dir;
  }
  return L;
}

}  // namespace

extern "C" {

// Renders width x height pixels at spp samples.  mode 0: path tracing;
// mode 1: photon mapping over the global (and caustics) photon arrays (SoA,
// 3n floats each, original order).  cam: position[3], forward[3], right[3],
// up[3], focal.  Outputs (height x width x 3, row-major): the pixel means and
// the variances of the means.  Returns the number of dropped samples.
int64_t orc_render(int mode, const void* tris, int n_tri, const void* mats, const float* ior,
                   int n_mat, const float* le, const float* mat_le, const float* cam, int width, int height, int spp,
                   int max_depth, int fg_depth, int k, int64_t n_global, const float* g_pos,
                   const float* g_wi, const float* g_flux, float n_emit_global,
                   int64_t n_caustics, const float* c_pos, const float* c_wi,
                   const float* c_flux, float n_emit_caustics, uint64_t seed, int nthreads,
                   float* out_mean, float* out_var) {
  Scene S{static_cast<const Tri*>(tris), n_tri, static_cast<const MaterialRec*>(mats), ior,
          n_mat, {}, {}};
  // a face whose material emits is a light of that emission
  // (include/scene.h:41-52); otherwise the is_light flag with le
  S.tri_le.assign(n_tri, V{0, 0, 0});
  for (int i = 0; i < n_tri; ++i) {
    const int m = S.mat_of(i);
    if (mat_le && (mat_le[3 * m] > 0 || mat_le[3 * m + 1] > 0 || mat_le[3 * m + 2] > 0))
      S.tri_le[i] = V{mat_le[3 * m], mat_le[3 * m + 1], mat_le[3 * m + 2]};
    else if (le && S.tris[i].is_light)
      S.tri_le[i] = V{le[0], le[1], le[2]};
    if (S.is_light(i)) S.lights.push_back(i);
  }
  Bvh bvh;
  bvh.build(S.tris, n_tri);
  S.bvh = &bvh;
  KdTreeOracle gt, ct;
  auto build = [](KdTreeOracle& t, int64_t n, const float* p) {
    t.n = static_cast<int>(n);
    t.x.assign(p, p + n);
    t.y.assign(p + n, p + 2 * n);
    t.z.assign(p + 2 * n, p + 3 * n);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    t.nodes.reserve(n);
    t.build(idx.data(), static_cast<int>(n), 0);
  };
  if (mode == 1) {
    build(gt, n_global, g_pos);
    build(ct, n_caustics, c_pos);
  }
  PM pm{S, Map{&gt, g_wi, g_flux, n_global, n_emit_global},
        Map{&ct, c_wi, c_flux, n_caustics, n_emit_caustics}, k, fg_depth, max_depth};
  const V pos{cam[0], cam[1], cam[2]}, fwd{cam[3], cam[4], cam[5]}, rgt{cam[6], cam[7], cam[8]},
      up{cam[9], cam[10], cam[11]};
  const float focal = cam[12];
  int64_t dropped = 0;
  const int threads = resolve_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : dropped)
  for (int64_t pix = 0; pix < static_cast<int64_t>(width) * height; ++pix) {
    const int i = static_cast<int>(pix / width), j = static_cast<int>(pix % width);
    Rng rng{seed * 0x2545f4914f6cdd1dull + static_cast<uint64_t>(pix) * 0xd1b54a32d192ed03ull};
    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (int sp = 0; sp < spp; ++sp) {
      const float u = (2.0f * (j + rng.next()) - width) / height;
      const float v = (2.0f * (i + rng.next()) - height) / height;
      const V sensor = pos + u * rgt + v * up;
      const V d = normalize(pos + focal * fwd - sensor);
      const V L = mode == 0 ? path_trace(S, sensor, d, rng, max_depth)
                            : pm.integrate(sensor, d, rng, 0);
      if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z) || L.x < 0 || L.y < 0 ||
          L.z < 0) {
        ++dropped;
        continue;
      }
      const double c[3] = {L.x, L.y, L.z};
      for (int ch = 0; ch < 3; ++ch) {
        s[ch] += c[ch];
        q[ch] += c[ch] * c[ch];
      }
    }
    for (int ch = 0; ch < 3; ++ch) {
      const double m = s[ch] / spp;
      out_mean[3 * pix + ch] = static_cast<float>(m);
      const double var = spp > 1 ? (q[ch] - spp * m * m) / (double(spp) * (spp - 1)) : 0.0;
      out_var[3 * pix + ch] = static_cast<float>(var > 0 ? var : 0.0);
    }
  }
  return dropped;
}

// Gather query points of a camera pass, restated for checking the device
// producer's distribution.  Per sample: draw a camera ray through the pixel
// (examples/main.cpp:40-41); a miss, a light or a non-Lambert primary hit is
// redrawn (at most 64 draws, then a placeholder of material -1).  depth 0:
// the primary hit is the query (include/integrator.h:305-308).  depth 1: up
// to 16 cosine-sampled final-gathering rays from the primary hit, each
// followed through at most 17 specular hits sampled as camera paths sample
// them, until one reaches a Lambert surface (include/integrator.h:215-252).
// out: m records of 10 floats {pos, shading normal, wo, material as int32
// bits}; fg_mat (may be null): the primary hit's material, -1 if none.
void orc_fg_queries(const void* tris, int n_tri, const void* mats, const float* ior, int n_mat,
                    const float* le, const float* mat_le, const float* cam, int width,
                    int height, int spp, int depth, uint64_t seed, int nthreads, float* out,
                    int32_t* fg_mat) {
  Scene S{static_cast<const Tri*>(tris), n_tri, static_cast<const MaterialRec*>(mats), ior,
          n_mat, {}, {}};
  S.tri_le.assign(n_tri, V{0, 0, 0});
  for (int i = 0; i < n_tri; ++i) {
    const int m = S.mat_of(i);
    if (mat_le && (mat_le[3 * m] > 0 || mat_le[3 * m + 1] > 0 || mat_le[3 * m + 2] > 0))
      S.tri_le[i] = V{mat_le[3 * m], mat_le[3 * m + 1], mat_le[3 * m + 2]};
    else if (le && S.tris[i].is_light)
      S.tri_le[i] = V{le[0], le[1], le[2]};
  }
  Bvh bvh;
  bvh.build(S.tris, n_tri);
  S.bvh = &bvh;
  const V pos{cam[0], cam[1], cam[2]}, fwd{cam[3], cam[4], cam[5]}, rgt{cam[6], cam[7], cam[8]},
      up{cam[9], cam[10], cam[11]};
  const float focal = cam[12];
  const int64_t total = static_cast<int64_t>(width) * height * spp;
  const int threads = resolve_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
  for (int64_t id = 0; id < total; ++id) {
    const int64_t pix = id / spp;
    const int i = static_cast<int>(pix / width), j = static_cast<int>(pix % width);
    Rng rng{seed * 0x9e3779b97f4a7c15ull + static_cast<uint64_t>(id) * 0xd1b54a32d192ed03ull};
    V qp{0, 0, 0}, qn{0, 1, 0}, qwo{0, 1, 0};
    int qmat = -1, pmat = -1;
    for (int attempt = 0; attempt < 64; ++attempt) {
      const float u = (2.0f * (j + rng.next()) - width) / height;
      const float v = (2.0f * (i + rng.next()) - height) / height;
      const V sensor = pos + u * rgt + v * up;
      const V d0 = normalize(pos + focal * fwd - sensor);
      const Hit h0 = closest(S, sensor, d0, std::numeric_limits<float>::max());
      if (h0.tri < 0 || S.tris[h0.tri].is_light || S.is_light(h0.tri)) continue;
      const Surf s0 = surface(S, sensor, d0, h0);
      if (S.mats[s0.mat].type != 0) continue;
      if (depth == 0) {
        qp = s0.pos; qn = s0.ns; qwo = d0 * -1.0f; qmat = S.tris[h0.tri].material;
        break;
      }
      bool found = false;
      for (int bounce = 0; bounce < 16 && !found; ++bounce) {
        float pdf;
        const V dl = cosine_hemisphere(rng.next(), rng.next(), pdf);
        V d = to_world(dl, s0);
        V o = s0.pos;
        for (int chain = 0; chain <= 16; ++chain) {
          const Hit h1 = closest(S, o, d, std::numeric_limits<float>::max());
          if (h1.tri < 0) break;
          const Surf s1 = surface(S, o, d, h1);
          if (S.mats[s1.mat].type == 0) {
            qp = s1.pos; qn = s1.ns; qwo = d * -1.0f; qmat = S.tris[h1.tri].material;
            pmat = S.tris[h0.tri].material;
            found = true;
            break;
          }
          V wi;
          float spdf;
          (void)sample_bxdf(S, s1, d * -1.0f, FROM_CAMERA, rng, wi, spdf);
          o = s1.pos;
          d = wi;
        }
      }
      if (found) break;
    }
    float* r = out + 10 * id;
    r[0] = qp.x; r[1] = qp.y; r[2] = qp.z;
    r[3] = qn.x; r[4] = qn.y; r[5] = qn.z;
    r[6] = qwo.x; r[7] = qwo.y; r[8] = qwo.z;
    std::memcpy(r + 9, &qmat, sizeof(int32_t));
    if (fg_mat) fg_mat[id] = pmat;
  }
}

// Closest hits of n rays (origins and directions as n x 3 floats), by the
// BVH (method 0) or the brute-force loop (method 1): triangle (-1 on a miss)
// and t (tmax on a miss).
void orc_closest_hits(const void* tris, int n_tri, int64_t n, const float* org,
                      const float* dir, float tmax, int method, int nthreads, int32_t* out_tri,
                      float* out_t) {
  Scene S{static_cast<const Tri*>(tris), n_tri, nullptr, nullptr, 0, {}, {}};
  Bvh bvh;
  if (method == 0) {
    bvh.build(S.tris, n_tri);
    S.bvh = &bvh;
  }
  const int threads = resolve_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) {
    const Hit h = closest(S, ld(org + 3 * i), ld(dir + 3 * i), tmax);
    out_tri[i] = h.tri;
    out_t[i] = h.t;
  }
}

}  // extern "C"
