// pm/render.hpp -- header-only C++ drop-in for the reference's Scene
// (include/scene.h), Camera (include/camera.h), PathTracing and
// PhotonMapping integrators (include/integrator.h) and Image
// (include/image.h), backed by libpm_gather.so.
//
// Same class names, method names and argument meaning.  The reference
// renders by calling Integrator::integrate(ray, scene, sampler) for every
// pixel sample from an OpenMP loop (examples/main.cpp:33-70); that call is
// here (one ray through the device renderer per call, serialised: the slow
// path), and Integrator::render(scene, camera, width, height, n_samples,
// seed) runs the whole loop on the GPU and returns the averaged Image
// (dropped NaN / negative samples and the division by n_samples as in the
// reference's loop).  Scene::loadModel reads OBJ + MTL with the reference's
// material semantics; Scene::build uploads the scene and builds its BVH.
// PhotonMapping::build traces nPhotonsGlobal (and nPhotonsGlobal x
// multiplier caustic) paths and builds both maps on the GPU.
//
// The vector type is PM_VEC3F_TYPE (default pm::Vec3f), as in
// pm/photon_map.hpp.  Errors throw pm::Error.
#ifndef PM_RENDER_HPP
#define PM_RENDER_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "pm/photon_map.hpp"
#include "pm_scene.h"

namespace pm {

constexpr float kPi = 3.14159265359f;  // include/core.h:7

namespace detail {
inline void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
inline void normalize3(float* a) {
  const float l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  for (int i = 0; i < 3; ++i) a[i] = a[i] / l;
}
}  // namespace detail

template <typename V>
struct BasicRay {
  V origin;
  V direction;
  BasicRay() = default;
  BasicRay(const V& origin, const V& direction) : origin(origin), direction(direction) {}
  V operator()(float t) const {
    V p;
    for (int i = 0; i < 3; ++i) p[i] = origin[i] + t * direction[i];
    return p;
  }
};

namespace detail {
// PCG32 with the XSH-RR output function (M. E. O'Neill, "PCG: A family of
// simple fast space-efficient statistically good algorithms for random
// number generation", 2014): 64-bit LCG state, increment 1, the way the
// reference's RNG seeds it (include/sampler.h:27-48).
struct Pcg32 {
  uint64_t state = 1;
  uint32_t next() {
    const uint64_t old = state;
    state = old * 6364136223846793005ull + 1ull;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
};

// Every device call of the per-ray path goes through one lock: the
// renderer's workspaces belong to the scene's context.
inline std::mutex& device_mutex() {
  static std::mutex m;
  return m;
}
}  // namespace detail

// Sampler interface (include/sampler.h:51-66): a seeded random stream of
// floats in [0, 1].
template <typename V2>
class BasicSampler {
 public:
  BasicSampler() = default;
  explicit BasicSampler(uint64_t seed) { rng.state = seed; }
  virtual ~BasicSampler() = default;
  uint64_t getSeed() const { return rng.state; }
  void setSeed(uint64_t seed) { rng.state = seed; }
  virtual std::unique_ptr<BasicSampler> clone() const = 0;
  virtual float getNext1D() = 0;
  virtual V2 getNext2D() = 0;

 protected:
  detail::Pcg32 rng;
};

// UniformSampler (include/sampler.h:68-80): u32 * (1 / (2^32 - 1)).
template <typename V2>
class BasicUniformSampler : public BasicSampler<V2> {
 public:
  BasicUniformSampler() = default;
  BasicUniformSampler(uint64_t seed) : BasicSampler<V2>(seed) {}
  std::unique_ptr<BasicSampler<V2>> clone() const override {
    return std::make_unique<BasicUniformSampler>();
  }
  float getNext1D() override {
    constexpr float divider = 1.0f / 4294967295.0f;
    return static_cast<float>(this->rng.next()) * divider;
  }
  V2 getNext2D() override {
    const float a = getNext1D();
    const float b = getNext1D();
    return V2(a, b);
  }
};

// Surface point of a light sample (include/core.h:268-276 SurfaceInfo, the
// members a light sample fills).
template <typename V, typename V2>
struct BasicSurfaceInfo {
  V position;
  V geometricNormal;
  V shadingNormal;  // interpolated, not renormalised
  V dpdu;
  V dpdv;
  V2 barycentric;
};

// AreaLight (include/light.h:17-45): one emitting triangle of constant
// radiance on both sides, sampled uniformly by area, emitting
// cosine-distributed about its shading normal.
template <typename V, typename V2>
class BasicAreaLight {
 public:
  using SurfaceInfo = BasicSurfaceInfo<V, V2>;
  BasicAreaLight(const pmx_triangle& tri, const float le[3]) : tri(tri) {
    for (int i = 0; i < 3; ++i) this->le[i] = le[i];
    float e1[3], e2[3], c[3];
    for (int a = 0; a < 3; ++a) {
      e1[a] = tri.v[1][a] - tri.v[0][a];
      e2[a] = tri.v[2][a] - tri.v[0][a];
    }
    detail::cross3(e1, e2, c);
    area = 0.5f * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  }
  V Le(const SurfaceInfo&, const V&) const { return le; }
  float surfaceArea() const { return area; }

  // Triangle::samplePoint (include/triangle.h:84-114): barycentric
  // (1 - sqrt(u0), u1 sqrt(u0)), pdf 1 / area
  SurfaceInfo samplePoint(BasicSampler<V2>& sampler, float& pdf) const {
    const V2 uv = sampler.getNext2D();
    const float su0 = std::sqrt(uv[0]);
    const float b0 = 1.0f - su0, b1 = uv[1] * su0;
    SurfaceInfo s;
    float ng[3], e1[3], e2[3], n[3], t[3], b[3];
    for (int a = 0; a < 3; ++a) {
      s.position[a] = (1.0f - b0 - b1) * tri.v[0][a] + b0 * tri.v[1][a] + b1 * tri.v[2][a];
      n[a] = (1.0f - b0 - b1) * tri.n[0][a] + b0 * tri.n[1][a] + b1 * tri.n[2][a];
      s.shadingNormal[a] = n[a];
      e1[a] = tri.v[1][a] - tri.v[0][a];
      e2[a] = tri.v[2][a] - tri.v[0][a];
    }
    detail::cross3(e1, e2, ng);
    detail::normalize3(ng);
    basis(n, t, b);
    for (int a = 0; a < 3; ++a) {
      s.geometricNormal[a] = ng[a];
      s.dpdu[a] = t[a];
      s.dpdv[a] = b[a];
    }
    s.barycentric = V2(b0, b1);
    pdf = 1.0f / area;
    return s;
  }

  // AreaLight::sampleDirection: cosine about the shading normal, pdf cos / pi
  V sampleDirection(const SurfaceInfo& s, BasicSampler<V2>& sampler, float& pdf) const {
    const V2 uv = sampler.getNext2D();
    const float theta = 0.5f * std::acos(std::min(1.0f, std::max(-1.0f, 1.0f - 2.0f * uv[0])));
    const float phi = 2.0f * kPi * uv[1];
    const float l[3] = {std::cos(phi) * std::sin(theta), std::cos(theta),
                        std::sin(phi) * std::sin(theta)};
    pdf = std::cos(theta) / kPi;
    V d;
    for (int a = 0; a < 3; ++a)
      d[a] = l[0] * s.dpdu[a] + l[1] * s.shadingNormal[a] + l[2] * s.dpdv[a];
    return d;
  }

 private:
  // orthonormalBasis (include/core.h:224-231)
  static void basis(const float n[3], float t[3], float b[3]) {
    const float up[3] = {0.f, 1.f, 0.f}, back[3] = {0.f, 0.f, -1.f};
    detail::cross3(n, std::fabs(n[1]) < 0.9f ? up : back, t);
    detail::normalize3(t);
    detail::cross3(t, n, b);
    detail::normalize3(b);
  }

  pmx_triangle tri;
  V le;
  float area = 0.0f;
};

// Pinhole camera (include/camera.h:10-47): right = normalize(forward x y),
// up = normalize(right x forward), focal length 1 / tan(FOV / 2); the sensor
// centre is `position` and the pinhole sits at position + focal * forward.
template <typename V>
class BasicCamera {
 public:
  BasicCamera(const V& position, const V& forward, float FOV = 0.5f * kPi) {
    for (int i = 0; i < 3; ++i) {
      cam.position[i] = position[i];
      cam.forward[i] = forward[i];
    }
    const float y[3] = {0.f, 1.f, 0.f};
    detail::cross3(cam.forward, y, cam.right);
    detail::normalize3(cam.right);
    detail::cross3(cam.right, cam.forward, cam.up);
    detail::normalize3(cam.up);
    cam.focal = 1.0f / std::tan(0.5f * FOV);
  }

  // uv in [-1, 1]^2 on the sensor (a Vec2f, or any two floats with
  // operator[]); pdf 1 (include/camera.h:41-47).
  template <typename UV>
  bool sampleRay(const UV& uv, BasicRay<V>& ray, float& pdf) const {
    float s[3], d[3];
    for (int i = 0; i < 3; ++i) {
      s[i] = cam.position[i] + uv[0] * cam.right[i] + uv[1] * cam.up[i];
      d[i] = cam.position[i] + cam.focal * cam.forward[i] - s[i];
    }
    detail::normalize3(d);
    for (int i = 0; i < 3; ++i) {
      ray.origin[i] = s[i];
      ray.direction[i] = d[i];
    }
    pdf = 1.0f;
    return true;
  }

  const pmx_camera& desc() const { return cam; }

 private:
  pmx_camera cam{};
};

// Hit record of Scene::intersect (include/core.h SurfaceInfo / IntersectInfo,
// without the primitive pointer: primID indexes Scene::triangles()).
template <typename V>
struct BasicIntersectInfo {
  float t = 0.0f;
  int primID = -1;
  V position;
  V geometricNormal;
  V shadingNormal;   // interpolated, not renormalised (include/triangle.h:64-71)
  float barycentric[2] = {0.0f, 0.0f};
};

// Triangle scene (include/scene.h:54-356).
class Scene {
 public:
  Scene() = default;
  Scene(const Scene&) = delete;
  Scene& operator=(const Scene&) = delete;
  ~Scene() { release(); }

  // Scene::loadModel (include/scene.h:109-270): OBJ + MTL, polygons
  // fan-triangulated, vertex normals when a face has them for every vertex,
  // else the winding normal; illum 5 = mirror (albedo 1), illum 7 = glass
  // (albedo 1, IOR Ni), otherwise Lambert Kd; faces without a material are
  // Lambert 0.9; a positive Ke makes the face an area light.  MTL files are
  // read from the OBJ file's directory.
  void loadModel(const std::string& filepath) {
    clear();
    std::ifstream in(filepath);
    if (!in) throw Error(PM_ERR_INVALID_ARG, "Scene::loadModel: cannot open " + filepath);
    const std::string dir = filepath.substr(0, filepath.find_last_of("/\\") + 1);
    std::vector<float> v, vn;
    std::map<std::string, Mtl> mtl;
    std::vector<std::string> used;
    std::vector<int> face_mat;  // index into used, -1: none
    std::string cur;
    bool has_cur = false;
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string tag;
      if (!(ls >> tag) || tag[0] == '#') continue;
      if (tag == "v" || tag == "vn") {
        float x, y, z;
        ls >> x >> y >> z;
        auto& dst = tag == "v" ? v : vn;
        dst.insert(dst.end(), {x, y, z});
      } else if (tag == "mtllib") {
        std::string f;
        while (ls >> f) read_mtl(dir + f, mtl);
      } else if (tag == "usemtl") {
        std::getline(ls >> std::ws, cur);
        has_cur = true;
      } else if (tag == "f") {
        std::vector<int> vi, ni;
        std::string ref;
        while (ls >> ref) {
          const size_t s1 = ref.find('/');
          const int a = std::stoi(ref.substr(0, s1));
          vi.push_back(a > 0 ? a - 1 : static_cast<int>(v.size() / 3) + a);
          int b = -1;
          if (s1 != std::string::npos) {
            const size_t s2 = ref.find('/', s1 + 1);
            if (s2 != std::string::npos && s2 + 1 < ref.size()) {
              const int bb = std::stoi(ref.substr(s2 + 1));
              b = bb > 0 ? bb - 1 : static_cast<int>(vn.size() / 3) + bb;
            }
          }
          ni.push_back(b);
        }
        int mi = -1;
        if (has_cur) {
          auto it = std::find(used.begin(), used.end(), cur);
          if (it == used.end()) {
            used.push_back(cur);
            mi = static_cast<int>(used.size()) - 1;
          } else {
            mi = static_cast<int>(it - used.begin());
          }
        }
        for (size_t j = 1; j + 1 < vi.size(); ++j) {
          const int idx[3] = {vi[0], vi[j], vi[j + 1]};
          const int nidx[3] = {ni[0], ni[j], ni[j + 1]};
          pmx_triangle t{};
          for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) t.v[c][a] = v[3 * idx[c] + a];
          if (nidx[0] >= 0 && nidx[1] >= 0 && nidx[2] >= 0) {
            for (int c = 0; c < 3; ++c)
              for (int a = 0; a < 3; ++a) t.n[c][a] = vn[3 * nidx[c] + a];
          } else {
            float e1[3], e2[3], n[3];
            for (int a = 0; a < 3; ++a) {
              e1[a] = t.v[1][a] - t.v[0][a];
              e2[a] = t.v[2][a] - t.v[0][a];
            }
            detail::normalize3(e1);
            detail::normalize3(e2);
            detail::cross3(e1, e2, n);
            detail::normalize3(n);
            for (int c = 0; c < 3; ++c)
              for (int a = 0; a < 3; ++a) t.n[c][a] = n[a];
          }
          tris.push_back(t);
          face_mat.push_back(mi);
        }
      }
    }
    // materials: the used MTL materials in first-use order, then Lambert 0.9
    for (const std::string& name : used) {
      const auto it = mtl.find(name);
      const Mtl m = it == mtl.end() ? Mtl{} : it->second;
      pm_material pmat{};
      if (m.illum == 5 || m.illum == 7) {
        pmat.rho[0] = pmat.rho[1] = pmat.rho[2] = 1.0f;
        pmat.type = m.illum == 5 ? PM_MATERIAL_MIRROR : PM_MATERIAL_GLASS;
      } else {
        for (int a = 0; a < 3; ++a) pmat.rho[a] = m.kd[a];
        pmat.type = PM_MATERIAL_LAMBERT;
      }
      mats.push_back(pmat);
      ior.push_back(m.ni);
      for (int a = 0; a < 3; ++a) mat_le.push_back(m.ke[a]);
    }
    const int dflt = static_cast<int>(used.size());
    mats.push_back(pm_material{{0.9f, 0.9f, 0.9f}, PM_MATERIAL_LAMBERT});
    ior.push_back(1.0f);
    mat_le.insert(mat_le.end(), {0.f, 0.f, 0.f});
    for (size_t f = 0; f < tris.size(); ++f) tris[f].material = face_mat[f] < 0 ? dflt : face_mat[f];
  }

  // Scene::build (include/scene.h:275-305): device copy and BVH.
  void build() {
    release();
    if (tris.empty()) throw Error(PM_ERR_INVALID_ARG, "Scene::build: no triangles");
    check(pmx_scene_create(default_context(), tris.data(), static_cast<int>(tris.size()),
                           mats.data(), ior.data(), static_cast<int>(mats.size()), nullptr,
                           mat_le.data(), &scene),
          "pmx_scene_create");
  }

  // Scene::intersect (include/scene.h:307-348) for one ray: the closest hit
  // in (1e-5, inf) over both faces.  Stages through the device and
  // synchronises; batches of rays go through intersectBatch.
  template <typename V>
  bool intersect(const BasicRay<V>& ray, BasicIntersectInfo<V>& info) const {
    std::vector<int32_t> prim;
    std::vector<float> t, u, v;
    intersectBatch(std::vector<BasicRay<V>>{ray}, prim, t, u, v);
    if (prim[0] < 0) return false;
    const pmx_triangle& tr = tris[prim[0]];
    info.t = t[0];
    info.primID = prim[0];
    info.barycentric[0] = u[0];
    info.barycentric[1] = v[0];
    float e1[3], e2[3], ng[3];
    for (int a = 0; a < 3; ++a) {
      info.position[a] = ray.origin[a] + t[0] * ray.direction[a];
      info.shadingNormal[a] =
          tr.n[0][a] * (1.0f - u[0] - v[0]) + tr.n[1][a] * u[0] + tr.n[2][a] * v[0];
      e1[a] = tr.v[1][a] - tr.v[0][a];
      e2[a] = tr.v[2][a] - tr.v[0][a];
    }
    detail::cross3(e1, e2, ng);
    detail::normalize3(ng);
    for (int a = 0; a < 3; ++a) info.geometricNormal[a] = ng[a];
    return true;
  }

  // Closest hits of many rays: triangle index (-1: miss), t, barycentrics.
  template <typename V>
  void intersectBatch(const std::vector<BasicRay<V>>& rays, std::vector<int32_t>& prim,
                      std::vector<float>& t, std::vector<float>& u,
                      std::vector<float>& v) const {
    const size_t n = rays.size();
    std::vector<float> org(3 * n), dir(3 * n);
    for (size_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        org[a * n + i] = rays[i].origin[a];
        dir[a * n + i] = rays[i].direction[a];
      }
    prim.assign(n, -1);
    t.assign(n, 0.0f);
    u.assign(n, 0.0f);
    v.assign(n, 0.0f);
    check(pmx_scene_intersect(handle(), static_cast<int64_t>(n), org.data(), dir.data(),
                              PMX_HIT_BVH, prim.data(), t.data(), u.data(), v.data(),
                              PM_MEM_HOST, nullptr),
          "pmx_scene_intersect");
  }

  // Scene::sampleLight (include/scene.h:350-355): one of the emitting
  // triangles chosen uniformly, pdf 1 / (number of lights).
  template <typename V = PM_VEC3F_TYPE, typename V2>
  std::shared_ptr<BasicAreaLight<V, V2>> sampleLight(BasicSampler<V2>& sampler,
                                                     float& pdf) const {
    const std::vector<int> lights = lightFaces();
    if (lights.empty()) throw Error(PM_ERR_INVALID_ARG, "Scene::sampleLight: no light");
    unsigned int idx = static_cast<unsigned int>(lights.size() * sampler.getNext1D());
    if (idx == lights.size()) --idx;
    pdf = 1.0f / static_cast<float>(lights.size());
    const pmx_triangle& t = tris[lights[idx]];
    return std::make_shared<BasicAreaLight<V, V2>>(t, &mat_le[3 * t.material]);
  }

  // the emitting triangles: those whose material has a positive Ke
  // (include/scene.h:41-52, 249-265)
  std::vector<int> lightFaces() const {
    std::vector<int> out;
    for (size_t f = 0; f < tris.size(); ++f) {
      const float* le = &mat_le[3 * tris[f].material];
      if (le[0] > 0.f || le[1] > 0.f || le[2] > 0.f) out.push_back(static_cast<int>(f));
    }
    return out;
  }

  uint32_t nVertices() const { return static_cast<uint32_t>(3 * tris.size()); }
  uint32_t nFaces() const { return static_cast<uint32_t>(tris.size()); }
  const std::vector<pmx_triangle>& triangles() const { return tris; }
  const std::vector<pm_material>& materials() const { return mats; }
  pmx_scene* handle() const {
    if (!scene) throw Error(PM_ERR_NOT_BUILT, "Scene used before build()");
    return scene;
  }

 private:
  struct Mtl {
    float kd[3] = {0.f, 0.f, 0.f};
    float ke[3] = {0.f, 0.f, 0.f};
    int illum = 0;
    float ni = 1.0f;
  };

  static void read_mtl(const std::string& path, std::map<std::string, Mtl>& out) {
    std::ifstream in(path);
    if (!in) throw Error(PM_ERR_INVALID_ARG, "Scene::loadModel: cannot open " + path);
    std::string line, name;
    Mtl* cur = nullptr;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string tag;
      if (!(ls >> tag) || tag[0] == '#') continue;
      if (tag == "newmtl") {
        std::getline(ls >> std::ws, name);
        cur = &out[name];
        *cur = Mtl{};
      } else if (cur && tag == "Kd") {
        ls >> cur->kd[0] >> cur->kd[1] >> cur->kd[2];
      } else if (cur && tag == "Ke") {
        ls >> cur->ke[0] >> cur->ke[1] >> cur->ke[2];
      } else if (cur && tag == "illum") {
        ls >> cur->illum;
      } else if (cur && tag == "Ni") {
        ls >> cur->ni;
      }
    }
  }

  void clear() {
    release();
    tris.clear();
    mats.clear();
    ior.clear();
    mat_le.clear();
  }
  void release() {
    if (scene) pmx_scene_destroy(scene);
    scene = nullptr;
  }

  std::vector<pmx_triangle> tris;
  std::vector<pm_material> mats;
  std::vector<float> ior;
  std::vector<float> mat_le;
  pmx_scene* scene = nullptr;
};

// Image (include/image.h): row-major RGB floats.
template <typename V>
class BasicImage {
 public:
  BasicImage(unsigned int width, unsigned int height)
      : w(width), h(height), pixels(3 * static_cast<size_t>(width) * height, 0.0f) {}
  unsigned int getWidth() const { return w; }
  unsigned int getHeight() const { return h; }
  V getPixel(unsigned int i, unsigned int j) const {
    const float* p = &pixels[3 * (static_cast<size_t>(i) * w + j)];
    return V(p[0], p[1], p[2]);
  }
  void setPixel(unsigned int i, unsigned int j, const V& c) {
    float* p = &pixels[3 * (static_cast<size_t>(i) * w + j)];
    for (int a = 0; a < 3; ++a) p[a] = c[a];
  }
  void addPixel(unsigned int i, unsigned int j, const V& c) {
    float* p = &pixels[3 * (static_cast<size_t>(i) * w + j)];
    for (int a = 0; a < 3; ++a) p[a] += c[a];
  }
  void divide(float k) {
    for (float& x : pixels) x /= k;
  }
  void gammaCorrection(float gamma) {
    for (float& x : pixels) x = std::pow(x, 1.0f / gamma);
  }
  // ASCII P3, uint(255 c) clamped to [0, 255] (include/image.h:71-90)
  void writePPM(const std::string& filename) const {
    std::ofstream file(filename);
    file << "P3\n" << w << " " << h << "\n255\n";
    for (size_t p = 0; p < static_cast<size_t>(w) * h; ++p) {
      unsigned int c[3];
      for (int a = 0; a < 3; ++a) {
        const float x = 255.0f * pixels[3 * p + a];
        c[a] = x <= 0.0f ? 0u : (x >= 255.0f ? 255u : static_cast<unsigned int>(x));
      }
      file << c[0] << " " << c[1] << " " << c[2] << "\n";
    }
  }
  std::vector<float>& data() { return pixels; }
  const std::vector<float>& data() const { return pixels; }

 private:
  unsigned int w, h;
  std::vector<float> pixels;
};

template <typename V, typename V2 = PM_VEC2F_TYPE>
class BasicIntegrator {
 public:
  using Image = BasicImage<V>;
  using Camera = BasicCamera<V>;
  using Ray = BasicRay<V>;
  using Sampler = BasicSampler<V2>;
  virtual ~BasicIntegrator() = default;
  // Integrator::build (include/integrator.h:13): the photon maps' random
  // streams start from sampler.getSeed()
  virtual void build(const Scene& scene, Sampler& sampler) = 0;
  // Integrator::integrate (include/integrator.h:16-17): one radiance sample
  // along the ray.  Each call runs one ray through the device renderer (its
  // random stream seeded from two of the sampler's numbers) and waits for
  // it; calls from several threads take turns.  A sample the renderer drops
  // (NaN, negative) comes back as 0, which the reference's loop skips.
  virtual V integrate(const Ray& ray, const Scene& scene, Sampler& sampler) const = 0;
  // every pixel sample of the reference's render loop, averaged
  virtual Image render(const Scene& scene, const Camera& camera, int width, int height,
                       int n_samples, uint64_t seed = 0) const = 0;

 protected:
  static pmx_render_params params(int width, int height, int n_samples, int max_depth,
                                  uint64_t seed) {
    pmx_render_params p{};
    p.width = width;
    p.height = height;
    p.spp = n_samples;
    p.max_depth = max_depth;
    p.seed = seed;
    p.out_mem = PM_MEM_HOST;
    return p;
  }

  static V integrate_one(const Scene& scene, int integrator, pm_photonmap* global,
                         pm_photonmap* caustics, pmx_render_params p, const Ray& ray,
                         Sampler& sampler) {
    uint32_t w[2];
    for (uint32_t& x : w) {
      const float f = sampler.getNext1D();
      std::memcpy(&x, &f, sizeof(x));
    }
    p.seed = (static_cast<uint64_t>(w[0]) << 32) | w[1];
    p.out_mem = PM_MEM_HOST;
    const float r[6] = {ray.origin[0],    ray.origin[1],    ray.origin[2],
                        ray.direction[0], ray.direction[1], ray.direction[2]};
    float rgb[3] = {0.f, 0.f, 0.f};
    {
      std::lock_guard<std::mutex> lock(detail::device_mutex());
      check(pmx_integrate_rays(scene.handle(), integrator, global, caustics, &p, 1, r, rgb),
            "pmx_integrate_rays");
    }
    V out;
    for (int i = 0; i < 3; ++i) out[i] = rgb[i];
    return out;
  }
};

// PathTracing (include/integrator.h:49-101).
template <typename V, typename V2 = PM_VEC2F_TYPE>
class BasicPathTracing : public BasicIntegrator<V, V2> {
 public:
  using typename BasicIntegrator<V, V2>::Image;
  using typename BasicIntegrator<V, V2>::Camera;
  using typename BasicIntegrator<V, V2>::Ray;
  using typename BasicIntegrator<V, V2>::Sampler;
  explicit BasicPathTracing(int maxDepth = 100) : maxDepth(maxDepth) {}
  void build(const Scene&, Sampler&) override {}
  V integrate(const Ray& ray, const Scene& scene, Sampler& sampler) const override {
    return this->integrate_one(scene, PMX_INTEGRATOR_PATH_TRACING, nullptr, nullptr,
                               this->params(1, 1, 1, maxDepth, 0), ray, sampler);
  }
  Image render(const Scene& scene, const Camera& camera, int width, int height, int n_samples,
               uint64_t seed = 0) const override {
    Image img(width, height);
    const pmx_render_params p = this->params(width, height, n_samples, maxDepth, seed);
    std::lock_guard<std::mutex> lock(detail::device_mutex());
    check(pmx_render_path_tracing(scene.handle(), &camera.desc(), &p, img.data().data(), nullptr,
                                  nullptr),
          "pmx_render_path_tracing");
    return img;
  }

 private:
  const int maxDepth;
};

// PhotonMapping (include/integrator.h:104-563).
template <typename V, typename V2 = PM_VEC2F_TYPE>
class BasicPhotonMapping : public BasicIntegrator<V, V2> {
 public:
  using typename BasicIntegrator<V, V2>::Image;
  using typename BasicIntegrator<V, V2>::Camera;
  using typename BasicIntegrator<V, V2>::Ray;
  using typename BasicIntegrator<V, V2>::Sampler;
  BasicPhotonMapping(int nPhotonsGlobal, int nEstimationGlobal, float nPhotonsCausticsMultiplier,
                     int nEstimationCaustics, int strictCalcDepth, int maxDepth)
      : nPhotonsGlobal(nPhotonsGlobal),
        nEstimationGlobal(nEstimationGlobal),
        nPhotonsCaustics(static_cast<int>(nPhotonsGlobal * nPhotonsCausticsMultiplier)),
        nEstimationCaustics(nEstimationCaustics),
        finalGatheringDepth(strictCalcDepth),
        maxDepth(maxDepth) {}
  BasicPhotonMapping(const BasicPhotonMapping&) = delete;
  BasicPhotonMapping& operator=(const BasicPhotonMapping&) = delete;
  ~BasicPhotonMapping() override { release(); }

  // photon tracing of both maps and their builds (include/integrator.h:393-563)
  void build(const Scene& scene, Sampler& sampler) override {
    release();
    std::lock_guard<std::mutex> lock(detail::device_mutex());
    check(pmx_trace_photon_map(scene.handle(), PMX_PHOTONS_GLOBAL, nPhotonsGlobal, maxDepth,
                               sampler.getSeed(), &global, &nGlobalStored),
          "pmx_trace_photon_map");
    check(pm_photonmap_build(global, 0, nullptr), "pm_photonmap_build");
    if (finalGatheringDepth > 0) {
      check(pmx_trace_photon_map(scene.handle(), PMX_PHOTONS_CAUSTICS, nPhotonsCaustics, maxDepth,
                                 sampler.getSeed(), &caustics, &nCausticsStored),
            "pmx_trace_photon_map");
      check(pm_photonmap_build(caustics, 0, nullptr), "pm_photonmap_build");
    }
  }

  // PhotonMapping::integrate (include/integrator.h:566-569)
  V integrate(const Ray& ray, const Scene& scene, Sampler& sampler) const override {
    if (!global) throw Error(PM_ERR_NOT_BUILT, "PhotonMapping used before build()");
    return this->integrate_one(scene, PMX_INTEGRATOR_PHOTON_MAPPING, global, caustics,
                               frameParams(1, 1, 1, 0), ray, sampler);
  }

  Image render(const Scene& scene, const Camera& camera, int width, int height, int n_samples,
               uint64_t seed = 0) const override {
    if (!global) throw Error(PM_ERR_NOT_BUILT, "PhotonMapping used before build()");
    Image img(width, height);
    const pmx_render_params p = frameParams(width, height, n_samples, seed);
    std::lock_guard<std::mutex> lock(detail::device_mutex());
    check(pmx_render_photon_mapping(scene.handle(), &camera.desc(), global, caustics, &p,
                                    img.data().data(), nullptr, nullptr),
          "pmx_render_photon_mapping");
    return img;
  }

  // the render parameters of this integrator (what render() passes to
  // pmx_render_photon_mapping), for callers of the C ABI
  pmx_render_params frameParams(int width, int height, int n_samples, uint64_t seed) const {
    pmx_render_params p = this->params(width, height, n_samples, maxDepth, seed);
    p.final_gathering_depth = finalGatheringDepth;
    p.k = nEstimationGlobal;  // both maps, as include/integrator.h:139, 161 do
    p.n_emit_global = nPhotonsGlobal;
    p.n_emit_caustics = nPhotonsCaustics;
    return p;
  }

  // the device photon maps (the reference's getPhotonMapPtr returns the global one)
  pm_photonmap* getPhotonMapPtr() const { return global; }
  pm_photonmap* getCausticsPhotonMapPtr() const { return caustics; }
  int64_t nGlobalPhotonsStored() const { return nGlobalStored; }
  int64_t nCausticsPhotonsStored() const { return nCausticsStored; }

 private:
  void release() {
    if (global) pm_photonmap_destroy(global);
    if (caustics) pm_photonmap_destroy(caustics);
    global = caustics = nullptr;
  }

  const int nPhotonsGlobal;
  const int nEstimationGlobal;
  const int nPhotonsCaustics;
  const int nEstimationCaustics;  // read by nothing, as in the reference
  const int finalGatheringDepth;
  const int maxDepth;
  pm_photonmap* global = nullptr;
  pm_photonmap* caustics = nullptr;
  int64_t nGlobalStored = 0, nCausticsStored = 0;
};

}  // namespace pm

#ifndef PM_NO_GLOBAL_NAMES
#ifdef PM_VEC2F_IS_DEFAULT
using Vec2f = pm::Vec2f;
#endif
#ifdef PM_VEC3F_IS_DEFAULT
using Vec3f = pm::Vec3f;
constexpr float PI = pm::kPi;  // include/core.h:7
#endif
using Ray = pm::BasicRay<PM_VEC3F_TYPE>;
using IntersectInfo = pm::BasicIntersectInfo<PM_VEC3F_TYPE>;
using SurfaceInfo = pm::BasicSurfaceInfo<PM_VEC3F_TYPE, PM_VEC2F_TYPE>;
using AreaLight = pm::BasicAreaLight<PM_VEC3F_TYPE, PM_VEC2F_TYPE>;
using Camera = pm::BasicCamera<PM_VEC3F_TYPE>;
using Scene = pm::Scene;
using Image = pm::BasicImage<PM_VEC3F_TYPE>;
using Sampler = pm::BasicSampler<PM_VEC2F_TYPE>;
using UniformSampler = pm::BasicUniformSampler<PM_VEC2F_TYPE>;
using Integrator = pm::BasicIntegrator<PM_VEC3F_TYPE, PM_VEC2F_TYPE>;
using PathTracing = pm::BasicPathTracing<PM_VEC3F_TYPE, PM_VEC2F_TYPE>;
using PhotonMapping = pm::BasicPhotonMapping<PM_VEC3F_TYPE, PM_VEC2F_TYPE>;
#endif

#endif  // PM_RENDER_HPP
