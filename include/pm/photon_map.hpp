// pm/photon_map.hpp -- header-only C++ drop-in for the reference's PhotonMap
// (include/photon_map.h:10-22, 186-212), backed by libpm_gather.so.
//
// Same class names, method names, argument meaning and results:
//   Photon{throughput, position, wi}
//   PhotonMap::getNPhotons / getIthPhoton / addPhoton / setPhotons / build /
//              queryKNearestPhotons(p, k, max_dist2)
// queryKNearestPhotons returns photon indices farthest first and the k-th
// smallest squared distance, exactly as the reference's kd-tree does (ties on
// d^2 are ordered by photon index).  The search runs on the GPU; the
// single-query call stages through host memory and synchronises, so batch
// users call queryKNearestPhotonsBatch / gatherRadianceBatch instead.
//
// The vector type defaults to pm::Vec3f (three floats, operator[]); code that
// already has the reference's Vec3f defines PM_VEC3F_TYPE to it before
// including this header.  Errors are reported by throwing pm::Error (the
// reference has no error channel; it never fails on valid input).
#ifndef PM_PHOTON_MAP_HPP
#define PM_PHOTON_MAP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm_gather.h"

namespace pm {

struct Vec3f {
  float v[3];
  Vec3f() : v{0.f, 0.f, 0.f} {}
  Vec3f(float x) : v{x, x, x} {}
  Vec3f(float x, float y, float z) : v{x, y, z} {}
  float operator[](int i) const { return v[i]; }
  float& operator[](int i) { return v[i]; }
  // component-wise arithmetic of the reference's Vec3 (include/core.h:114-218)
  Vec3f operator+(const Vec3f& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
  Vec3f operator-(const Vec3f& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
  Vec3f operator*(const Vec3f& o) const { return {v[0] * o.v[0], v[1] * o.v[1], v[2] * o.v[2]}; }
  Vec3f operator*(float s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  Vec3f operator/(float s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  Vec3f& operator+=(const Vec3f& o) {
    for (int i = 0; i < 3; ++i) v[i] += o.v[i];
    return *this;
  }
};
inline Vec3f operator*(float s, const Vec3f& a) { return a * s; }

// Two floats with operator[] (the reference's Vec2f, include/core.h:25-112).
struct Vec2f {
  float v[2];
  Vec2f() : v{0.f, 0.f} {}
  Vec2f(float x) : v{x, x} {}
  Vec2f(float x, float y) : v{x, y} {}
  float operator[](int i) const { return v[i]; }
  float& operator[](int i) { return v[i]; }
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& where)
      : std::runtime_error(where + ": " + pm_last_error()), code(c) {}
};

inline void check(int rc, const char* where) {
  if (rc != PM_OK) throw Error(rc, where);
}

// One context per process and device, created on first use.
inline pm_context* default_context(int device = 0) {
  static pm_context* ctx = [device] {
    pm_context* c = nullptr;
    check(pm_context_create(device, &c), "pm_context_create");
    return c;
  }();
  return ctx;
}

template <typename V>
struct BasicPhoton {
  V throughput;  // flux
  V position;
  V wi;          // incident direction
  static constexpr int dim = 3;
  float operator[](int i) const { return position[i]; }
  BasicPhoton() {}
  BasicPhoton(const V& throughput, const V& position, const V& wi)
      : throughput(throughput), position(position), wi(wi) {}
};

template <typename V>
class BasicPhotonMap {
 public:
  using PhotonT = BasicPhoton<V>;

  BasicPhotonMap() = default;
  BasicPhotonMap(const BasicPhotonMap&) = delete;
  BasicPhotonMap& operator=(const BasicPhotonMap&) = delete;
  ~BasicPhotonMap() { release(); }

  int getNPhotons() const { return static_cast<int>(photons.size()); }
  const PhotonT& getIthPhoton(int i) const { return photons[i]; }

  // Like the reference, adding photons invalidates a built map until the
  // next build().
  void addPhoton(const PhotonT& photon) {
    photons.push_back(photon);
    release();
  }
  void setPhotons(const std::vector<PhotonT>& ph) {
    photons = ph;
    release();
  }

  void build(int leaf_size = 0) {
    release();
    std::vector<float> aos(9 * photons.size());
    for (size_t i = 0; i < photons.size(); ++i) {
      for (int c = 0; c < 3; ++c) {
        aos[9 * i + c] = photons[i].throughput[c];
        aos[9 * i + 3 + c] = photons[i].position[c];
        aos[9 * i + 6 + c] = photons[i].wi[c];
      }
    }
    check(pm_photonmap_create_aos(default_context(), static_cast<int64_t>(photons.size()),
                                  aos.data(), PM_MEM_HOST, &map),
          "pm_photonmap_create_aos");
    check(pm_photonmap_build(map, leaf_size, nullptr), "pm_photonmap_build");
  }

  std::vector<int> queryKNearestPhotons(const V& p, int k, float& max_dist2) const {
    const float q[3] = {p[0], p[1], p[2]};
    std::vector<int32_t> idx(static_cast<size_t>(k));
    check(pm_knn(built(), 1, q, k, idx.data(), &max_dist2, PM_MEM_HOST, nullptr), "pm_knn");
    std::vector<int> out;
    for (int32_t i : idx)
      if (i >= 0) out.push_back(i);
    return out;
  }

  // Batched kNN over host points: rows of k indices (farthest first, -1
  // padded) and the k-th squared distance of each point.
  void queryKNearestPhotonsBatch(const std::vector<V>& points, int k, std::vector<int32_t>& idx,
                                 std::vector<float>& max_dist2) const {
    const size_t m = points.size();
    std::vector<float> q(3 * m);
    for (size_t i = 0; i < m; ++i)
      for (int c = 0; c < 3; ++c) q[c * m + i] = points[i][c];
    idx.assign(m * k, -1);
    max_dist2.assign(m, 0.f);
    check(pm_knn(built(), static_cast<int64_t>(m), q.data(), k, idx.data(), max_dist2.data(),
                 PM_MEM_HOST, nullptr),
          "pm_knn");
  }

  // Batched PhotonMapping::computeRadianceWithPhotonMap
  // (include/integrator.h:133-152) over host query records.
  std::vector<float> gatherRadianceBatch(const std::vector<pm_query>& queries,
                                         const std::vector<pm_material>& materials, int k,
                                         double n_emit, std::vector<float>* r2 = nullptr) const {
    const size_t m = queries.size();
    std::vector<float> rgb(3 * m);
    if (r2) r2->assign(m, 0.f);
    check(pm_gather_radiance(built(), static_cast<int64_t>(m), queries.data(), materials.data(),
                             static_cast<int>(materials.size()), k, n_emit, rgb.data(),
                             r2 ? r2->data() : nullptr, nullptr, PM_MEM_HOST, nullptr),
          "pm_gather_radiance");
    return rgb;
  }

  pm_photonmap* handle() const { return map; }

 private:
  std::vector<PhotonT> photons;
  pm_photonmap* map = nullptr;

  pm_photonmap* built() const {
    if (!map) throw Error(PM_ERR_NOT_BUILT, "PhotonMap used before build()");
    return map;
  }
  void release() {
    if (map) pm_photonmap_destroy(map);
    map = nullptr;
  }
};

}  // namespace pm

#ifndef PM_VEC2F_TYPE
#define PM_VEC2F_TYPE pm::Vec2f
#define PM_VEC2F_IS_DEFAULT 1
#endif
#ifndef PM_VEC3F_TYPE
#define PM_VEC3F_TYPE pm::Vec3f
#define PM_VEC3F_IS_DEFAULT 1
#endif

#ifndef PM_NO_GLOBAL_NAMES
using Photon = pm::BasicPhoton<PM_VEC3F_TYPE>;
using PhotonMap = pm::BasicPhotonMap<PM_VEC3F_TYPE>;
#endif

#endif  // PM_PHOTON_MAP_HPP
