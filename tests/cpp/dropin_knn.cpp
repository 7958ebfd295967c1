// Uses the header-only drop-in exactly like code written against the
// reference's PhotonMap (include/photon_map.h:186-212): fill photons, build,
// query.  Prints one line per query: r^2 (as hex bits) and the indices.
// Input on stdin: n, then n lines "px py pz wx wy wz fr fg fb"; m, then m
// lines "qx qy qz"; then k.
#include <cstdio>
#include <cstring>
#include <vector>

#include "pm/photon_map.hpp"

int main() {
  int n = 0;
  if (std::scanf("%d", &n) != 1) return 1;
  std::vector<Photon> photons;
  for (int i = 0; i < n; ++i) {
    float p[9];
    for (float& v : p)
      if (std::scanf("%a", &v) != 1) return 1;
    photons.emplace_back(pm::Vec3f(p[6], p[7], p[8]), pm::Vec3f(p[0], p[1], p[2]),
                         pm::Vec3f(p[3], p[4], p[5]));
  }
  int m = 0;
  if (std::scanf("%d", &m) != 1) return 1;
  std::vector<pm::Vec3f> q(m);
  for (auto& v : q)
    for (int c = 0; c < 3; ++c)
      if (std::scanf("%a", &v[c]) != 1) return 1;
  int k = 0;
  if (std::scanf("%d", &k) != 1) return 1;

  PhotonMap map;
  map.setPhotons(photons);
  map.build();
  if (map.getNPhotons() != n) return 2;
  for (int i = 0; i < m; ++i) {
    float r2 = -1.f;
    const std::vector<int> idx = map.queryKNearestPhotons(q[i], k, r2);
    unsigned bits;
    std::memcpy(&bits, &r2, 4);
    std::printf("%08x", bits);
    for (int j : idx) std::printf(" %d", j);
    std::printf("\n");
  }
  // the same through the batched call
  std::vector<int32_t> bidx;
  std::vector<float> br2;
  map.queryKNearestPhotonsBatch(q, k, bidx, br2);
  for (int i = 0; i < m; ++i) {
    unsigned bits;
    std::memcpy(&bits, &br2[i], 4);
    std::printf("B%08x", bits);
    for (int j = 0; j < k; ++j)
      if (bidx[static_cast<size_t>(i) * k + j] >= 0) std::printf(" %d", bidx[static_cast<size_t>(i) * k + j]);
    std::printf("\n");
  }
  return 0;
}
