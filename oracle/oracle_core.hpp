// Shared pieces of the CPU oracle (TEST INFRASTRUCTURE ONLY): the FP32
// distance and dot product in the reference's operation order, the
// reference kd-tree (include/photon_map.h:61-183) and the FP32 radiance
// estimate (include/integrator.h:133-152).  Included by pm_oracle.cpp and
// render_oracle.cpp; built into libpm_oracle.so only.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// The reference's PI is a float constant (include/core.h:7).
constexpr float kPiF = 3.14159265359f;

// Query record, 40 bytes: pos[3], shading normal[3], wo[3], material id.
struct QueryRec {
  float pos[3];
  float ns[3];
  float wo[3];
  int32_t material;
};
static_assert(sizeof(QueryRec) == 40, "query record layout");

// Material record, 16 bytes: albedo rgb and a type (0 = Lambert).
struct MaterialRec {
  float rho[3];
  int32_t type;
};
static_assert(sizeof(MaterialRec) == 16, "material record layout");

// FP32 squared distance in the reference's operation order: the accumulator
// starts at 0 and adds (q[i] - p[i])^2 for i = 0, 1, 2 (photon_map.h:36-39).
inline float dist2_f32(float qx, float qy, float qz, float px, float py,
                       float pz) {
  float acc = 0.0f;
  const float dx = qx - px;
  acc += dx * dx;
  const float dy = qy - py;
  acc += dy * dy;
  const float dz = qz - pz;
  acc += dz * dz;
  return acc;
}

// FP32 dot product, left to right (core.h:206-209).
inline float dot3_f32(const float* a, const float* b) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

// Lambert visibility predicate: the BxDF is rho/pi unless either local
// direction has a negative y (material.h:83-85).  NaN compares false, so a
// NaN cosine does not zero the term, exactly as in the reference.
inline bool lambert_passes(float cos_o, float cos_i) {
  return !(cos_o < 0.0f || cos_i < 0.0f);
}

// ---------------------------------------------------------------------------
// kd-tree restatement.
// One node per photon in pre-order; a node stores the split axis, the photon
// it holds and the node indices of its two subtrees (-1 when empty).
struct KdNode {
  int axis;
  int idx;
  int left;
  int right;
};

struct KdTreeOracle {
  std::vector<KdNode> nodes;
  std::vector<float> x, y, z;  // copies of the photon positions
  int n = 0;

  float coord(int idx, int axis) const {
    return axis == 0 ? x[idx] : (axis == 1 ? y[idx] : z[idx]);
  }

  // Pre-order build over indices[0, count): axis cycles with depth, the
  // range is fully sorted on that axis with a strict < comparator, the node
  // takes the element at (count - 1) / 2, the left subtree gets everything
  // before it and the right subtree everything after it.
  void build(int* indices, int count, int depth) {
    if (count <= 0) return;
    const int axis = depth % 3;
    std::sort(indices, indices + count, [&](int a, int b) {
      return coord(a, axis) < coord(b, axis);
    });
    const int median = (count - 1) / 2;
    const int self = static_cast<int>(nodes.size());
    nodes.push_back(KdNode{axis, indices[median], -1, -1});

    const int before_left = static_cast<int>(nodes.size());
    build(indices, median, depth + 1);
    nodes[self].left =
        (static_cast<int>(nodes.size()) == before_left) ? -1 : before_left;

    const int before_right = static_cast<int>(nodes.size());
    build(indices + median + 1, count - median - 1, depth + 1);
    nodes[self].right =
        (static_cast<int>(nodes.size()) == before_right) ? -1 : before_right;
  }

  using Heap = std::priority_queue<std::pair<float, int>>;

  // Depth-first search.  Every visited node's photon is pushed; the heap is
  // cut back to k entries after each push; the near side is the one the
  // query falls in by a strict < on the split axis; the far side is searched
  // when the current heap maximum exceeds the squared plane distance.  As in
  // the reference there is no heap-size check before that far-side test.
  void search(int node_id, float qx, float qy, float qz, int k,
              Heap& heap) const {
    if (node_id < 0 || node_id >= static_cast<int>(nodes.size())) return;
    const KdNode& nd = nodes[node_id];
    const float mx = x[nd.idx], my = y[nd.idx], mz = z[nd.idx];
    heap.emplace(dist2_f32(qx, qy, qz, mx, my, mz), nd.idx);
    if (static_cast<int>(heap.size()) > k) heap.pop();

    const float q_axis = nd.axis == 0 ? qx : (nd.axis == 1 ? qy : qz);
    const float m_axis = nd.axis == 0 ? mx : (nd.axis == 1 ? my : mz);
    const bool go_left = q_axis < m_axis;
    search(go_left ? nd.left : nd.right, qx, qy, qz, k, heap);

    const float plane = m_axis - q_axis;
    if (heap.top().first > plane * plane) {
      search(go_left ? nd.right : nd.left, qx, qy, qz, k, heap);
    }
  }

  // Drains the heap: indices come out largest (d^2, index) first; r^2 is the
  // largest d^2 seen while draining (0 for an empty result).
  int query(float qx, float qy, float qz, int k, int32_t* out_idx,
            float* out_r2) const {
    Heap heap;
    search(0, qx, qy, qz, k, heap);
    const int count = static_cast<int>(heap.size());
    float r2 = 0.0f;
    for (int i = 0; i < count; ++i) {
      const auto& top = heap.top();
      out_idx[i] = top.second;
      r2 = std::max(r2, top.first);
      heap.pop();
    }
    for (int i = count; i < k; ++i) out_idx[i] = -1;
    *out_r2 = r2;
    return count;
  }
};

inline int resolve_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

// FP32 estimate in the reference's order: Lo accumulates f * flux over the
// neighbour list as given (farthest first), then is divided by
// (float(n_emit) * PI) * r^2 when the list is non-empty.
inline void estimate_ref32(const float* wi_soa, const float* flux_soa, int64_t n,
                    const QueryRec& q, const MaterialRec* mats, int n_mats,
                    const int32_t* idx, int count, float r2, float n_emit,
                    float* out) {
  float lo[3] = {0.0f, 0.0f, 0.0f};
  const bool lambert = q.material >= 0 && q.material < n_mats &&
                       mats[q.material].type == 0;
  const float cos_o = dot3_f32(q.wo, q.ns);
  for (int j = 0; j < count; ++j) {
    const int64_t i = idx[j];
    const float wi[3] = {wi_soa[i], wi_soa[n + i], wi_soa[2 * n + i]};
    float f[3] = {0.0f, 0.0f, 0.0f};
    if (lambert && lambert_passes(cos_o, dot3_f32(wi, q.ns))) {
      for (int c = 0; c < 3; ++c) f[c] = mats[q.material].rho[c] / kPiF;
    }
    for (int c = 0; c < 3; ++c) lo[c] += f[c] * flux_soa[c * n + i];
  }
  if (count > 0) {
    const float denom = (n_emit * kPiF) * r2;
    for (int c = 0; c < 3; ++c) lo[c] /= denom;
  }
  for (int c = 0; c < 3; ++c) out[c] = lo[c];
}

}  // namespace orc
