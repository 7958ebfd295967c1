// Host build of the two-wide triangle BVH (layout in pm_bvh.hpp).
//
// Binned surface-area heuristic: at every node the triangle centroids are
// binned in 16 slabs along each axis and the split with the lowest
// (left area x left count + right area x right count) is taken.  A node
// switches to a median split when the heuristic's split could push the
// tree deeper than the device traversal stack allows; ranges of at most
// kBvhLeafMax triangles become leaves.  Triangle order within the tree is
// the order of the index array after the in-place partitions, so every leaf
// is a contiguous range of the packed triangle array.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "pm_bvh.hpp"
#include "pm_internal.hpp"

namespace pm {
namespace {

constexpr int kSahBins = 16;
constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
  float lo[3] = {kInf, kInf, kInf};
  float hi[3] = {-kInf, -kInf, -kInf};
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  bool empty() const { return lo[0] > hi[0]; }
  double area() const {
    if (empty()) return 0.0;
    const double x = double(hi[0]) - lo[0], y = double(hi[1]) - lo[1], z = double(hi[2]) - lo[2];
    return 2.0 * (x * y + y * z + z * x);
  }
};

// Internal levels a median-split subtree over c triangles needs.
int median_levels(int c) {
  int l = 0;
  while (c > kBvhLeafMax) {
    c = (c + 1) / 2;
    ++l;
  }
  return l;
}

struct Builder {
  const pmx_triangle* tris;
  int n;
  float pad;
  std::vector<Box> tbox;
  std::vector<float> cen;  // 3 per triangle
  std::vector<int> idx;

  Box range_box(int b, int e) const {
    Box r;
    for (int i = b; i < e; ++i) r.grow(tbox[idx[i]]);
    for (int a = 0; a < 3; ++a) {
      r.lo[a] -= pad;
      r.hi[a] += pad;
    }
    return r;
  }

  int bin_of(int tri, int axis, float lo, float scale) const {
    const int k = static_cast<int>((cen[3 * tri + axis] - lo) * scale);
    return std::min(std::max(k, 0), kSahBins - 1);
  }

  // Split of [b, e): the surface-area heuristic, unless `median` (or no
  // useful heuristic split exists): then halves by centroid along the
  // widest centroid axis.  Returns the first index of the right part.
  int split(int b, int e, bool median) {
    const int c = e - b;
    Box cb;
    for (int i = b; i < e; ++i) cb.grow(&cen[3 * idx[i]]);
    int wide = 0;
    for (int a = 1; a < 3; ++a)
      if (cb.hi[a] - cb.lo[a] > cb.hi[wide] - cb.lo[wide]) wide = a;
    if (!(cb.hi[wide] > cb.lo[wide])) return b + c / 2;  // coincident centroids
    if (!median) {
      double best = std::numeric_limits<double>::infinity();
      int best_axis = -1, best_bin = 0;
      for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.0f)) continue;
        const float scale = kSahBins / ext;
        Box bb[kSahBins];
        int cnt[kSahBins] = {};
        for (int i = b; i < e; ++i) {
          const int k = bin_of(idx[i], a, cb.lo[a], scale);
          bb[k].grow(tbox[idx[i]]);
          ++cnt[k];
        }
        double right_area[kSahBins];
        int right_cnt[kSahBins];
        Box acc;
        int nacc = 0;
        for (int k = kSahBins - 1; k >= 1; --k) {
          acc.grow(bb[k]);
          nacc += cnt[k];
          right_area[k] = acc.area();
          right_cnt[k] = nacc;
        }
        Box left;
        int nl = 0;
        for (int s = 1; s < kSahBins; ++s) {
          left.grow(bb[s - 1]);
          nl += cnt[s - 1];
          const int nr = right_cnt[s];
          if (nl == 0 || nr == 0) continue;
          const double cost = left.area() * nl + right_area[s] * nr;
          if (cost < best) {
            best = cost;
            best_axis = a;
            best_bin = s;
          }
        }
      }
      if (best_axis >= 0) {
        const float scale = kSahBins / (cb.hi[best_axis] - cb.lo[best_axis]);
        const float lo = cb.lo[best_axis];
        auto* mid = std::partition(idx.data() + b, idx.data() + e, [&](int t) {
          return bin_of(t, best_axis, lo, scale) < best_bin;
        });
        const int m = static_cast<int>(mid - idx.data());
        if (m > b && m < e) return m;
      }
    }
    const int m = b + c / 2;
    std::nth_element(idx.data() + b, idx.data() + m, idx.data() + e, [&](int x, int y) {
      const float cx = cen[3 * x + wide], cy = cen[3 * y + wide];
      return cx < cy || (cx == cy && x < y);
    });
    return m;
  }
};

void set_child(BvhNode& nd, int side, const Box& bx, int ref) {
  float* x = &nd.bx.x;
  float* y = &nd.by.x;
  float* z = &nd.bz.x;
  x[2 * side] = bx.lo[0];
  x[2 * side + 1] = bx.hi[0];
  y[2 * side] = bx.lo[1];
  y[2 * side + 1] = bx.hi[1];
  z[2 * side] = bx.lo[2];
  z[2 * side + 1] = bx.hi[2];
  (side == 0 ? nd.ref.x : nd.ref.y) = ref;
}

int leaf_ref(int first, int count) { return ~((first << 4) | count); }

}  // namespace

int build_bvh(const pmx_triangle* tris, int n, BvhHost* out) {
  if (!tris || !out || n < 1) return fail(PM_ERR_INVALID_ARG, "BVH needs at least one triangle");
  if (n >= (1 << 27)) return fail(PM_ERR_INVALID_ARG, "too many triangles for the BVH (2^27)");
  Builder B;
  B.tris = tris;
  B.n = n;
  B.tbox.resize(n);
  B.cen.resize(3 * static_cast<size_t>(n));
  B.idx.resize(n);
  std::iota(B.idx.begin(), B.idx.end(), 0);
  Box scene;
  for (int i = 0; i < n; ++i) {
    for (int v = 0; v < 3; ++v) {
      for (int a = 0; a < 3; ++a)
        if (!std::isfinite(tris[i].v[v][a]))
          return fail(PM_ERR_INVALID_ARG, "triangle vertices must be finite");
      B.tbox[i].grow(tris[i].v[v]);
    }
    for (int a = 0; a < 3; ++a) B.cen[3 * i + a] = 0.5f * (B.tbox[i].lo[a] + B.tbox[i].hi[a]);
    scene.grow(B.tbox[i]);
  }
  float extent = 0.0f;
  for (int a = 0; a < 3; ++a) extent = std::max(extent, scene.hi[a] - scene.lo[a]);
  B.pad = std::max(1e-5f * extent, 1e-30f);

  std::vector<BvhNode>& nodes = out->nodes;
  nodes.clear();
  nodes.reserve(2 * static_cast<size_t>(n) / kBvhLeafMax + 2);
  nodes.emplace_back();
  out->depth = 0;
  out->leaves = 0;
  if (n <= kBvhLeafMax) {
    const Box bx = B.range_box(0, n);
    set_child(nodes[0], 0, bx, leaf_ref(0, n));
    set_child(nodes[0], 1, bx, leaf_ref(0, 0));
    out->leaves = 1;
  } else {
    struct Task {
      int b, e, depth, node;
    };
    std::vector<Task> todo{{0, n, 0, 0}};
    while (!todo.empty()) {
      const Task t = todo.back();
      todo.pop_back();
      out->depth = std::max(out->depth, t.depth);
      // children sit at depth + 1 and must keep their subtrees within the stack
      const auto fits = [&](int cc) {
        return t.depth + 1 + median_levels(cc) - 1 <= kBvhStack - 2;
      };
      int m = B.split(t.b, t.e, false);
      if (!fits(m - t.b) || !fits(t.e - m)) m = B.split(t.b, t.e, true);
      const int range[2][2] = {{t.b, m}, {m, t.e}};
      for (int side = 1; side >= 0; --side) {
        const int b = range[side][0], e = range[side][1];
        const Box bx = B.range_box(b, e);
        int ref;
        if (e - b <= kBvhLeafMax) {
          ref = leaf_ref(b, e - b);
          ++out->leaves;
        } else {
          ref = static_cast<int>(nodes.size());
          nodes.emplace_back();
          todo.push_back({b, e, t.depth + 1, ref});
        }
        set_child(nodes[t.node], side, bx, ref);
      }
    }
  }

  // expected tests per ray under the surface-area heuristic
  {
    Box root;
    Box l, r;
    const BvhNode& nd = nodes[0];
    l.lo[0] = nd.bx.x; l.hi[0] = nd.bx.y; l.lo[1] = nd.by.x; l.hi[1] = nd.by.y;
    l.lo[2] = nd.bz.x; l.hi[2] = nd.bz.y;
    r.lo[0] = nd.bx.z; r.hi[0] = nd.bx.w; r.lo[1] = nd.by.z; r.hi[1] = nd.by.w;
    r.lo[2] = nd.bz.z; r.hi[2] = nd.bz.w;
    root.grow(l);
    root.grow(r);
    const double a_root = std::max(root.area(), 1e-300);
    double cost = 1.0;
    for (const BvhNode& q : nodes) {
      for (int side = 0; side < 2; ++side) {
        Box b;
        const float* x = &q.bx.x;
        const float* y = &q.by.x;
        const float* z = &q.bz.x;
        b.lo[0] = x[2 * side]; b.hi[0] = x[2 * side + 1];
        b.lo[1] = y[2 * side]; b.hi[1] = y[2 * side + 1];
        b.lo[2] = z[2 * side]; b.hi[2] = z[2 * side + 1];
        const int ref = side == 0 ? q.ref.x : q.ref.y;
        const double w = b.area() / a_root;
        cost += ref >= 0 ? w : w * ((~ref) & 15);
      }
    }
    out->sah_cost = cost;
  }

  // packed hit-test operands in leaf order
  out->tri4.resize(3 * static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    const pmx_triangle& t = tris[B.idx[i]];
    const float* v0 = t.v[0];
    const float* v1 = t.v[1];
    const float* v2 = t.v[2];
    int orig = B.idx[i];
    float orig_bits;
    static_assert(sizeof(orig_bits) == sizeof(orig), "bit cast");
    std::memcpy(&orig_bits, &orig, sizeof(orig));
    out->tri4[3 * i + 0] = make_float4(v0[0], v0[1], v0[2], orig_bits);
    out->tri4[3 * i + 1] = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], 0.0f);
    out->tri4[3 * i + 2] = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.0f);
  }
  return PM_OK;
}

}  // namespace pm
