// Two-wide bounding-volume hierarchy over scene triangles: the layout shared
// by the host builder (pm_bvh.hip) and the device traversal (pm_scene.hip).
// It replaces the reference's Embree scene (include/scene.h:275-305).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pm_scene.h"

namespace pm {

// Leaves hold at most this many triangles.
constexpr int kBvhLeafMax = 4;
// Traversal stack entries per ray.  The builder bounds the internal-node
// depth to kBvhStack - 2, so a stack of this size never overflows.
constexpr int kBvhStack = 40;

// One internal node (64 bytes): the boxes of both children, axis by axis,
// then the two child references.  A reference >= 0 is an internal node; a
// reference < 0 is a leaf, ~ref = (first triangle << 4) | count, count 0..15
// (0: an empty child, used only under a root with a single leaf).
struct alignas(16) BvhNode {
  float4 bx;  // (left lo.x, left hi.x, right lo.x, right hi.x)
  float4 by;
  float4 bz;
  int4 ref;   // (left, right, unused, unused)
};
static_assert(sizeof(BvhNode) == 64, "node layout");

struct BvhHost {
  std::vector<BvhNode> nodes;
  // per triangle in leaf order, three float4: (v0, original index as int
  // bits), (v1 - v0, 0), (v2 - v0, 0), the operands of the hit test
  std::vector<float4> tri4;
  int depth = 0;       // deepest internal node (root = 0)
  int leaves = 0;
  double sah_cost = 0.0;  // expected node + triangle tests per ray (surface-area heuristic)
};

// Binned-SAH build over n >= 1 triangles.  Boxes are padded by 1e-5 of the
// scene's largest extent, so that every hit the triangle test reports lies
// inside the boxes on its path.  Returns PM_OK or an error code.
int build_bvh(const pmx_triangle* tris, int n, BvhHost* out);

}  // namespace pm
