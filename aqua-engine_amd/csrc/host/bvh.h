// Binned-SAH BVH builder (Wald 2007, "On fast Construction of SAH-based
// Bounding Volume Hierarchies": 32 bins over the centroid bounds, split on
// the plane of least A_L*N_L + A_R*N_R) emitting the device node layout.
#pragma once

#include <cstdint>
#include <vector>

namespace aqua {

// Device node, 64 bytes, one per inner node; it holds the boxes of both
// children so one fetch decides which children to visit:
//   f[0..3]  = left  lo.x, hi.x, lo.y, hi.y
//   f[4..7]  = right lo.x, hi.x, lo.y, hi.y
//   f[8..11] = left lo.z, hi.z, right lo.z, hi.z
//   i[12..15]= left idx, right idx, left count, right count
// count > 0: leaf of `count` triangles starting at triangle slot idx;
// count == 0 and idx >= 0: inner node idx; idx < 0: empty child.
struct BvhNode {
  float f[12];
  int32_t child[2];
  int32_t count[2];
};
static_assert(sizeof(BvhNode) == 64, "node layout");

struct BvhBuildOpts {
  int bins = 32;
  int max_leaf = 4;
  float traversal_cost = 1.0f;  // relative to one triangle test
};

struct Bvh {
  std::vector<BvhNode> nodes;       // node 0 is the root
  std::vector<uint32_t> tri_order;  // triangle slot -> triangle id (load order)
  uint32_t max_depth = 0;
  double sah_cost = 0.0;            // expected cost per ray, relative units
};

// verts: 9 floats per triangle.
void build_bvh(const float* verts, uint32_t n_tris, const BvhBuildOpts& opts, Bvh* out);

}  // namespace aqua
