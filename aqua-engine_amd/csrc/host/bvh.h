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

// 4-wide node, 128 bytes: the boxes of up to four children as six float4
// rows (lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4]) so one node
// fetch feeds four slab tests, then child[4] and count[4] with the same
// meaning as in BvhNode (count > 0 leaf, count == 0 inner, child < 0 empty).
struct Bvh4Node {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int32_t child[4];
  int32_t count[4];
};
static_assert(sizeof(Bvh4Node) == 128, "node layout");

struct Bvh4 {
  std::vector<Bvh4Node> nodes;  // node 0 is the root; top levels first, then preorder
  std::vector<uint32_t> tri_order;  // triangle slot -> triangle id (load order)
  uint32_t max_inner_depth = 0; // deepest inner node (root = 0)
};

// Device node format: 64 bytes = 16 words per 4-wide node.
//   w0..2  origin (float): the low corner of the union of the child boxes
//   w3     per-axis power-of-two scales as signed exponent bytes (bits 0-7
//          x, 8-15 y, 16-23 z, two's complement): plane = origin + q * 2^e
//   w4..9  child planes as bytes (byte k = child k): lo.x, hi.x, lo.y, hi.y,
//          lo.z, hi.z, each rounded outward (lo down, hi up)
//   w10..13 child references: an inner child is its node index (< 2^31); a
//          leaf child is kLeafFlag | the mask of its triangle slots over the
//          node's leaf run (bit i = slot w14 + i); an empty child is kNoChild
//   w14    first triangle slot of the node's leaf children (one run of
//          slots; 0 when there are none)
//   w15    0
// Decoded boxes enclose the original ones, so traversal stays conservative.
constexpr uint32_t kNodeWords = 16;
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kLeafMaskBits = 31;
// An empty child: a leaf of no triangles (flag, empty mask), so traversal
// treats it like any culled leaf and needs no test of its own (its planes
// are all 0).
constexpr uint32_t kNoChild = kLeafFlag;
// false when a node's leaf children do not form one run that fits w15.
bool quantize_bvh4(const Bvh4& in, std::vector<uint32_t>* out);

// Collapses a binary tree into 4-wide nodes with the least SAH cost (a
// dynamic program over the binary tree choosing, for every wide node, which
// binary descendants become its up to four children and which subtrees of
// at most max_leaf triangles become single leaves). Triangle slots are laid
// out anew: the leaf children of every 4-wide node take one contiguous run of
// slots, in child order, so traversal can test a node's hit leaves as one
// bit mask over that run.
void collapse_bvh4(const Bvh& bin, Bvh4* out, int max_leaf = 4);

// Node order: the nodes of depth <= kTopDepth (at most 1 + 4 + 16 + 64 = 85)
// come first, breadth-first, so the traversal kernels can stage a prefix of
// the node array in LDS; the rest keep their preorder. A child's index is
// always larger than its parent's. collapse_bvh4 applies it.
constexpr int kTopDepth = 3;
void order_top_levels(Bvh4* b, int depth);

}  // namespace aqua
