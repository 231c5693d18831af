#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

namespace aqua {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int k = 0; k < 3; ++k) {
      lo[k] = INFINITY;
      hi[k] = -INFINITY;
    }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  void grow(const float p[3]) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  // Half surface area (the SAH only needs ratios).
  double half_area() const {
    if (!(hi[0] >= lo[0])) return 0.0;
    double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    return dx * dy + dy * dz + dz * dx;
  }
};

struct BNode {
  Box box;
  int32_t left = -1, right = -1;  // builder node ids; -1 for a leaf
  uint32_t begin = 0, count = 0;  // leaf triangle range in `order`
  uint32_t depth = 0;
};

struct Builder {
  const BvhBuildOpts& opts;
  std::vector<Box> tri_box;
  std::vector<float> centroid;  // 3 per triangle
  std::vector<uint32_t> order;
  std::vector<BNode> nodes;
  std::atomic<uint32_t> next{0};
  std::atomic<int> threads_live{0};
  unsigned hw = 1;

  explicit Builder(const BvhBuildOpts& o) : opts(o) {}

  uint32_t alloc() { return next.fetch_add(1); }

  // Builds the subtree over order[begin, begin+count) into node `id`.
  void build(uint32_t id, uint32_t begin, uint32_t count, uint32_t depth) {
    BNode& nd = nodes[id];
    nd.depth = depth;
    Box b, cb;
    b.reset();
    cb.reset();
    for (uint32_t i = begin; i < begin + count; ++i) {
      b.grow(tri_box[order[i]]);
      cb.grow(&centroid[3 * size_t(order[i])]);
    }
    nd.box = b;
    auto make_leaf = [&]() {
      nd.left = nd.right = -1;
      nd.begin = begin;
      nd.count = count;
    };
    if (count <= 1) return make_leaf();

    const int B = opts.bins;
    double best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    for (int a = 0; a < 3; ++a) {
      float ext = cb.hi[a] - cb.lo[a];
      if (!(ext > 0.f)) continue;
      float scale = float(B) / ext;
      Box bb[64];
      uint32_t bn[64];
      for (int k = 0; k < B; ++k) {
        bb[k].reset();
        bn[k] = 0;
      }
      for (uint32_t i = begin; i < begin + count; ++i) {
        uint32_t t = order[i];
        int k = int((centroid[3 * size_t(t) + a] - cb.lo[a]) * scale);
        k = k < 0 ? 0 : (k >= B ? B - 1 : k);
        bb[k].grow(tri_box[t]);
        ++bn[k];
      }
      double right_area[64];
      uint32_t right_n[64];
      Box acc;
      acc.reset();
      uint32_t n = 0;
      for (int k = B - 1; k > 0; --k) {
        acc.grow(bb[k]);
        n += bn[k];
        right_area[k] = acc.half_area();
        right_n[k] = n;
      }
      acc.reset();
      n = 0;
      for (int k = 0; k < B - 1; ++k) {
        acc.grow(bb[k]);
        n += bn[k];
        if (n == 0 || right_n[k + 1] == 0) continue;
        double c = acc.half_area() * n + right_area[k + 1] * right_n[k + 1];
        if (c < best_cost) {
          best_cost = c;
          best_axis = a;
          best_split = k;
        }
      }
    }
    double parent_area = b.half_area();
    double leaf_cost = double(count);
    double split_cost = opts.traversal_cost + (parent_area > 0 ? best_cost / parent_area : INFINITY);
    uint32_t mid;
    if (best_axis < 0) {
      // All centroids coincide: no SAH plane separates them.
      if (count <= uint32_t(opts.max_leaf)) return make_leaf();
      mid = begin + count / 2;
    } else {
      if (count <= uint32_t(opts.max_leaf) && leaf_cost <= split_cost) return make_leaf();
      float ext = cb.hi[best_axis] - cb.lo[best_axis];
      float scale = float(B) / ext;
      float lo = cb.lo[best_axis];
      uint32_t* m = std::partition(order.data() + begin, order.data() + begin + count, [&](uint32_t t) {
        int k = int((centroid[3 * size_t(t) + best_axis] - lo) * scale);
        k = k < 0 ? 0 : (k >= B ? B - 1 : k);
        return k <= best_split;
      });
      mid = uint32_t(m - order.data());
      if (mid == begin || mid == begin + count) mid = begin + count / 2;
    }
    uint32_t l = alloc(), r = alloc();
    nd.left = int32_t(l);
    nd.right = int32_t(r);
    uint32_t nl = mid - begin, nr = count - nl;
    // Large subtrees are built on their own thread.
    if (nl > 65536 && nr > 65536 && threads_live.load() < int(hw)) {
      threads_live.fetch_add(1);
      std::thread th([=]() { build(l, begin, nl, depth + 1); });
      build(r, mid, nr, depth + 1);
      th.join();
      threads_live.fetch_sub(1);
    } else {
      build(l, begin, nl, depth + 1);
      build(r, mid, nr, depth + 1);
    }
  }
};

void set_child(BvhNode& dn, int side, const BNode& c, int32_t idx, int32_t count) {
  if (side == 0) {
    dn.f[0] = c.box.lo[0];
    dn.f[1] = c.box.hi[0];
    dn.f[2] = c.box.lo[1];
    dn.f[3] = c.box.hi[1];
    dn.f[8] = c.box.lo[2];
    dn.f[9] = c.box.hi[2];
  } else {
    dn.f[4] = c.box.lo[0];
    dn.f[5] = c.box.hi[0];
    dn.f[6] = c.box.lo[1];
    dn.f[7] = c.box.hi[1];
    dn.f[10] = c.box.lo[2];
    dn.f[11] = c.box.hi[2];
  }
  dn.child[side] = idx;
  dn.count[side] = count;
}

}  // namespace

void build_bvh(const float* verts, uint32_t n, const BvhBuildOpts& opts, Bvh* out) {
  *out = Bvh();
  Builder b(opts);
  b.hw = std::max(1u, std::thread::hardware_concurrency());
  b.tri_box.resize(n);
  b.centroid.resize(3 * size_t(n));
  b.order.resize(n);
  for (uint32_t t = 0; t < n; ++t) {
    Box bx;
    bx.reset();
    for (int v = 0; v < 3; ++v) bx.grow(&verts[9 * size_t(t) + 3 * size_t(v)]);
    b.tri_box[t] = bx;
    for (int k = 0; k < 3; ++k) b.centroid[3 * size_t(t) + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
    b.order[t] = t;
  }
  b.nodes.resize(n > 0 ? 2 * size_t(n) - 1 : 1);
  uint32_t root = b.alloc();
  if (n > 0) b.build(root, 0, n, 0);
  b.nodes.resize(b.next.load());

  // Flatten: device nodes are the builder's inner nodes in depth-first
  // preorder; a root that is a leaf gets a device node of its own.
  out->tri_order.resize(n);
  std::vector<uint32_t> slot_of(b.nodes.size(), 0);
  std::vector<uint32_t> stack;
  const BNode& rt = b.nodes[root];
  BvhNode empty;
  memset(&empty, 0, sizeof empty);
  empty.child[0] = empty.child[1] = -1;
  // Emit triangles in leaf order while assigning device slots.
  uint32_t tri_slot = 0;
  std::vector<uint32_t> leaf_slot(b.nodes.size(), 0);
  {
    // Leaves in DFS order (left first) get contiguous triangle slots.
    stack.push_back(root);
    while (!stack.empty()) {
      uint32_t id = stack.back();
      stack.pop_back();
      const BNode& nd = b.nodes[id];
      if (nd.left < 0) {
        leaf_slot[id] = tri_slot;
        for (uint32_t i = 0; i < nd.count; ++i) out->tri_order[tri_slot++] = b.order[nd.begin + i];
        out->max_depth = std::max(out->max_depth, nd.depth);
      } else {
        stack.push_back(uint32_t(nd.right));
        stack.push_back(uint32_t(nd.left));
      }
    }
  }
  if (rt.left < 0) {
    BvhNode dn = empty;
    set_child(dn, 0, rt, 0, int32_t(rt.count));
    if (rt.count == 0) dn.child[0] = -1;
    for (int k = 0; k < 3; ++k) {  // empty right child: boxes unused
      (void)k;
    }
    out->nodes.push_back(dn);
  } else {
    // Preorder over inner nodes.
    std::vector<uint32_t> inner;
    stack.push_back(root);
    while (!stack.empty()) {
      uint32_t id = stack.back();
      stack.pop_back();
      const BNode& nd = b.nodes[id];
      if (nd.left < 0) continue;
      slot_of[id] = uint32_t(inner.size());
      inner.push_back(id);
      stack.push_back(uint32_t(nd.right));
      stack.push_back(uint32_t(nd.left));
    }
    out->nodes.resize(inner.size());
    for (size_t i = 0; i < inner.size(); ++i) {
      const BNode& nd = b.nodes[inner[i]];
      BvhNode dn = empty;
      for (int side = 0; side < 2; ++side) {
        const BNode& c = b.nodes[side == 0 ? nd.left : nd.right];
        if (c.left < 0)
          set_child(dn, side, c, int32_t(leaf_slot[side == 0 ? nd.left : nd.right]), int32_t(c.count));
        else
          set_child(dn, side, c, int32_t(slot_of[side == 0 ? nd.left : nd.right]), 0);
      }
      out->nodes[i] = dn;
    }
  }
  // SAH cost of the finished tree (traversal cost per inner node visit +
  // 1 per triangle), relative to the root area.
  double ra = rt.box.half_area();
  double cost = 0.0;
  if (ra > 0) {
    for (const BNode& nd : b.nodes) {
      double a = nd.box.half_area() / ra;
      if (nd.left < 0)
        cost += a * nd.count;
      else
        cost += a * opts.traversal_cost;
    }
  }
  out->sah_cost = cost;
}

namespace {

struct Ref {
  float lo[3], hi[3];
  int32_t child, count;
  double area() const {
    double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    return dx * dy + dy * dz + dz * dx;
  }
};

// ---- SAH-optimal collapse to 4-wide nodes -----------------------------------
//
// An entry is one side of a binary node (a child box with a triangle range or
// a binary subtree). D(e, k) is the least SAH cost of covering entry e's
// subtree with at most k child slots of the enclosing 4-wide node:
//   leaf entry:          D(e, k) = A(e) N(e) c_tri
//   inner entry, k = 1:  min(A(e) N(e) c_tri   [whole subtree as one leaf, N <= max_leaf],
//                            A(e) c_node + min_j D(L, j) + D(R, 4 - j)   [a 4-wide node])
//   inner entry, k >= 2: min(D(e, k - 1), min_j D(L, j) + D(R, k - j))   [e opened]
// (the dynamic program of Ylitie, Karras & Laine 2017, for width 4). Costs
// are in units of one triangle test; a 4-wide node visit costs c_node.

constexpr double kNodeCost = 1.0;

struct Collapse {
  const Bvh& bin;
  int max_leaf;
  // per entry (2 per binary node)
  std::vector<double> d;         // D(e, 1..4)
  std::vector<int8_t> kj;        // for k = 2..4 (index k - 1): 0 = as for k - 1, else j slots to the left side
  std::vector<int8_t> node_j;    // the split of the 4-wide node made from e
  std::vector<uint8_t> as_leaf;  // D(e, 1) covers e with one leaf
  std::vector<uint32_t> n_tris, first;

  Collapse(const Bvh& b, int ml) : bin(b), max_leaf(ml) {
    size_t ne = 2 * b.nodes.size();
    d.assign(4 * ne, 0.0);
    kj.assign(4 * ne, 0);
    node_j.assign(ne, 1);
    as_leaf.assign(ne, 0);
    n_tris.assign(ne, 0);
    first.assign(ne, 0);
  }

  Ref ref_of(uint32_t e) const {
    const BvhNode& n = bin.nodes[e >> 1];
    int side = int(e & 1u);
    Ref r;
    r.lo[0] = n.f[side * 4 + 0];
    r.hi[0] = n.f[side * 4 + 1];
    r.lo[1] = n.f[side * 4 + 2];
    r.hi[1] = n.f[side * 4 + 3];
    r.lo[2] = n.f[8 + side * 2];
    r.hi[2] = n.f[9 + side * 2];
    r.child = n.child[side];
    r.count = n.count[side];
    return r;
  }

  // Bottom-up over the binary tree (children before parents).
  void solve() {
    std::vector<uint32_t> order;
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
      uint32_t m = stack.back();
      stack.pop_back();
      order.push_back(m);
      for (int side = 0; side < 2; ++side)
        if (bin.nodes[m].child[side] >= 0 && bin.nodes[m].count[side] == 0)
          stack.push_back(uint32_t(bin.nodes[m].child[side]));
    }
    for (size_t i = order.size(); i-- > 0;)
      for (int side = 0; side < 2; ++side) solve_entry(2 * order[i] + uint32_t(side));
  }

  void solve_entry(uint32_t e) {
    const Ref r = ref_of(e);
    double* D = &d[4 * size_t(e)];
    if (r.child < 0) return;  // empty side of a single-leaf root: costs nothing
    const double A = r.area();
    if (r.count > 0) {
      n_tris[e] = uint32_t(r.count);
      first[e] = uint32_t(r.child);
      as_leaf[e] = 1;
      for (int k = 0; k < 4; ++k) D[k] = A * r.count;
      return;
    }
    const uint32_t L = 2 * uint32_t(r.child), R = L + 1;
    n_tris[e] = n_tris[L] + n_tris[R];
    first[e] = first[L];
    const double* DL = &d[4 * size_t(L)];
    const double* DR = &d[4 * size_t(R)];
    double open[5];
    int8_t oj[5];
    for (int k = 2; k <= 4; ++k) {
      open[k] = INFINITY;
      oj[k] = 1;
      for (int j = 1; j < k; ++j) {
        double c = DL[j - 1] + DR[k - j - 1];
        if (c < open[k]) {
          open[k] = c;
          oj[k] = int8_t(j);
        }
      }
    }
    node_j[e] = oj[4];
    const double node = A * kNodeCost + open[4];
    const double leaf = n_tris[e] <= uint32_t(max_leaf) ? A * double(n_tris[e]) : INFINITY;
    as_leaf[e] = leaf <= node ? 1 : 0;
    D[0] = std::min(leaf, node);
    for (int k = 2; k <= 4; ++k) {
      if (open[k] < D[k - 2]) {
        D[k - 1] = open[k];
        kj[4 * size_t(e) + k - 1] = oj[k];
      } else {
        D[k - 1] = D[k - 2];
        kj[4 * size_t(e) + k - 1] = 0;
      }
    }
  }

  // The slots that cover entry e with at most k of them.
  void distribute(uint32_t e, int k, std::vector<uint32_t>* slots) const {
    const Ref r = ref_of(e);
    if (r.child < 0) return;
    if (r.count > 0 || k == 1) {
      slots->push_back(e);
      return;
    }
    const int j = kj[4 * size_t(e) + k - 1];
    if (j == 0) {
      distribute(e, k - 1, slots);
      return;
    }
    const uint32_t L = 2 * uint32_t(r.child), R = L + 1;
    distribute(L, j, slots);
    distribute(R, k - j, slots);
  }

  // Emits the 4-wide node made from binary node m, its sides sharing the
  // four slots j : 4 - j; returns its index (preorder).
  uint32_t emit(uint32_t m, int j, uint32_t depth, Bvh4* out) const {
    std::vector<uint32_t> slots;
    if (bin.nodes[m].child[1] < 0) {
      distribute(2 * m, 4, &slots);
    } else {
      distribute(2 * m, j, &slots);
      distribute(2 * m + 1, 4 - j, &slots);
    }
    const uint32_t id = uint32_t(out->nodes.size());
    out->nodes.emplace_back();
    out->max_inner_depth = std::max(out->max_inner_depth, depth);
    Bvh4Node node;
    memset(&node, 0, sizeof node);
    for (int k = 0; k < 4; ++k) node.child[k] = -1;
    // leaf children first, into one run of slots, before any inner child's
    // subtree takes slots of its own
    for (size_t k = 0; k < slots.size(); ++k) {
      const uint32_t e = slots[k];
      const Ref r = ref_of(e);
      node.lo_x[k] = r.lo[0];
      node.hi_x[k] = r.hi[0];
      node.lo_y[k] = r.lo[1];
      node.hi_y[k] = r.hi[1];
      node.lo_z[k] = r.lo[2];
      node.hi_z[k] = r.hi[2];
      if (as_leaf[e]) {
        node.child[k] = int32_t(out->tri_order.size());
        node.count[k] = int32_t(n_tris[e]);
        for (uint32_t i = 0; i < n_tris[e]; ++i) out->tri_order.push_back(bin.tri_order[first[e] + i]);
      }
    }
    for (size_t k = 0; k < slots.size(); ++k) {
      const uint32_t e = slots[k];
      if (!as_leaf[e]) {
        node.child[k] = int32_t(emit(uint32_t(ref_of(e).child), node_j[e], depth + 1, out));
        node.count[k] = 0;
      }
    }
    out->nodes[id] = node;
    return id;
  }
};

}  // namespace

bool quantize_bvh4(const Bvh4& in, std::vector<uint32_t>* out) {
  out->assign(size_t(kNodeWords) * in.nodes.size(), 0u);
  for (size_t i = 0; i < in.nodes.size(); ++i) {
    const Bvh4Node& n = in.nodes[i];
    uint32_t* w = out->data() + kNodeWords * i;
    const float* lo[3] = {n.lo_x, n.lo_y, n.lo_z};
    const float* hi[3] = {n.hi_x, n.hi_y, n.hi_z};
    float origin[3];
    int ex[3];
    for (int a = 0; a < 3; ++a) {
      float o = INFINITY, h = -INFINITY;
      for (int k = 0; k < 4; ++k)
        if (n.child[k] >= 0) {
          o = std::min(o, lo[a][k]);
          h = std::max(h, hi[a][k]);
        }
      if (!(o <= h)) o = h = 0.0f;  // no children
      origin[a] = o;
      // smallest power of two with 255 steps covering the extent
      double ext = double(h) - double(o);
      int e = -100;
      if (ext > 0.0) {
        e = std::ilogb(ext) - 8;
        while (std::ldexp(255.0, e) < ext) ++e;
        while (e > -100 && std::ldexp(255.0, e - 1) >= ext) --e;
      }
      ex[a] = std::max(-100, std::min(100, e));
    }
    for (int a = 0; a < 3; ++a) memcpy(&w[a], &origin[a], 4);
    w[3] = (uint32_t(ex[0]) & 0xFFu) | ((uint32_t(ex[1]) & 0xFFu) << 8) | ((uint32_t(ex[2]) & 0xFFu) << 16);
    for (int k = 0; k < 4; ++k) {
      if (n.child[k] < 0) {
        w[10 + k] = kNoChild;
        continue;
      }
      for (int a = 0; a < 3; ++a) {
        double sc = std::ldexp(1.0, ex[a]);
        double ql = std::floor((double(lo[a][k]) - origin[a]) / sc);
        double qh = std::ceil((double(hi[a][k]) - origin[a]) / sc);
        uint32_t bl = uint32_t(std::max(0.0, std::min(255.0, ql)));
        uint32_t bh = uint32_t(std::max(0.0, std::min(255.0, qh)));
        w[4 + 2 * a] |= bl << (8 * k);
        w[5 + 2 * a] |= bh << (8 * k);
      }
      if (n.count[k] == 0) w[10 + k] = uint32_t(n.child[k]);  // inner node
    }
    // the leaf run: collapse_bvh4 gives a node's leaf children consecutive
    // slots in child order, at most 4 x 7 of them; a leaf child's reference
    // is the flag plus its slots as a mask over the run
    uint32_t base = 0xFFFFFFFFu;
    for (int k = 0; k < 4; ++k)
      if (n.child[k] >= 0 && n.count[k] > 0) base = std::min(base, uint32_t(n.child[k]));
    if (base != 0xFFFFFFFFu) {
      w[14] = base;
      for (int k = 0; k < 4; ++k)
        if (n.child[k] >= 0 && n.count[k] > 0) {
          const uint32_t off = uint32_t(n.child[k]) - base;
          if (n.count[k] > 7 || off + uint32_t(n.count[k]) > kLeafMaskBits) return false;
          w[10 + k] = kLeafFlag | (((1u << n.count[k]) - 1u) << off);
        }
    }
  }
  return true;
}

void collapse_bvh4(const Bvh& bin, Bvh4* out, int max_leaf) {
  *out = Bvh4();
  if (bin.nodes.empty()) return;
  out->nodes.reserve(bin.nodes.size() / 2 + 1);
  out->tri_order.reserve(bin.tri_order.size());
  Collapse c(bin, max_leaf);
  c.solve();
  // the root is always a 4-wide node: best split of its 4 slots
  const double* DL = &c.d[0];
  const double* DR = &c.d[4];
  int j = 1;
  double best = INFINITY;
  for (int a = 1; a < 4; ++a) {
    double v = DL[a - 1] + DR[4 - a - 1];
    if (v < best) {
      best = v;
      j = a;
    }
  }
  c.emit(0, j, 0, out);
  order_top_levels(out, kTopDepth);
}

void order_top_levels(Bvh4* b, int depth) {
  const size_t nn = b->nodes.size();
  if (nn == 0) return;
  std::vector<int> d(nn, -1);
  std::vector<uint32_t> bfs{0};
  d[0] = 0;
  for (size_t i = 0; i < bfs.size(); ++i) {
    const Bvh4Node& nd = b->nodes[bfs[i]];
    if (d[bfs[i]] >= depth) continue;
    for (int k = 0; k < 4; ++k)
      if (nd.child[k] >= 0 && nd.count[k] == 0) {
        d[size_t(nd.child[k])] = d[bfs[i]] + 1;
        bfs.push_back(uint32_t(nd.child[k]));
      }
  }
  std::vector<uint32_t> id(nn, 0xFFFFFFFFu);
  uint32_t next = 0;
  for (uint32_t v : bfs) id[v] = next++;
  for (size_t v = 0; v < nn; ++v)
    if (id[v] == 0xFFFFFFFFu) id[v] = next++;
  std::vector<Bvh4Node> out(nn);
  for (size_t v = 0; v < nn; ++v) {
    Bvh4Node nd = b->nodes[v];
    for (int k = 0; k < 4; ++k)
      if (nd.child[k] >= 0 && nd.count[k] == 0) nd.child[k] = int32_t(id[size_t(nd.child[k])]);
    out[id[v]] = nd;
  }
  b->nodes.swap(out);
}

}  // namespace aqua
