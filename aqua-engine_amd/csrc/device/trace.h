// Ray traversal of the 4-wide BVH and the watertight ray/triangle test.
//
// Closest hit is the least (t, triangle id) pair over every triangle the
// watertight test accepts with 0 < t < tmax. Node culling is conservative
// (Ize 2013, "Robust BVH Ray Traversal": far distances scaled by 1 + 2*gamma_3,
// plus an absolute slack `delta` proportional to the scene's extent that
// covers the rounding of the triangle test's t), so the answer is the same
// as testing every triangle, whatever the tree or the visiting order.
#pragma once

#include "common.h"

namespace aqd {

// 1 + 2 * gamma(3), gamma(n) = n*eps / (1 - n*eps), eps = 2^-24.
constexpr float kFarScale = 1.00000036f;

struct RayPre {
  v3 o, d;
  v3 inv;     // 1 / d, with |d| clamped away from 0
  int kx, ky, kz;
  float sx, sy, sz;
};

AQ_DEV float safe_inv(float d) {
  float dd = __builtin_fabsf(d) < 1e-20f ? __builtin_copysignf(1e-20f, d) : d;
  return 1.0f / dd;
}

AQ_DEV float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

AQ_DEV RayPre ray_pre(v3 o, v3 d) {
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = v3{safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
  // Woop, Benthin & Wald 2013, "Watertight Ray/Triangle Intersection":
  // kz = dominant axis of d, kx/ky the others (swapped if d[kz] < 0 to keep
  // the winding), shear constants Sx, Sy, Sz.
  float ax = __builtin_fabsf(d.x), ay = __builtin_fabsf(d.y), az = __builtin_fabsf(d.z);
  int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dz = comp(d, kz);
  if (dz < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.sx = comp(d, kx) / dz;
  r.sy = comp(d, ky) / dz;
  r.sz = 1.0f / dz;
  return r;
}

// Watertight test. Returns true and t when the ray crosses the triangle with
// all three edge functions of one sign (zero counts as either sign).
AQ_DEV bool tri_hit(const RayPre& r, v3 v0, v3 v1, v3 v2, float* t_out) {
  v3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.sx * Akz;
  float Ay = comp(A, r.ky) - r.sy * Akz;
  float Bx = comp(B, r.kx) - r.sx * Bkz;
  float By = comp(B, r.ky) - r.sy * Bkz;
  float Cx = comp(C, r.kx) - r.sx * Ckz;
  float Cy = comp(C, r.ky) - r.sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = (U + V) + W;
  if (det == 0.0f) return false;
  float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  float T = __builtin_fmaf(U, Az, __builtin_fmaf(V, Bz, W * Cz));
  *t_out = T / det;
  return true;
}

// Slab test of one box from its near and far plane distances along the ray
// per axis: *tn = the entry distance, true when the box is not culled. The t
// values are finite (|1/d| is clamped), so the hardware min / max (v_min3 /
// v_max3) apply. Culling only has to be conservative: the far bound is
// widened by kFarScale and delta, which also covers the rounding of the
// plane distances.
AQ_DEV bool slab_test(float nx, float fx, float ny, float fy, float nz, float fz, float tbest, float delta,
                      float* tn) {
  const float e = __builtin_fmaxf(__builtin_fmaxf(nx, ny), __builtin_fmaxf(nz, 0.0f));
  const float f = __builtin_fminf(__builtin_fminf(fx, fy), __builtin_fminf(fz, tbest));
  *tn = e;
  return e <= __builtin_fmaf(f, kFarScale, delta);
}

// A child reference with this bit set is a leaf: the rest of the word is the
// mask of its triangle slots over the node's leaf run (an empty child is a
// leaf of no triangles). Without it, the word is an inner node's index.
constexpr uint32_t kLeafFlag = 0x80000000u;

// The first kTopNodes nodes (the builder puts the top three levels there)
// are staged in LDS by every traversal block: every ray starts at the root,
// so these are the most-fetched nodes.
constexpr uint32_t kTopNodes = 85;

struct DevBvh {
  const float4* nodes;  // 4 per 4-wide node (64-byte quantised nodes)
  const float4* top;    // nodes [0, n_top) in LDS (set by the kernel)
  uint32_t n_top;       // min(kTopNodes, node count)
  const float4* tris;   // 3 per triangle slot: (v0, id), (v1, material), (v2, 0)
  float delta;
  uint32_t* spill;      // traversal spill stacks: kSpillStack x (grid threads) words
};

struct TraceCount {
  uint32_t nodes, tris;
  uint32_t deep;  // node visits while the ray's stack had entries spilled to global memory
};

// ---- traversal stack: a 16-entry ring in LDS (one column per thread) whose
// oldest entries spill to global memory when it is full. With at most three
// pushes per visited node the host bounds the total depth (3 x tree depth)
// by kLdsStack + kSpillStack.
//
// The spill area is a global buffer the renderer owns, entry i of grid
// thread t at spill[i * threads + t] (lanes at the same depth coalesce), not
// a private array: private arrays put the kernel on scratch memory, whose
// per-launch set-up cost (~0.1 ms at the traversal grids) exceeds the whole
// traversal of a short queue.
constexpr int kLdsStack = 16;
constexpr int kSpillStack = 64;
// Threads per block of every kernel that traverses (the LDS column stride).
constexpr int kStackThreads = 256;

// Only scalars live in the struct, so they stay in registers.
struct TStack {
  uint32_t* lds;    // this thread's column
  uint32_t* spill;  // this thread's first spill entry
  uint32_t sstride; // threads per grid
  int top, cnt, gsp;
};

// Every thread of the block calls it: copies nodes [0, bvh.n_top) to `lds`
// and points bvh.top at it.
AQ_DEV void stage_top(DevBvh& bvh, float4* lds) {
  for (uint32_t i = threadIdx.x; i < 4 * bvh.n_top; i += blockDim.x) lds[i] = bvh.nodes[i];
  __syncthreads();
  bvh.top = lds;
}

// Thread (blockIdx, threadIdx) of a grid of gridDim x kStackThreads threads.
AQ_DEV void st_init(TStack& s, uint32_t* lds_base, uint32_t* spill_base) {
  const uint32_t tid = threadIdx.x;
  s.lds = lds_base + tid;
  s.sstride = gridDim.x * uint32_t(kStackThreads);
  s.spill = spill_base + size_t(blockIdx.x) * kStackThreads + tid;
  s.top = kLdsStack - 1;
  s.cnt = 0;
  s.gsp = 0;
}
static_assert((kLdsStack & (kLdsStack - 1)) == 0, "the ring index wraps with a mask");
AQ_DEV void st_push(TStack& s, uint32_t v) {
  const int slot = (s.top + 1) & (kLdsStack - 1);
  if (s.cnt == kLdsStack)
    s.spill[size_t(s.gsp++) * s.sstride] = s.lds[slot * kStackThreads];
  else
    ++s.cnt;
  s.lds[slot * kStackThreads] = v;
  s.top = slot;
}
AQ_DEV bool st_empty(const TStack& s) { return s.cnt == 0 && s.gsp == 0; }
AQ_DEV uint32_t st_pop(TStack& s) {
  if (s.cnt > 0) {
    uint32_t v = s.lds[s.top * kStackThreads];
    s.top = (s.top - 1) & (kLdsStack - 1);
    --s.cnt;
    return v;
  }
  return s.spill[size_t(--s.gsp) * s.sstride];
}

// Branch-free choice of v[k], k in 0..3 (three selects).
AQ_DEV uint32_t pick4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t k) {
  uint32_t lo = (k & 1u) ? b : a, hi = (k & 1u) ? d : c;
  return (k & 2u) ? hi : lo;
}

// Byte k of a word as a float (v_cvt_f32_ubyteK).
AQ_DEV float ubyte(uint32_t w, int k) { return float((w >> (8 * k)) & 0xFFu); }

// 2^e x v for the signed exponent byte at bit `sh` of w (v_bfe_i32 and
// v_ldexp_f32; exact but for overflow and underflow, like the product).
AQ_DEV float scale_by(float v, uint32_t w, int sh) {
  return __builtin_amdgcn_ldexpf(v, int(int8_t(uint8_t(w >> sh))));
}

// Four slab tests of one 64-byte quantised node (four 16-byte loads issued
// together, straight-line code). A child plane is origin + q * 2^e, so its
// distance along the ray is q * (2^e / d) + (origin - o) / d: one FMA per
// plane after six per-node products. hit[k] tells whether child k's box
// survives culling and tn[k] is its entry distance; ref[k] is the child
// reference (an inner node's index, or kLeafFlag | the leaf's slot mask over
// the run starting at *lbase). An empty child is a leaf of no triangles, so
// it needs no test here: whatever its distance, it adds no triangle and is
// never pushed.
AQ_DEV void node_test4(const RayPre& r, const float4* np, float tbest, float delta, float tn[4], bool hit[4],
                       uint32_t ref[4], uint32_t* lbase) {
  const float4 h = np[0];
  const uint4 q0 = *reinterpret_cast<const uint4*>(np + 1);  // lo.x hi.x lo.y hi.y
  const uint4 q1 = *reinterpret_cast<const uint4*>(np + 2);  // lo.z hi.z ref0 ref1
  const uint4 q2 = *reinterpret_cast<const uint4*>(np + 3);  // ref2 ref3 leaf run -
  const uint32_t e = __float_as_uint(h.w);
  const float ax = scale_by(r.inv.x, e, 0), ay = scale_by(r.inv.y, e, 8), az = scale_by(r.inv.z, e, 16);
  const float bx = (h.x - r.o.x) * r.inv.x, by = (h.y - r.o.y) * r.inv.y, bz = (h.z - r.o.z) * r.inv.z;
  ref[0] = q1.z;
  ref[1] = q1.w;
  ref[2] = q2.x;
  ref[3] = q2.y;
  *lbase = q2.z;
  // Near and far planes per axis from the sign of the direction: the plane
  // distance fma(q, a, b) is monotone in q, and a has the sign of 1/d, so
  // this picks exactly min / max of the two distances.
  const bool fx = r.inv.x < 0.0f, fy = r.inv.y < 0.0f, fz = r.inv.z < 0.0f;
  const uint32_t nx = fx ? q0.y : q0.x, xx = fx ? q0.x : q0.y;
  const uint32_t ny = fy ? q0.w : q0.z, xy = fy ? q0.z : q0.w;
  const uint32_t nz = fz ? q1.y : q1.x, xz = fz ? q1.x : q1.y;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    hit[k] = slab_test(__builtin_fmaf(ubyte(nx, k), ax, bx), __builtin_fmaf(ubyte(xx, k), ax, bx),
                       __builtin_fmaf(ubyte(ny, k), ay, by), __builtin_fmaf(ubyte(xy, k), ay, by),
                       __builtin_fmaf(ubyte(nz, k), az, bz), __builtin_fmaf(ubyte(xz, k), az, bz), tbest, delta,
                       &tn[k]);
}

AQ_DEV void cswap(uint32_t& a, uint32_t& b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

// The triangles of a node's hit leaf children as one mask over its leaf run.
AQ_DEV uint32_t leaf_mask(const bool hit[4], const uint32_t ref[4]) {
  uint32_t m = 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) m |= hit[k] && (ref[k] & kLeafFlag) ? ref[k] : 0u;
  return m & ~kLeafFlag;
}

// ---- resumable traversal: the state of one ray between steps, so a
// persistent wave can give a lane a new ray as soon as its last one ends
// instead of idling until the wave's slowest ray is done. A step is either
// one node visit or one round of (up to) two leaf triangles: a node visit
// only records the hit leaves' triangles as a bit mask, and the wave tests
// them in rounds of its own once the lanes holding such triangles outnumber
// half the lanes ready to visit a node (kLeafRatio), so a leaf round serves
// many lanes instead of the few whose last node had leaves. Neither answer
// depends on the order of triangle tests and node visits: the closest hit
// is the least (t, triangle id) over every triangle the watertight test
// accepts in a box the conservative culling keeps, and deferring a leaf
// only culls less.

constexpr uint32_t kNoNode = 0xFFFFFFFFu;
// A wave takes a leaf round once kLeafRatio x (lanes with leaf triangles
// pending) exceeds the lanes that can visit a node (measured: 2 beats 3 by
// 0.4 % on the 1 M-triangle interior, both kernels; 1 and 1.5 lose).
constexpr uint32_t kLeafRatio = 2;

struct ClosestRay {
  RayPre r;
  float tmax, tbest;
  uint32_t best_id, best_slot;
  uint32_t node;          // next node to visit, kNoNode when none is left
  uint32_t lbase, lmask;  // leaf triangles still to test: slots lbase + set bits
};

AQ_DEV void closest_begin(ClosestRay& c, v3 o, v3 d, float tmax) {
  c.r = ray_pre(o, d);
  c.tmax = tmax;
  c.tbest = tmax;
  c.best_id = 0xFFFFFFFFu;
  c.best_slot = 0xFFFFFFFFu;
  c.node = 0;
  c.lbase = 0;
  c.lmask = 0;
}

// The ray is finished: its result is (best_slot, tbest), best_slot
// 0xFFFFFFFF = miss, and its stack is empty.
AQ_DEV bool closest_done(const ClosestRay& c) { return c.node == kNoNode && c.lmask == 0u; }

// One node visit (the ray has no leaf triangles pending): records the hit
// leaves, pushes the other hit inner children far-first and moves to the
// nearest one (or the top of the stack).
template <bool COUNT>
AQ_DEV void closest_node(const DevBvh& bvh, TStack& st, ClosestRay& c, TraceCount* cnt) {
  float tn[4];
  bool hit[4];
  uint32_t ref[4];
  node_test4(c.r, (c.node < bvh.n_top ? bvh.top : bvh.nodes) + 4 * c.node, c.tbest, bvh.delta, tn, hit, ref,
             &c.lbase);
  if (COUNT) cnt->nodes++;
  c.lmask = leaf_mask(hit, ref);
  // key = entry distance bits with the child index in the low two bits
  // (entry distances are finite and >= 0) for a hit inner child, all ones
  // otherwise: keys below kKeyInf are the hit inner children
  uint32_t key[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    key[k] = hit[k] && !(ref[k] & kLeafFlag) ? ((__float_as_uint(tn[k]) & ~3u) | uint32_t(k)) : 0xFFFFFFFFu;
  cswap(key[0], key[1]);
  cswap(key[2], key[3]);
  cswap(key[0], key[2]);
  cswap(key[1], key[3]);
  cswap(key[1], key[2]);
  constexpr uint32_t kKeyInf = 0x7F800000u;
#pragma unroll
  for (int j = 3; j >= 1; --j)
    if (key[j] < kKeyInf) st_push(st, pick4(ref[0], ref[1], ref[2], ref[3], key[j] & 3u));
  if (COUNT && st.gsp > 0) cnt->deep++;
  if (key[0] < kKeyInf)
    c.node = pick4(ref[0], ref[1], ref[2], ref[3], key[0] & 3u);
  else
    c.node = st_empty(st) ? kNoNode : st_pop(st);
}

// The next two of the ray's pending leaf triangles, both fetched before
// either is tested.
template <bool COUNT>
AQ_DEV void closest_leaf(const DevBvh& bvh, ClosestRay& c, TraceCount* cnt) {
  uint32_t m = c.lmask;
  const uint32_t i0 = __builtin_ctz(m);
  m &= m - 1u;
  const bool two = m != 0u;
  const uint32_t i1 = two ? uint32_t(__builtin_ctz(m)) : i0;
  c.lmask = m & (m - 1u);
  const float4* ta = bvh.tris + 3 * (c.lbase + i0);
  const float4* tb = bvh.tris + 3 * (c.lbase + i1);
  const float4 a0 = ta[0], a1 = ta[1], a2 = ta[2];
  const float4 b0 = tb[0], b1 = tb[1], b2 = tb[2];
  if (COUNT) cnt->tris += two ? 2u : 1u;
  float t;
  if (tri_hit(c.r, xyz(a0), xyz(a1), xyz(a2), &t)) {
    uint32_t id = __float_as_uint(a0.w);
    if (t > 0.0f && t < c.tmax && (t < c.tbest || (t == c.tbest && id < c.best_id))) {
      c.tbest = t;
      c.best_id = id;
      c.best_slot = c.lbase + i0;
    }
  }
  if (two && tri_hit(c.r, xyz(b0), xyz(b1), xyz(b2), &t)) {
    uint32_t id = __float_as_uint(b0.w);
    if (t > 0.0f && t < c.tmax && (t < c.tbest || (t == c.tbest && id < c.best_id))) {
      c.tbest = t;
      c.best_id = id;
      c.best_slot = c.lbase + i1;
    }
  }
}

// One step of a wave (every lane calls it together; inactive lanes hold no
// ray): a leaf round once the lanes with leaf triangles pending (which wait
// meanwhile) number more than 1 / kLeafRatio of those that can visit a node,
// else a node visit.
template <bool COUNT>
AQ_DEV void closest_wave_step(const DevBvh& bvh, TStack& st, ClosestRay& c, bool active, TraceCount* cnt) {
  const bool wl = active && c.lmask != 0u;
  const bool wn = active && !wl && c.node != kNoNode;
  if (kLeafRatio * __popcll(__ballot(wl)) > __popcll(__ballot(wn))) {
    if (wl) closest_leaf<COUNT>(bvh, c, cnt);
  } else if (wn) {
    closest_node<COUNT>(bvh, st, c, cnt);
  }
}

struct AnyRay {
  RayPre r;
  float tmax;
  uint32_t node;          // next node to visit, kNoNode when none is left
  uint32_t lbase, lmask;  // leaf triangles still to test
  bool occluded;
};

AQ_DEV void any_begin(AnyRay& a, v3 o, v3 d, float tmax) {
  a.r = ray_pre(o, d);
  a.tmax = tmax;
  a.node = 0;
  a.lbase = 0;
  a.lmask = 0;
  a.occluded = false;
}

// Finished: the answer is a.occluded, and the stack is empty.
AQ_DEV bool any_done(const AnyRay& a) { return a.node == kNoNode && a.lmask == 0u; }

template <bool COUNT>
AQ_DEV void any_node(const DevBvh& bvh, TStack& st, AnyRay& a, TraceCount* cnt) {
  float tn[4];
  bool hit[4];
  uint32_t ref[4];
  node_test4(a.r, (a.node < bvh.n_top ? bvh.top : bvh.nodes) + 4 * a.node, a.tmax, bvh.delta, tn, hit, ref,
             &a.lbase);
  if (COUNT) cnt->nodes++;
  a.lmask = leaf_mask(hit, ref);
  uint32_t next = kNoNode;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!hit[k] || (ref[k] & kLeafFlag)) continue;
    if (next == kNoNode)
      next = ref[k];
    else
      st_push(st, ref[k]);
  }
  if (COUNT && st.gsp > 0) cnt->deep++;
  a.node = next != kNoNode ? next : (st_empty(st) ? kNoNode : st_pop(st));
}

// Any hit in (0, tmax) among the next two pending leaf triangles ends the
// ray (and empties its stack).
template <bool COUNT>
AQ_DEV void any_leaf(const DevBvh& bvh, TStack& st, AnyRay& a, TraceCount* cnt) {
  uint32_t m = a.lmask;
  const uint32_t i0 = __builtin_ctz(m);
  m &= m - 1u;
  const bool two = m != 0u;
  const uint32_t i1 = two ? uint32_t(__builtin_ctz(m)) : i0;
  a.lmask = m & (m - 1u);
  const float4* ta = bvh.tris + 3 * (a.lbase + i0);
  const float4* tb = bvh.tris + 3 * (a.lbase + i1);
  const float4 a0 = ta[0], a1 = ta[1], a2 = ta[2];
  const float4 b0 = tb[0], b1 = tb[1], b2 = tb[2];
  if (COUNT) cnt->tris += two ? 2u : 1u;
  float t;
  bool hit = tri_hit(a.r, xyz(a0), xyz(a1), xyz(a2), &t) && t > 0.0f && t < a.tmax;
  if (!hit && two) hit = tri_hit(a.r, xyz(b0), xyz(b1), xyz(b2), &t) && t > 0.0f && t < a.tmax;
  if (hit) {
    a.occluded = true;
    a.lmask = 0u;
    a.node = kNoNode;
    st.cnt = 0;
    st.gsp = 0;
  }
}

template <bool COUNT>
AQ_DEV void any_wave_step(const DevBvh& bvh, TStack& st, AnyRay& a, bool active, TraceCount* cnt) {
  const bool wl = active && a.lmask != 0u;
  const bool wn = active && !wl && a.node != kNoNode;
  if (kLeafRatio * __popcll(__ballot(wl)) > __popcll(__ballot(wn))) {
    if (wl) any_leaf<COUNT>(bvh, st, a, cnt);
  } else if (wn) {
    any_node<COUNT>(bvh, st, a, cnt);
  }
}

}  // namespace aqd
