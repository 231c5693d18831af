// Ray traversal of the BVH2 and the watertight ray/triangle test.
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
  v3 oi;      // o * inv
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
  r.oi = mul(o, r.inv);
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

// Slab test of one child box; returns the entry distance or +inf if culled.
AQ_DEV float box_enter(const RayPre& r, float lx, float hx, float ly, float hy, float lz, float hz,
                       float tbest, float delta) {
  float t0x = __builtin_fmaf(lx, r.inv.x, -r.oi.x), t1x = __builtin_fmaf(hx, r.inv.x, -r.oi.x);
  float t0y = __builtin_fmaf(ly, r.inv.y, -r.oi.y), t1y = __builtin_fmaf(hy, r.inv.y, -r.oi.y);
  float t0z = __builtin_fmaf(lz, r.inv.z, -r.oi.z), t1z = __builtin_fmaf(hz, r.inv.z, -r.oi.z);
  float tn = fmax2(fmax2(fmin2(t0x, t1x), fmin2(t0y, t1y)), fmax2(fmin2(t0z, t1z), 0.0f));
  float tf = fmin2(fmin2(fmax2(t0x, t1x), fmax2(t0y, t1y)), fmin2(fmax2(t0z, t1z), tbest));
  return tn <= tf * kFarScale + delta ? tn : __builtin_inff();
}

struct DevBvh {
  const float4* nodes;  // 4 per node
  const float4* tris;   // 3 per triangle slot: (v0, id), (v1, material), (v2, 0)
  float delta;
};

struct TraceCount {
  uint32_t nodes, tris;
};

constexpr int kStackSize = 64;

// Closest hit. Returns the triangle slot (0xFFFFFFFF on a miss) and t.
template <bool COUNT>
AQ_DEV uint32_t trace_closest(const DevBvh& bvh, v3 o, v3 d, float tmax, float* t_out,
                              TraceCount* cnt) {
  RayPre r = ray_pre(o, d);
  float tbest = tmax;
  uint32_t best_id = 0xFFFFFFFFu, best_slot = 0xFFFFFFFFu;
  int stack[kStackSize];
  int sp = 0;
  int node = 0;
  while (true) {
    const float4* np = bvh.nodes + 4 * node;
    float4 a = np[0], b = np[1], c = np[2];
    int4 e = *reinterpret_cast<const int4*>(np + 3);
    if (COUNT) cnt->nodes++;
    float tl = e.x >= 0 ? box_enter(r, a.x, a.y, a.z, a.w, c.x, c.y, tbest, bvh.delta) : __builtin_inff();
    float tr = e.y >= 0 ? box_enter(r, b.x, b.y, b.z, b.w, c.z, c.w, tbest, bvh.delta) : __builtin_inff();
    bool hl = tl != __builtin_inff(), hr = tr != __builtin_inff();
    // Leaf children are tested right away.
    for (int side = 0; side < 2; ++side) {
      bool h = side == 0 ? hl : hr;
      int cntc = side == 0 ? e.z : e.w;
      if (!h || cntc == 0) continue;
      int first = side == 0 ? e.x : e.y;
      for (int k = 0; k < cntc; ++k) {
        const float4* tp = bvh.tris + 3 * (first + k);
        float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        if (COUNT) cnt->tris++;
        float t;
        if (tri_hit(r, xyz(p0), xyz(p1), xyz(p2), &t)) {
          uint32_t id = __float_as_uint(p0.w);
          if (t > 0.0f && t < tmax && (t < tbest || (t == tbest && id < best_id))) {
            tbest = t;
            best_id = id;
            best_slot = uint32_t(first + k);
          }
        }
      }
      if (side == 0)
        hl = false;
      else
        hr = false;
    }
    if (hl && hr) {
      bool left_first = tl <= tr;
      int near = left_first ? e.x : e.y, far = left_first ? e.y : e.x;
      stack[sp++] = far;
      node = near;
    } else if (hl) {
      node = e.x;
    } else if (hr) {
      node = e.y;
    } else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  *t_out = tbest;
  return best_slot;
}

// Any hit in (0, tmax).
template <bool COUNT>
AQ_DEV bool trace_any(const DevBvh& bvh, v3 o, v3 d, float tmax, TraceCount* cnt) {
  RayPre r = ray_pre(o, d);
  int stack[kStackSize];
  int sp = 0;
  int node = 0;
  while (true) {
    const float4* np = bvh.nodes + 4 * node;
    float4 a = np[0], b = np[1], c = np[2];
    int4 e = *reinterpret_cast<const int4*>(np + 3);
    if (COUNT) cnt->nodes++;
    float tl = e.x >= 0 ? box_enter(r, a.x, a.y, a.z, a.w, c.x, c.y, tmax, bvh.delta) : __builtin_inff();
    float tr = e.y >= 0 ? box_enter(r, b.x, b.y, b.z, b.w, c.z, c.w, tmax, bvh.delta) : __builtin_inff();
    bool hl = tl != __builtin_inff(), hr = tr != __builtin_inff();
    for (int side = 0; side < 2; ++side) {
      bool h = side == 0 ? hl : hr;
      int cntc = side == 0 ? e.z : e.w;
      if (!h || cntc == 0) continue;
      int first = side == 0 ? e.x : e.y;
      for (int k = 0; k < cntc; ++k) {
        const float4* tp = bvh.tris + 3 * (first + k);
        float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        if (COUNT) cnt->tris++;
        float t;
        if (tri_hit(r, xyz(p0), xyz(p1), xyz(p2), &t) && t > 0.0f && t < tmax) return true;
      }
      if (side == 0)
        hl = false;
      else
        hr = false;
    }
    if (hl && hr) {
      bool left_first = tl <= tr;
      stack[sp++] = left_first ? e.y : e.x;
      node = left_first ? e.x : e.y;
    } else if (hl) {
      node = e.x;
    } else if (hr) {
      node = e.y;
    } else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  return false;
}

}  // namespace aqd
