// Persistent traversal kernels: the work queue that feeds their waves
// (WaveWork, LaneFeed), closest-hit traversal of a ray queue (k_intersect;
// path queues or caller rays) and any-hit traversal (k_shadow; shadow-ray
// queues or caller rays). Included by renderer.hip only; kept apart from the shading
// kernels so that the sources of the traversal kernels' instruction stream
// are these device headers alone.
#pragma once

#include "kernels.h"

namespace aqd {

// Persistent traversal waves. The queue [0, n) is cut into 8 contiguous
// 64-aligned ranges; range x belongs to the blocks b with b % 8 == x, which
// the hardware deals to XCD x, so a range's rays share that XCD's L2. A wave
// first takes its static 64-ray chunk of its own range, then further chunks
// from its range's cursor, then helps the other ranges through theirs. The
// own range's cursor is incremented without a look first: every claim is a
// round trip to a line that all of the XCD's waves hit, and a read before
// each increment had doubled that cost (it was most of the time of a queue
// of short rays). Another range's cursor is read before it is incremented,
// so exhausted ranges cost no atomics; each cursor has its own 128-byte line.
// The grid is a multiple of 8 blocks.
struct WaveWork {
  uint32_t n, per, waves_x, wave_x, x;
  uint32_t* cur;  // 8 cursors, kCursorStride words apart
  uint32_t phase; // 0: static chunk; 1..8: cursor of range (x + phase - 1) % 8
};

__device__ __forceinline__ WaveWork ww_init(uint32_t n, uint32_t* cursors) {
  WaveWork w;
  w.n = n;
  w.per = (((n + 7u) >> 3) + 63u) & ~63u;
  w.x = blockIdx.x & 7u;
  w.waves_x = (gridDim.x >> 3) * (blockDim.x >> 6);
  w.wave_x = (blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  w.cur = cursors;
  w.phase = 0;
  return w;
}

// Base of the wave's next chunk, or kMiss when every range is done.
__device__ __forceinline__ uint32_t ww_next(WaveWork& w) {
  if (w.phase == 0) {
    w.phase = 1;
    uint64_t lo = min(uint64_t(w.n), uint64_t(w.x) * w.per), hi = min(uint64_t(w.n), lo + w.per);
    uint64_t base = lo + 64ull * w.wave_x;
    if (base < hi) return uint32_t(base);
  }
  while (w.phase <= 8u) {
    const uint32_t r = (w.x + w.phase - 1u) & 7u;
    const uint64_t lo = min(uint64_t(w.n), uint64_t(r) * w.per), hi = min(uint64_t(w.n), lo + w.per);
    uint32_t c = kMiss;
    if (__lane_id() == 0) {
      uint32_t* p = w.cur + r * kCursorStride;
      if (w.phase == 1u ||
          lo + 64ull * (uint64_t(w.waves_x) + __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < hi)
        c = atomicAdd(p, 1u);
    }
    c = __shfl(c, 0);
    if (c != kMiss) {
      uint64_t base = lo + 64ull * (uint64_t(w.waves_x) + c);
      if (base < hi) return uint32_t(base);
    }
    ++w.phase;
  }
  return kMiss;
}

// Per-lane ray refill on top of WaveWork: the lanes that want a ray take the
// next rays of the wave's current 64-ray chunk in lane order, and the wave
// fetches further chunks as that one runs out. Every lane of the wave calls
// it together (the bookkeeping is wave-uniform).
struct LaneFeed {
  WaveWork ww;
  uint32_t pos, end;  // the wave's current chunk: rays [pos, end) not yet handed out
  bool more;          // chunks may remain
};

__device__ __forceinline__ LaneFeed feed_init(uint32_t n, uint32_t* cursors) {
  LaneFeed f;
  f.ww = ww_init(n, cursors);
  f.pos = 0;
  f.end = 0;
  f.more = true;
  return f;
}

// This lane's new ray index, or kMiss (it did not ask, or the queue is done).
__device__ __forceinline__ uint32_t feed_take(LaneFeed& f, bool wants) {
  const unsigned long long m = __ballot(wants);
  const uint32_t k = uint32_t(__popcll(m));
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
  uint32_t got = kMiss, given = 0;
  while (given < k && f.more) {
    if (f.pos >= f.end) {
      uint32_t base = ww_next(f.ww);
      if (base == kMiss) {
        f.more = false;
        break;
      }
      f.pos = base;
      f.end = min(base + 64u, f.ww.n);
    }
    uint32_t take = min(k - given, f.end - f.pos);
    if (wants && rank >= given && rank < given + take) got = f.pos + (rank - given);
    f.pos += take;
    given += take;
  }
  return got;
}

// A wave keeps traversing while more than this many of its lanes hold a ray,
// then refills the idle ones (until the queue is drained). With chunk claims
// a single atomic each, refilling once 16 lanes are idle beats waiting for 24.
constexpr uint32_t kRefillBelow = 48;

// Where the closest-hit kernel reads its rays and writes its answers.
// Path queues: origin / direction arrays, and the hit record (triangle slot,
// t) the shading kernel reads.
struct QueueIO {
  const float4* __restrict__ ro;
  const float4* __restrict__ rd;
  uint2* hit;
  __device__ __forceinline__ void load(uint32_t r, float4* o, float4* d) const {
    *o = ro[r];
    *d = rd[r];
  }
  __device__ __forceinline__ void store(uint32_t r, const ClosestRay& c) const {
    hit[r] = make_uint2(c.best_slot, __float_as_uint(c.tbest));
  }
};
// Caller rays, 8 floats each (origin, tmax, direction, -), answered straight
// into the caller's arrays: triangle id (load order; kMiss = none) and t
// (tmax for a miss), the values aqua_trace_closest returns.
struct RayIO {
  const float4* __restrict__ rays;
  uint32_t* ids;
  float* ts;
  __device__ __forceinline__ void load(uint32_t r, float4* o, float4* d) const {
    *o = rays[2 * size_t(r)];
    *d = rays[2 * size_t(r) + 1];
  }
  __device__ __forceinline__ void store(uint32_t r, const ClosestRay& c) const {
    ids[r] = c.best_id;
    ts[r] = c.tbest;
  }
};

template <bool COUNT, class IO>
__global__ void __launch_bounds__(kStackThreads, 7) k_intersect(DevBvh bvh, IO io, const uint32_t* count, uint32_t* cursor,
                                                      unsigned long long* stats) {
  __shared__ float4 lds_top[4 * kTopNodes];
  stage_top(bvh, lds_top);
  __shared__ uint32_t lds_stack[kLdsStack * kStackThreads];
  TStack st;
  st_init(st, lds_stack, bvh.spill);
  const uint32_t n = *count;
  TraceCount tc{0u, 0u, 0u};
  LaneFeed feed = feed_init(n, cursor);
  ClosestRay cr;
  bool active = false;
  uint32_t ray = 0;
  while (true) {
    if (feed.more) {
      uint32_t r = feed_take(feed, !active);
      if (r != kMiss) {
        float4 o, d;
        io.load(r, &o, &d);
        closest_begin(cr, xyz(o), xyz(d), o.w);
        ray = r;
        active = true;
      }
    }
    const uint32_t keep = feed.more ? kRefillBelow : 0u;
    uint32_t live = uint32_t(__popcll(__ballot(active)));
    if (live == 0) break;
    while (live > keep) {
      closest_wave_step<COUNT>(bvh, st, cr, active, &tc);
      if (active && closest_done(cr)) {
        io.store(ray, cr);
        active = false;
      }
      live = uint32_t(__popcll(__ballot(active)));
    }
  }
  if (COUNT) {
    atomicAdd(&stats[kStNodes], (unsigned long long)tc.nodes);
    atomicAdd(&stats[kStTris], (unsigned long long)tc.tris);
    atomicAdd(&stats[kStSpill], (unsigned long long)tc.deep);
  }
}

// Where the any-hit kernel reads its rays and writes its answers, with the
// state a lane keeps for its ray meanwhile. Shadow queues: an unoccluded ray
// adds its contribution to the radiance of its path (the next bounce's slot,
// or the pass buffer for a path that ended). Each path has at most one shadow
// ray per bounce and nothing else writes that radiance meanwhile, so the sum
// is formed when the ray is taken (its loads overlap the traversal) and only
// stored when it ends.
struct ShadowIO {
  Queues q;
  uint32_t nx;  // radiance buffer of the next bounce's paths
  struct Lane {
    float4* Lp;
    float4 sum;
  };
  __device__ __forceinline__ void load(uint32_t r, float4* o, float4* d, Lane* l) const {
    *o = q.sho[r];
    *d = q.shd[r];
    const float4 cc = q.shc[r];
    const uint32_t dst = __float_as_uint(d->w);
    l->Lp = (dst & 0x80000000u) ? &q.Lbuf[dst & 0x7FFFFFFFu] : &q.L[nx][dst];
    const float4 L = *l->Lp;
    l->sum = make_float4(L.x + cc.x, L.y + cc.y, L.z + cc.z, 0.0f);
  }
  __device__ __forceinline__ void store(uint32_t, const Lane& l, bool occluded) const {
    if (!occluded) *l.Lp = l.sum;
  }
};
// Caller rays (origin and tmax, direction): 1 where the open segment
// (0, tmax) hits a triangle, else 0 -- what aqua_trace_occluded returns.
struct OcclusionIO {
  const float4* __restrict__ ro;
  const float4* __restrict__ rd;
  uint8_t* occluded;
  struct Lane {};
  __device__ __forceinline__ void load(uint32_t r, float4* o, float4* d, Lane*) const {
    *o = ro[r];
    *d = rd[r];
  }
  __device__ __forceinline__ void store(uint32_t r, const Lane&, bool occ) const { occluded[r] = occ ? 1u : 0u; }
};

// Any-hit traversal of the rays [0, *count) that IO supplies, persistent like
// k_intersect (cursor: the launch's 8 range cursors, zeroed).
template <bool COUNT, class IO>
__global__ void __launch_bounds__(kStackThreads, 7) k_shadow(DevBvh bvh, IO io, const uint32_t* count,
                                                             uint32_t* cursor, unsigned long long* stats) {
  __shared__ float4 lds_top[4 * kTopNodes];
  stage_top(bvh, lds_top);
  __shared__ uint32_t lds_stack[kLdsStack * kStackThreads];
  TStack st;
  st_init(st, lds_stack, bvh.spill);
  const uint32_t n = *count;
  TraceCount tc{0u, 0u, 0u};
  LaneFeed feed = feed_init(n, cursor);
  AnyRay ar;
  typename IO::Lane ls{};
  bool active = false;
  uint32_t ray = 0;
  while (true) {
    if (feed.more) {
      uint32_t r = feed_take(feed, !active);
      if (r != kMiss) {
        float4 o, d;
        io.load(r, &o, &d, &ls);
        any_begin(ar, xyz(o), xyz(d), o.w);
        ray = r;
        active = true;
      }
    }
    const uint32_t keep = feed.more ? kRefillBelow : 0u;
    uint32_t live = uint32_t(__popcll(__ballot(active)));
    if (live == 0) break;
    while (live > keep) {
      any_wave_step<COUNT>(bvh, st, ar, active, &tc);
      if (active && any_done(ar)) {
        io.store(ray, ls, ar.occluded);
        active = false;
      }
      live = uint32_t(__popcll(__ballot(active)));
    }
  }
  if (COUNT) {
    atomicAdd(&stats[kStShNodes], (unsigned long long)tc.nodes);
    atomicAdd(&stats[kStShTris], (unsigned long long)tc.tris);
    atomicAdd(&stats[kStSpill], (unsigned long long)tc.deep);
  }
}

}  // namespace aqd
