// Wavefront path-tracing kernels (declarations shared by renderer.hip).
//
// Path state lives in structure-of-arrays queues in HBM, one float4/uint4 per
// field so every access is a coalesced 16-byte load or store per lane. Each
// bounce runs closest-hit traversal over the ray queue, then shading, which
// compacts surviving paths into the other half of a ping-pong queue (wave64
// ballot + mbcnt prefix + one atomic per wave) and emits next-event shadow
// rays into a compacted shadow queue, then any-hit traversal of the shadow
// queue. Queue lengths stay in device memory, so a frame needs no host sync.
#pragma once

#include "bsdf.h"
#include "morton.h"
#include "texture.h"
#include "trace.h"

namespace aqd {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kMaxBounces = 64;
// counters[] layout. Words [0, 2 kMaxBounces + 4) hold the queue lengths:
// rays of bounce b at 2 b, shadow rays emitted at bounce b at 2 b + 3, so
// the ray count of bounce b + 1 and the shadow count of bounce b form one
// 8-byte-aligned pair that the shading kernel bumps with a single 64-bit
// atomic (ray count in the low word).
__host__ __device__ constexpr uint32_t cnt_rays(uint32_t b) { return 2 * b; }
__host__ __device__ constexpr uint32_t cnt_shadow(uint32_t b) { return 2 * b + 3; }
constexpr uint32_t kCntQuery = 304;     // cache queries emitted this pass
constexpr uint32_t kCntTrain = 305;     // cache training samples emitted this pass
constexpr uint32_t kCntWords = 320;
// cursors[] layout: the persistent traversal kernels' work cursors, 8 per
// launch (one per XCD range), each on its own 128-byte line.
constexpr uint32_t kCursorStride = 32;                 // words between cursors
constexpr uint32_t kCursorSet = 8 * kCursorStride;     // one launch
constexpr uint32_t kCursorWords = 2 * kMaxBounces * kCursorSet;
// cursor set of the closest-hit (kind 0) or any-hit (kind 1) launch of bounce b
__host__ __device__ constexpr uint32_t cursor_set(uint32_t b, uint32_t kind) { return (2 * b + kind) * kCursorSet; }
// stats64[] layout
enum : uint32_t {
  kStNodes = 0,
  kStTris = 1,
  kStShNodes = 2,
  kStShTris = 3,
  kStNonfinite = 4,
  kStSpill = 5,    // node visits with spilled stack entries (both traversal kernels)
  kStWords = 8
};

struct DevScene {
  DevBvh bvh;
  const float4* mats;     // 3 per material
  const float4* lights;   // 4 per light: area (v0|area, v1, v2, Le); point (pos, -, -, I)
  const float* light_cdf;
  const float* light_pmf;
  uint32_t n_area, n_lights;
  DevTex tex;             // image textures (texels == nullptr: none)
};

struct CamParams {
  float m[12];
  float tan_half, aspect;
  uint32_t width, height;
};

struct PassParams {
  CamParams cam;
  uint32_t npix;
  uint32_t spp_in_pass;
  uint32_t sample0;
  uint32_t frame;
  uint32_t seed;
  uint32_t max_depth;
  uint32_t rr_depth;
  uint32_t dims;  // PCG32 draws reserved per sample
  // neural radiance cache (nrc == 0: plain path tracing)
  uint32_t nrc;
  uint32_t query_depth;   // depth (0 = first hit) at which non-training paths query
  uint32_t train_stride;  // path p trains when p % stride == offset ...
  uint32_t train_offset;
  uint32_t n_train;       // ... and p / stride < n_train
  float bmin[3];          // position normalisation to [0,1]^3
  float bscale[3];
};

struct Queues {
  float4* ro[2];   // origin, tmax
  float4* rd[2];   // direction, -
  float4* tp[2];   // throughput, pdf of the sampled direction
  float4* L[2];    // radiance so far
  uint4* rng[2];   // pcg state lo, hi, path id, depth
  uint2* hit;      // triangle slot, t bits
  float4* sho;     // shadow origin, tmax
  float4* shd;     // shadow direction, destination bits
  float4* shc;     // shadow contribution
  float4* Lbuf;    // radiance of finished paths, by path id
  const uint32_t* order;  // shading order of the hit queue (material sort), or null
  uint32_t* counters;
  uint32_t* cursors;
  unsigned long long* stats;
  // cache queries (compacted) and training slots (train index x depth)
  float4* qrec;    // 4 per query: (pos, roughness), (dir, -), (normal, -), (albedo, -)
  float4* qaux;    // (throughput, path id bits)
  uint32_t* qkey;  // Morton key of each query's position (sort input)
  float4* trec;    // 4 per training slot, as qrec
  float4* taux;    // 2 per training slot: (throughput, valid), (radiance after emission, -)
};

}  // namespace aqd
