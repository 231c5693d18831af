// Renderer: HIP kernels of the wavefront path tracer and their host-side
// orchestration behind the C ABI of include/aqua/aqua.h.
//
// One renderer = one device, one HIP stream; every launch of a frame is
// queued on that stream and ordered by it, and queue lengths are read by the
// kernels from device memory, so the host never waits inside a frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "aqua/aqua.h"
#include "device/kernels.h"
#include "device/sampler.h"
#include "device/traverse.h"
#include "host/bvh.h"
#include "host/internal.h"

using namespace aqd;

namespace {

constexpr int kBlock = 256;
static_assert(kBlock == aqd::kStackThreads, "traversal kernels run kStackThreads threads per block");

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Wave-aggregated queue append: every lane of the wave must call it; lanes
// with `want` get consecutive slots starting at *counter.
// Block-wide queue appends: each thread that wants a slot gets base +
// (wanting threads before it in the block); one atomic per block. Every
// thread of the block must call (the calls synchronise the block).
__device__ __forceinline__ uint32_t block_prefix(unsigned long long mask, const uint32_t* wave_tot, uint32_t w) {
  uint32_t p = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
  for (uint32_t k = 0; k < w; ++k) p += wave_tot[k];
  return p;
}

// Two queues at once: `pair` holds the first queue's length in its low word
// and the second's in its high word.
__device__ __forceinline__ void block_append2(bool a, bool b, unsigned long long* pair, uint32_t* ja,
                                              uint32_t* jb) {
  __shared__ uint32_t ta[kBlock / 64], tb[kBlock / 64];
  __shared__ unsigned long long base;
  const unsigned long long ma = __ballot(a), mb = __ballot(b);
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    ta[w] = uint32_t(__popcll(ma));
    tb[w] = uint32_t(__popcll(mb));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sb = 0;
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
      sa += ta[k];
      sb += tb[k];
    }
    base = (sa | sb) ? atomicAdd(pair, sa | (sb << 32)) : 0ull;
  }
  __syncthreads();
  *ja = uint32_t(base) + block_prefix(ma, ta, w);
  *jb = uint32_t(base >> 32) + block_prefix(mb, tb, w);
  __syncthreads();
}

__device__ __forceinline__ uint32_t block_append(bool a, uint32_t* counter) {
  __shared__ uint32_t ta[kBlock / 64];
  __shared__ uint32_t base;
  const unsigned long long ma = __ballot(a);
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) ta[w] = uint32_t(__popcll(ma));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sa = 0;
    for (uint32_t k = 0; k < kBlock / 64; ++k) sa += ta[k];
    base = sa ? atomicAdd(counter, sa) : 0u;
  }
  __syncthreads();
  uint32_t j = base + block_prefix(ma, ta, w);
  __syncthreads();
  return j;
}

// Queue counters of a pass (bounce 0 holds initial_rays) and the first
// cursor_words work cursors.
__global__ void k_reset(uint32_t* counters, uint32_t initial_rays, uint32_t* cursors, uint32_t cursor_words) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kCntWords + cursor_words;
       i += gridDim.x * blockDim.x) {
    if (i < kCntWords)
      counters[i] = (i == cnt_rays(0)) ? initial_rays : 0u;
    else
      cursors[i - kCntWords] = 0u;
  }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_zero_stats(unsigned long long* st) {
  uint32_t i = threadIdx.x;
  if (i < kStWords) st[i] = 0ull;
}

__device__ __forceinline__ v3 camera_dir(const CamParams& c, uint32_t px, uint32_t py, float u0,
                                         float u1) {
  float fx = (float(px) + u0) / float(c.width);
  float fy = (float(py) + u1) / float(c.height);
  float nx = 2.0f * fx - 1.0f;
  float ny = 1.0f - 2.0f * fy;
  float cx = (nx * c.tan_half) * c.aspect;
  float cy = ny * c.tan_half;
  v3 d = v3{(c.m[0] * cx + c.m[1] * cy) - c.m[2], (c.m[4] * cx + c.m[5] * cy) - c.m[6],
            (c.m[8] * cx + c.m[9] * cy) - c.m[10]};
  return normalize(d);
}

// Camera rays + fresh path state for samples [sample0, sample0 + S) of
// every pixel; path id = s_local * npix + pixel.
__global__ void k_camera(PassParams pp, uint32_t sampler, Queues q) {
  uint32_t n = pp.npix * pp.spp_in_pass;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t s_local = i / pp.npix;
    uint32_t pixel = i - s_local * pp.npix;
    uint32_t px = pixel % pp.cam.width, py = pixel / pp.cam.width;
    uint32_t sample = pp.sample0 + s_local;
    Pcg rng = pcg_seed(pp.seed, pixel);
    pcg_advance(rng, ((uint64_t(pp.frame) << 20) | uint64_t(sample)) * uint64_t(pp.dims));
    float u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = u01(pcg_next(rng));
    if (sampler) stratified2(pp.seed, pp.frame, pixel, 0u, sample, &u[0], &u[1]);
    v3 d = camera_dir(pp.cam, px, py, u[0], u[1]);
    q.ro[0][i] = make_float4(pp.cam.m[3], pp.cam.m[7], pp.cam.m[11], kFltMax);
    q.rd[0][i] = make_float4(d.x, d.y, d.z, 0.0f);
    q.tp[0][i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    q.L[0][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    q.rng[0][i] = make_uint4(uint32_t(rng.state), uint32_t(rng.state >> 32), i, 0u);
  }
}

__device__ __forceinline__ uint32_t pick_light(const float* cdf, uint32_t n, float u) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (u < cdf[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// One path vertex per thread: emission (MIS-weighted), then, below the
// depth limit, next-event estimation, BSDF sampling and Russian roulette.
// NRC: the cache is on (pp.nrc): query records and training slots; the plain
// tracer's instance carries none of their registers.
template <bool NRC>
__global__ void __launch_bounds__(kBlock)
    k_shade(DevScene sc, PassParams pp, uint32_t sampler, uint32_t b, Queues q) {
  const uint32_t c = b & 1u, nx = c ^ 1u;
  const uint32_t n = q.counters[cnt_rays(b)];
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    bool alive = false, shadow = false, query = false;
    float4 n_ro, n_rd, n_tp, s_o, s_d, s_c;
    float4 qr0, qr1, qr2, qr3, qa;
    uint4 n_rng;
    v3 L = v3{0.0f, 0.0f, 0.0f};
    uint32_t path_id = 0;
    if (valid) {
      const uint32_t src = q.order ? q.order[i] : i;
      float4 ro4 = q.ro[c][src], rd4 = q.rd[c][src], tp4 = q.tp[c][src], L4 = q.L[c][src];
      uint4 rg = q.rng[c][src];
      uint2 h = q.hit[src];
      v3 o = xyz(ro4), d = xyz(rd4), tp = xyz(tp4);
      float pdf_prev = tp4.w;
      L = xyz(L4);
      path_id = rg.z;
      uint32_t depth = rg.w;
      if (h.x != kMiss) {
        float t = __uint_as_float(h.y);
        const float4* tr = sc.bvh.tris + 3 * h.x;
        float4 p0 = tr[0], p1 = tr[1], p2 = tr[2];
        v3 v0 = xyz(p0), v1 = xyz(p1), v2 = xyz(p2);
        Mat m = load_mat(sc.mats, __float_as_uint(p1.w));
        v3 p = v3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
        if (m.tex) {
          float b1, b2;
          bary_of(v0, v1, v2, p, &b1, &b2);
          m.base = tex_base(sc.tex, h.x, m.tex - 1u, b1, b2);
        }
        v3 wo = neg(d);
        v3 ng = normalize(cross(sub(v1, v0), sub(v2, v0)));
        float cg = dot(ng, wo);
        if ((m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f) && cg > 0.0f) {
          float w = 1.0f;
          if (depth > 0u) {
            float pl = ((m.light_pdf * t) * t) / cg;
            float b2 = pdf_prev * pdf_prev, l2 = pl * pl;
            w = b2 / (b2 + l2);
          }
          L = v3{L.x + (tp.x * m.emission.x) * w, L.y + (tp.y * m.emission.y) * w,
                 L.z + (tp.z * m.emission.z) * w};
        }
        // ---- neural radiance cache: query (and end the path) or record a
        // training sample of the reflected radiance at this vertex
        if (NRC) {
          bool is_train = (path_id % pp.train_stride) == pp.train_offset && path_id / pp.train_stride < pp.n_train;
          uint32_t tslot = 0xFFFFFFFFu;
          if (!is_train && depth == pp.query_depth)
            query = true;
          else if (is_train && depth >= pp.query_depth && depth < pp.max_depth)
            tslot = (path_id / pp.train_stride) * (pp.max_depth + 1u) + depth;
          if (query || tslot != 0xFFFFFFFFu) {
            v3 nsq = cg < 0.0f ? neg(ng) : ng;
            float px = fminf(fmaxf((p.x - pp.bmin[0]) * pp.bscale[0], 0.0f), 1.0f);
            float py = fminf(fmaxf((p.y - pp.bmin[1]) * pp.bscale[1], 0.0f), 1.0f);
            float pz = fminf(fmaxf((p.z - pp.bmin[2]) * pp.bscale[2], 0.0f), 1.0f);
            qr0 = make_float4(px, py, pz, m.roughness);
            qr1 = make_float4(wo.x, wo.y, wo.z, 0.0f);
            qr2 = make_float4(nsq.x, nsq.y, nsq.z, 0.0f);
            qr3 = make_float4(m.base.x, m.base.y, m.base.z, 0.0f);
            qa = make_float4(tp.x, tp.y, tp.z, __uint_as_float(path_id));
            if (tslot != 0xFFFFFFFFu) {
              q.trec[4 * size_t(tslot) + 0] = qr0;
              q.trec[4 * size_t(tslot) + 1] = qr1;
              q.trec[4 * size_t(tslot) + 2] = qr2;
              q.trec[4 * size_t(tslot) + 3] = qr3;
              q.taux[2 * size_t(tslot) + 0] = make_float4(tp.x, tp.y, tp.z, 1.0f);
              q.taux[2 * size_t(tslot) + 1] = make_float4(L.x, L.y, L.z, 0.0f);
            }
          }
        }
        if (depth < pp.max_depth && !query) {
          uint32_t pixel = path_id % pp.npix;
          Pcg rng;
          rng.state = (uint64_t(rg.y) << 32) | uint64_t(rg.x);
          rng.inc = (uint64_t(pixel) << 1u) | 1u;
          float u[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) u[k] = u01(pcg_next(rng));
          if (sampler && depth == 0u) {
            const uint32_t sample = pp.sample0 + path_id / pp.npix;
            stratified2(pp.seed, pp.frame, pixel, 1u, sample, &u[1], &u[2]);
            stratified2(pp.seed, pp.frame, pixel, 2u, sample, &u[4], &u[5]);
            stratified2(pp.seed, pp.frame, pixel, 3u, sample, &u[0], &u[3]);
          }
          v3 ns = cg < 0.0f ? neg(ng) : ng;
          Bsdf bs = make_bsdf(m);
          // ---- next-event estimation ----
          if (sc.n_lights > 0u) {
            uint32_t li = pick_light(sc.light_cdf, sc.n_lights, u[0]);
            float psel = sc.light_pmf[li];
            const float4* lp = sc.lights + 4 * li;
            bool is_area = li < sc.n_area;
            v3 wi, Le;
            float d2, dist, pdf_l = 0.0f;
            bool ok;
            if (is_area) {
              float4 a0 = lp[0], a1 = lp[1], a2 = lp[2], a3 = lp[3];
              v3 lv0 = xyz(a0), lv1 = xyz(a1), lv2 = xyz(a2);
              float area = a0.w;
              Le = xyz(a3);
              float su = __builtin_sqrtf(u[1]);
              float b0 = 1.0f - su;
              float b1 = u[2] * su;
              float b2 = (1.0f - b0) - b1;
              v3 qp = v3{(lv0.x * b0 + lv1.x * b1) + lv2.x * b2, (lv0.y * b0 + lv1.y * b1) + lv2.y * b2,
                         (lv0.z * b0 + lv1.z * b1) + lv2.z * b2};
              v3 nl = normalize(cross(sub(lv1, lv0), sub(lv2, lv0)));
              v3 dv = sub(qp, p);
              d2 = dot(dv, dv);
              dist = __builtin_sqrtf(d2);
              wi = scl(dv, 1.0f / dist);
              float cl = -dot(nl, wi);
              ok = cl > 0.0f && d2 > 0.0f;
              if (ok) pdf_l = (psel * d2) / (area * cl);
            } else {
              float4 a0 = lp[0], a3 = lp[3];
              Le = xyz(a3);
              v3 dv = sub(xyz(a0), p);
              d2 = dot(dv, dv);
              dist = __builtin_sqrtf(d2);
              wi = scl(dv, 1.0f / dist);
              ok = d2 > 0.0f;
            }
            if (ok) {
              float pb, ci;
              v3 f = bsdf_eval(bs, ns, wo, wi, &pb, &ci);
              if (ci > 0.0f) {
                float sfac;
                if (is_area) {
                  float l2 = pdf_l * pdf_l, bb2 = pb * pb;
                  float w = l2 / (l2 + bb2);
                  sfac = (ci * w) / pdf_l;
                } else {
                  sfac = ci / (d2 * psel);
                }
                v3 con = v3{((tp.x * f.x) * Le.x) * sfac, ((tp.y * f.y) * Le.y) * sfac,
                            ((tp.z * f.z) * Le.z) * sfac};
                if (con.x > 0.0f || con.y > 0.0f || con.z > 0.0f) {
                  shadow = true;
                  v3 so = offset_ray(p, ns);
                  s_o = make_float4(so.x, so.y, so.z, dist * 0.9999f);
                  s_d = make_float4(wi.x, wi.y, wi.z, 0.0f);
                  s_c = make_float4(con.x, con.y, con.z, 0.0f);
                }
              }
            }
          }
          // ---- BSDF sampling + Russian roulette ----
          v3 wi, f;
          float pdf, ci;
          if (bsdf_sample(bs, ns, wo, u[3], u[4], u[5], &wi, &f, &pdf, &ci)) {
            float s = ci / pdf;
            v3 tpn = v3{tp.x * (f.x * s), tp.y * (f.y * s), tp.z * (f.z * s)};
            bool cont = tpn.x > 0.0f || tpn.y > 0.0f || tpn.z > 0.0f;
            if (cont && depth >= pp.rr_depth) {
              float qq = fmax2(tpn.x, fmax2(tpn.y, tpn.z));
              if (qq > 0.95f) qq = 0.95f;
              if (u[6] >= qq)
                cont = false;
              else
                tpn = v3{tpn.x / qq, tpn.y / qq, tpn.z / qq};
            }
            if (cont) {
              alive = true;
              v3 no = offset_ray(p, ns);
              n_ro = make_float4(no.x, no.y, no.z, kFltMax);
              n_rd = make_float4(wi.x, wi.y, wi.z, 0.0f);
              n_tp = make_float4(tpn.x, tpn.y, tpn.z, pdf);
              n_rng = make_uint4(uint32_t(rng.state), uint32_t(rng.state >> 32), path_id, depth + 1u);
            }
          }
        }
      }
    }
    uint32_t j, k;
    block_append2(alive, shadow, reinterpret_cast<unsigned long long*>(q.counters + cnt_rays(b + 1)), &j, &k);
    if (alive) {
      q.ro[nx][j] = n_ro;
      q.rd[nx][j] = n_rd;
      q.tp[nx][j] = n_tp;
      q.L[nx][j] = make_float4(L.x, L.y, L.z, 0.0f);
      q.rng[nx][j] = n_rng;
    } else if (valid) {
      q.Lbuf[path_id] = make_float4(L.x, L.y, L.z, 0.0f);
    }
    if (shadow) {
      uint32_t dst = alive ? j : (0x80000000u | path_id);
      s_d.w = __uint_as_float(dst);
      q.sho[k] = s_o;
      q.shd[k] = s_d;
      q.shc[k] = s_c;
    }
    if (NRC) {
      uint32_t qi = block_append(query, &q.counters[kCntQuery]);
      if (query) {
        q.qrec[4 * size_t(qi) + 0] = qr0;
        q.qrec[4 * size_t(qi) + 1] = qr1;
        q.qrec[4 * size_t(qi) + 2] = qr2;
        q.qrec[4 * size_t(qi) + 3] = qr3;
        q.qaux[qi] = qa;
        q.qkey[qi] = morton30(qr0.x, qr0.y, qr0.z);
      }
    }
  }
}

// ---- material sort of the hit queue ------------------------------------------
//
// Optional (aqua_renderer_opts.sort_materials): before shading, the paths of
// a bounce are put in order of the material they hit (misses last), so a
// shading wave runs one material's code and texture. A counting sort:
// k_mat_hist counts each bin, k_mat_scan turns the counts into bin starts
// (and clears the counts for the next bounce), k_mat_place hands each path a
// slot in its bin. The order within a bin is arbitrary; results do not
// depend on it (paths are independent; the film adds by path id).
constexpr uint32_t kMatBins = 256;  // materials 254 and up share bin 254; 255 = miss

__device__ __forceinline__ uint32_t mat_bin(const DevBvh& bvh, uint2 h) {
  if (h.x == kMiss) return kMatBins - 1u;
  return min(__float_as_uint(bvh.tris[3 * size_t(h.x) + 1].w), kMatBins - 2u);
}

__global__ void __launch_bounds__(kBlock) k_mat_hist(DevBvh bvh, const uint2* hit, const uint32_t* count,
                                                     uint32_t* bins) {
  __shared__ uint32_t h[kMatBins];
  for (uint32_t i = threadIdx.x; i < kMatBins; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&h[mat_bin(bvh, hit[i])], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kMatBins; i += blockDim.x)
    if (h[i]) atomicAdd(&bins[i], h[i]);
}

// One block of kMatBins threads.
__global__ void __launch_bounds__(kMatBins) k_mat_scan(uint32_t* bins, uint32_t* start) {
  __shared__ uint32_t s[kMatBins];
  const uint32_t t = threadIdx.x;
  const uint32_t v = bins[t];
  s[t] = v;
  __syncthreads();
  for (uint32_t o = 1; o < kMatBins; o <<= 1) {
    const uint32_t add = t >= o ? s[t - o] : 0u;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  start[t] = s[t] - v;
  bins[t] = 0u;
}

// Each block takes blockDim.x paths at a time: counts its paths per bin in
// LDS, claims a range of each bin it uses with one global atomic, then
// writes its paths' slots there.
__global__ void __launch_bounds__(kBlock) k_mat_place(DevBvh bvh, const uint2* hit, const uint32_t* count,
                                                      uint32_t* start, uint32_t* order) {
  __shared__ uint32_t h[kMatBins], base[kMatBins];
  const uint32_t n = *count;
  for (uint32_t first = blockIdx.x * blockDim.x; first < n; first += gridDim.x * blockDim.x) {
    for (uint32_t i = threadIdx.x; i < kMatBins; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const uint32_t i = first + threadIdx.x;
    uint32_t bin = 0u, rank = 0u;
    if (i < n) {
      bin = mat_bin(bvh, hit[i]);
      rank = atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kMatBins; k += blockDim.x)
      if (h[k]) base[k] = atomicAdd(&start[k], h[k]);
    __syncthreads();
    if (i < n) order[base[bin] + rank] = i;
    __syncthreads();
  }
}

// Adds throughput x max(cache, 0) to the radiance of each querying path.
__global__ void k_nrc_apply(const float4* qaux, const float* qout, const uint32_t* count, float4* Lbuf) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 a = qaux[i];
    uint32_t path = __float_as_uint(a.w);
    float r = fmaxf(qout[3 * size_t(i)], 0.0f), g = fmaxf(qout[3 * size_t(i) + 1], 0.0f),
          bl = fmaxf(qout[3 * size_t(i) + 2], 0.0f);
    float4 l = Lbuf[path];
    Lbuf[path] = make_float4(l.x + a.x * r, l.y + a.y * g, l.z + a.z * bl, 0.0f);
  }
}

// Training targets: the radiance a training path gathered after a recorded
// vertex, divided by the throughput up to it = the reflected radiance the
// cache should predict there. Valid samples are compacted into the batch.
__global__ void k_nrc_targets(const float4* trec, const float4* taux, uint32_t n_slots, uint32_t slots_per_path,
                              uint32_t stride, uint32_t offset, const float4* Lbuf, float4* out_rec,
                              float4* out_tgt, uint32_t* counter) {
  for (uint32_t base = blockIdx.x * blockDim.x; base < n_slots; base += gridDim.x * blockDim.x) {
    uint32_t s = base + threadIdx.x;
    bool ok = false;
    float4 tgt = make_float4(0, 0, 0, 0);
    if (s < n_slots) {
      float4 T = taux[2 * size_t(s)];
      if (T.w == 1.0f) {
        float4 Lpre = taux[2 * size_t(s) + 1];
        uint32_t path = (s / slots_per_path) * stride + offset;
        float4 Lf = Lbuf[path];
        float tv[3] = {T.x, T.y, T.z}, lf[3] = {Lf.x, Lf.y, Lf.z}, lp[3] = {Lpre.x, Lpre.y, Lpre.z}, o[3];
        ok = true;
        for (int c = 0; c < 3; ++c) {
          o[c] = tv[c] > 1e-7f ? fmaxf(lf[c] - lp[c], 0.0f) / tv[c] : 0.0f;
          if (!__builtin_isfinite(o[c])) ok = false;
        }
        tgt = make_float4(o[0], o[1], o[2], 0.0f);
      }
    }
    uint32_t j = block_append(ok, counter);
    if (ok) {
      for (int k = 0; k < 4; ++k) out_rec[4 * size_t(j) + k] = trec[4 * size_t(s) + k];
      out_tgt[j] = tgt;
    }
  }
}

__global__ void k_clear_taux(float4* taux, uint32_t n_slots) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n_slots) taux[2 * size_t(s)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ bool finite3(float4 v) {
  return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

// Adds the pass's samples of each pixel to the film in sample order.
__global__ void k_film32(PassParams pp, const float4* Lbuf, float4* film, unsigned long long* stats) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pp.npix) return;
  float4 acc = film[p];
  uint32_t bad = 0;
  for (uint32_t s = 0; s < pp.spp_in_pass; ++s) {
    float4 l = Lbuf[size_t(s) * pp.npix + p];
    if (!finite3(l)) {
      ++bad;
      continue;
    }
    acc.x = acc.x + l.x;
    acc.y = acc.y + l.y;
    acc.z = acc.z + l.z;
  }
  film[p] = acc;
  if (bad) atomicAdd(&stats[kStNonfinite], (unsigned long long)bad);
}

__global__ void k_film64(PassParams pp, const float4* Lbuf, double* film, unsigned long long* stats) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pp.npix) return;
  double ax = film[3 * size_t(p)], ay = film[3 * size_t(p) + 1], az = film[3 * size_t(p) + 2];
  uint32_t bad = 0;
  for (uint32_t s = 0; s < pp.spp_in_pass; ++s) {
    float4 l = Lbuf[size_t(s) * pp.npix + p];
    if (!finite3(l)) {
      ++bad;
      continue;
    }
    ax += double(l.x);
    ay += double(l.y);
    az += double(l.z);
  }
  film[3 * size_t(p)] = ax;
  film[3 * size_t(p) + 1] = ay;
  film[3 * size_t(p) + 2] = az;
  if (bad) atomicAdd(&stats[kStNonfinite], (unsigned long long)bad);
}

__global__ void k_zero_f4(float4* p, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__global__ void k_zero_f64(double* p, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Hit slots -> triangle ids (load order) and distances.
__global__ void k_hits_out(const uint2* hit, const float4* tris, uint32_t n, uint32_t* ids, float* ts) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint2 h = hit[i];
  ids[i] = h.x == kMiss ? kMiss : __float_as_uint(tris[3 * h.x].w);
  ts[i] = __uint_as_float(h.y);
}

// relMSE / relRMSE partial sums in FP64, one triple per block.
__global__ void k_err_partial(const float4* film, float inv_spp, const double* ref, double inv_ref_spp,
                              uint32_t npix, double* partial) {
  __shared__ double sh[3][kBlock];
  double a = 0.0, bsum = 0.0, c = 0.0;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    float4 f = film[p];
    float xs[3] = {f.x * inv_spp, f.y * inv_spp, f.z * inv_spp};
    for (int k = 0; k < 3; ++k) {
      double x = double(xs[k]);
      double y = ref[3 * size_t(p) + k] * inv_ref_spp;
      double dxy = x - y;
      a += dxy * dxy / (y * y + 1e-2);
      bsum += dxy * dxy;
      c += y * y;
    }
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = bsum;
  sh[2][threadIdx.x] = c;
  __syncthreads();
  for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) partial[3 * blockIdx.x + k] = sh[k][0];
}

__global__ void k_err_final(const double* partial, uint32_t nblocks, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0, c = 0.0;
  for (uint32_t i = 0; i < nblocks; ++i) {
    a += partial[3 * i];
    b += partial[3 * i + 1];
    c += partial[3 * i + 2];
  }
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

// Display encoding thresholds: code k + 1 starts at the smallest float x
// whose sRGB encoding reaches (k + 0.5) / 255 (computed on the host in
// double), so the 8-bit code of x is the number of thresholds <= x: the
// correctly rounded round(255 * srgb(x)) with no transcendental on the GPU.
struct ToneLut {
  float t[256];
};

__global__ void k_tonemap(const float4* film, float inv_spp, float scale, uint32_t npix, ToneLut lut,
                          uint32_t* out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float4 f = film[p];
  float c[3] = {(f.x * inv_spp) * scale, (f.y * inv_spp) * scale, (f.z * inv_spp) * scale};
  uint32_t px = 0xFF000000u;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t lo = 0, hi = 255;  // code = #{thresholds <= c}, thresholds ascending
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (c[k] >= lut.t[mid])
        lo = mid + 1;
      else
        hi = mid;
    }
    px |= lo << (8 * k);
  }
  out[p] = px;
}

// Elementwise FP32 arithmetic of the device for host-side comparison:
// out[5*i + {0..4}] = a/b, sqrt(a), fma(a, b, c), sin(2 pi u), cos(2 pi u).
__global__ void k_fp_probe(const float* a, const float* b, const float* cc, const float* u, uint32_t n,
                           float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s, c;
  sincos_2pi(u[i], &s, &c);
  out[5 * i + 0] = a[i] / b[i];
  out[5 * i + 1] = __builtin_sqrtf(a[i]);
  out[5 * i + 2] = __builtin_fmaf(a[i], b[i], cc[i]);
  out[5 * i + 3] = s;
  out[5 * i + 4] = c;
}

// ---- allocation-free closest hit (aqua_trace_closest_noalloc) --------------
// Its scratch: the queue counters (k_reset writes them, with the ray count,
// at every call), one set of work cursors (cleared likewise) and the
// traversal spill stacks of its fixed grid of kNaGrid blocks.
constexpr uint32_t kNaGrid = 512;
constexpr size_t kNaCounterWord = 0;     // kCntWords counters
constexpr size_t kNaCursorWord = 512;    // kCursorSet cursor words
constexpr size_t kNaSpillWord = 1024;    // kSpillStack x kNaGrid x kBlock words
constexpr size_t kNaWords = kNaSpillWord + size_t(kSpillStack) * kNaGrid * kBlock;
static_assert(kNaCounterWord + kCntWords <= kNaCursorWord && kNaCursorWord + kCursorSet <= kNaSpillWord,
              "scratch regions overlap");
static_assert(kNaGrid % 8 == 0, "traversal grids are a multiple of 8 blocks (one work range per XCD)");
// Largest ray count one call takes (32 GiB of rays); the work ranges index
// rays in 32 bits.
constexpr uint32_t kNaMaxRays = 1u << 30;

}  // namespace

// Device scratch of aqua_trace_closest_noalloc: 33.6 MB, held by every GPU
// the library runs on until the process ends. Host code never uses this
// name as an address (it is a host-side stand-in): the per-device address
// comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) uint32_t scratch[kNaWords];

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return aqua::fail(AQUA_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
  } while (0)

namespace aqua {
hipError_t query_sort_storage(uint32_t n, size_t* bytes);
hipError_t query_sort(void* temp, size_t temp_bytes, const float4* rec, const uint32_t* d_count, uint32_t n,
                      uint32_t* keys, uint32_t* idx, hipStream_t s, const uint32_t** perm);
}  // namespace aqua

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t count) {
    if (count <= n && p) return AQUA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return aqua::fail(e == hipErrorOutOfMemory ? AQUA_ERR_OOM : AQUA_ERR_HIP,
                        "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    n = count;
    return AQUA_OK;
  }
  int upload(const T* src, size_t count, hipStream_t s) {
    if (int e = ensure(count)) return e;
    if (count == 0) return AQUA_OK;
    HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return AQUA_OK;
  }
};

// One set of path queues with its counters, work cursors, traversal spill
// stacks and streams. A pass runs on one lane; without per-kernel timing,
// consecutive passes of fewer than 16 M paths alternate between two lanes and
// run concurrently (only their film accumulations are ordered), so one pass's
// memory-bound shading overlaps the other's traversal.
struct Lane {
  hipStream_t s = nullptr;   // camera, closest hit, shading (and the film, unless on s2)
  // Shadow rays of bounce b run on s2, overlapped with the closest-hit rays
  // of bounce b + 1 (independent work that fills each other's tail); the
  // fork / join events order them against the shading kernels.
  hipStream_t s2 = nullptr;
  std::vector<hipEvent_t> fork_ev, join_ev;
  hipEvent_t film_done = nullptr, done = nullptr;
  // cache queries on s2 beside the training paths' last bounces; tail_done:
  // those bounces are traced (the film, then on s2, waits for it)
  hipEvent_t q_fork = nullptr, q_done = nullptr, tail_done = nullptr;
  size_t cap = 0;
  DevBuf<float4> ro[2], rd[2], tp[2], L[2], sho, shd, shc, Lbuf;
  DevBuf<uint4> rng[2];
  DevBuf<uint2> hit;
  DevBuf<uint32_t> counters, cursors;
  DevBuf<uint32_t> spill;   // traversal spill stacks (kSpillStack words per grid thread)
  DevBuf<uint32_t> spill2;  // the same for the shadow kernels, which may run concurrently
  // cache queries of this lane's pass: records, (throughput, path), outputs,
  // Morton sort buffers
  size_t nrc_cap = 0;
  DevBuf<float4> qrec, qaux;
  DevBuf<float> qout;
  DevBuf<uint32_t> skey, sidx;  // Morton keys, slot order (sort double buffers)
  DevBuf<uint8_t> stemp;
  // material sort: shading order, bin counts (zero between bounces), bin starts
  size_t order_cap = 0;
  DevBuf<uint32_t> order, mbins, mstart;

  int ensure_order(size_t n, hipStream_t st) {
    if (!mbins.p) {
      if (int e = mbins.ensure(kMatBins)) return e;
      if (int e = mstart.ensure(kMatBins)) return e;
      k_zero_u32<<<1, kMatBins, 0, st>>>(mbins.p, kMatBins);
    }
    if (n <= order_cap) return AQUA_OK;
    if (int e = order.ensure(n)) return e;
    order_cap = n;
    return AQUA_OK;
  }

  int ensure_nrc(size_t n) {
    if (n <= nrc_cap) return AQUA_OK;
    if (int e = qrec.ensure(4 * n)) return e;
    if (int e = qaux.ensure(n)) return e;
    if (int e = qout.ensure(3 * n)) return e;
    if (int e = skey.ensure(2 * n)) return e;
    if (int e = sidx.ensure(2 * n)) return e;
    nrc_cap = n;
    return AQUA_OK;
  }

  int init(int persistent_blocks, int max_bounces) {
    if (s) return AQUA_OK;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    fork_ev.resize(size_t(max_bounces) + 1);
    join_ev.resize(size_t(max_bounces) + 1);
    for (hipEvent_t& e : fork_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (hipEvent_t& e : join_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&film_done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&q_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&q_done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&tail_done, hipEventDisableTiming));
    if (int e = counters.ensure(kCntWords)) return e;
    if (int e = cursors.ensure(kCursorWords)) return e;
    if (int e = spill.ensure(size_t(kSpillStack) * persistent_blocks * kBlock)) return e;
    if (int e = spill2.ensure(size_t(kSpillStack) * persistent_blocks * kBlock)) return e;
    return AQUA_OK;
  }

  int ensure(size_t n) {
    if (n <= cap) return AQUA_OK;
    for (int k = 0; k < 2; ++k) {
      if (int e = ro[k].ensure(n)) return e;
      if (int e = rd[k].ensure(n)) return e;
      if (int e = tp[k].ensure(n)) return e;
      if (int e = L[k].ensure(n)) return e;
      if (int e = rng[k].ensure(n)) return e;
    }
    if (int e = hit.ensure(n)) return e;
    if (int e = sho.ensure(n)) return e;
    if (int e = shd.ensure(n)) return e;
    if (int e = shc.ensure(n)) return e;
    if (int e = Lbuf.ensure(n)) return e;
    cap = n;
    return AQUA_OK;
  }

  ~Lane() {
    if (s) (void)hipStreamSynchronize(s);
    if (s2) (void)hipStreamSynchronize(s2);
    for (hipEvent_t e : fork_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
    if (film_done) (void)hipEventDestroy(film_done);
    if (done) (void)hipEventDestroy(done);
    if (q_fork) (void)hipEventDestroy(q_fork);
    if (q_done) (void)hipEventDestroy(q_done);
    if (tail_done) (void)hipEventDestroy(tail_done);
    if (s2) (void)hipStreamDestroy(s2);
    if (s) (void)hipStreamDestroy(s);
  }
};

struct aqua_renderer {
  int device = 0;
  hipStream_t stream = nullptr;  // lane[0].s
  Lane lane[2];
  aqua_integrator_cfg cfg{};
  aqua_camera camera{};
  aqua_renderer_opts opts{};
  uint32_t n_tris = 0, n_area = 0, n_lights = 0;
  float delta = 0.0f;
  int persistent_blocks = 2048;  // resident blocks of the traversal kernels
  uint32_t* na_scratch = nullptr;  // this device's `scratch` (aqua_trace_closest_noalloc)
  std::vector<uint32_t> tri_order;
  DevBuf<float4> tris, nodes, mats, lights;
  DevBuf<float> cdf, pmf;
  // image textures: texels of every image, (first texel, w, h) per image,
  // texcoords per triangle slot, sRGB decode table
  DevBuf<uint32_t> texels;
  DevBuf<uint4> tex_info;
  DevBuf<float4> tuv;
  DevBuf<float> tex_lut;
  DevBuf<unsigned long long> stats;
  // pinned host copies of each pass's counters (read after the frame)
  uint32_t* pass_counters = nullptr;
  float* loss_host = nullptr;  // pinned: the frame's last training loss (a sum over loss_rows x 3)
  uint32_t loss_rows = 0;
  bool presort = false;  // the pass queues the sort of its training rows (one step on all of them)
  size_t pass_counters_cap = 0;
  // films
  uint32_t film_w = 0, film_h = 0, film_spp = 0, oracle_spp = 0;
  DevBuf<float4> film;
  DevBuf<double> film64;
  DevBuf<double> err_partial;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> kev;
  hipEvent_t nev[4] = {};  // around the cache queries, around the cache training
  // neural radiance cache
  aqua_nrc* nrc = nullptr;
  bool nrc_enabled = true;
  uint64_t nrc_steps = 0;  // Adam steps the renderer has taken
  uint64_t ck_nrc_steps = 0;  // ... and cache on / off at the last checkpoint
  bool ck_nrc_enabled = true;
  float bmin[3] = {0, 0, 0}, bscale[3] = {1, 1, 1};
  DevBuf<float4> trec, taux, brec, btgt;  // training slots of a frame's first pass, batch
  uint32_t batch_n = 0;
  // opts.train_ahead: the frame's batch waits in pbrec / pbtgt (pending_nt
  // rows) and is trained at the start of the next frame on tstream, beside
  // that frame's camera rays and first bounces; its queries wait for the
  // update (train_done), so they see the same weights as without
  DevBuf<float4> pbrec, pbtgt, ck_brec, ck_btgt;
  uint32_t pending_nt = 0, ck_pending_nt = 0;
  hipStream_t tstream = nullptr;
  hipEvent_t train_go = nullptr, train_done = nullptr;
  bool train_launched = false;
  uint32_t train_nt = 0;           // training samples of the frame's first pass
  hipEvent_t pass0_done = nullptr;  // two lanes: the frame's first pass is done

  ~aqua_renderer() {
    (void)hipSetDevice(device);
    if (tstream) {
      (void)hipStreamSynchronize(tstream);
      (void)hipStreamDestroy(tstream);
    }
    if (train_go) (void)hipEventDestroy(train_go);
    if (train_done) (void)hipEventDestroy(train_done);
    if (lane[1].s) (void)hipStreamSynchronize(lane[1].s);
    if (stream) (void)hipStreamSynchronize(stream);
    if (pass_counters) (void)hipHostFree(pass_counters);
    if (loss_host) (void)hipHostFree(loss_host);
    if (pass0_done) (void)hipEventDestroy(pass0_done);
    if (nrc) aqua_nrc_destroy(nrc);
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    for (hipEvent_t e : nev)
      if (e) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }

  // Blocks of a persistent traversal launch over n rays: a multiple of 8
  // (one range per XCD), at most the resident count.
  int trav_grid(size_t n) const {
    int g = int(std::min<size_t>((n + kBlock - 1) / kBlock, size_t(persistent_blocks)));
    return std::min(std::max(8, (g + 7) & ~7), persistent_blocks);
  }

  DevScene dev_scene() const {
    DevScene s;
    s.bvh.nodes = nodes.p;
    s.bvh.top = nullptr;
    s.bvh.n_top = uint32_t(std::min<size_t>(kTopNodes, nodes.n / 4));  // 4 float4 per node
    s.bvh.tris = tris.p;
    s.bvh.delta = delta;
    s.bvh.spill = lane[0].spill.p;
    s.mats = mats.p;
    s.lights = lights.p;
    s.light_cdf = cdf.p;
    s.light_pmf = pmf.p;
    s.n_area = n_area;
    s.n_lights = n_lights;
    s.tex.texels = texels.p;
    s.tex.info = tex_info.p;
    s.tex.uv = tuv.p;
    s.tex.lut = tex_lut.p;
    return s;
  }

  Queues queues(int li = 0) {
    const Lane& l = lane[li];
    Queues q;
    for (int k = 0; k < 2; ++k) {
      q.ro[k] = l.ro[k].p;
      q.rd[k] = l.rd[k].p;
      q.tp[k] = l.tp[k].p;
      q.L[k] = l.L[k].p;
      q.rng[k] = l.rng[k].p;
    }
    q.hit = l.hit.p;
    q.sho = l.sho.p;
    q.shd = l.shd.p;
    q.shc = l.shc.p;
    q.Lbuf = l.Lbuf.p;
    q.order = nullptr;
    q.counters = l.counters.p;
    q.cursors = l.cursors.p;
    q.stats = stats.p;
    q.qrec = l.qrec.p;
    q.qaux = l.qaux.p;
    q.qkey = l.skey.p;
    q.trec = trec.p;
    q.taux = taux.p;
    return q;
  }

  // Lane li's queues (and the cache's query buffers) for n paths.
  int ensure_queues(size_t n, int li = 0) {
    if (int e = lane[li].ensure(n)) return e;
    if (nrc)
      if (int e = lane[li].ensure_nrc(n)) return e;
    return AQUA_OK;
  }
};

namespace aqua {
int nrc_infer_device(aqua_nrc* nrc, uint32_t n_max, const float4* d_rec, float* d_out, uint32_t precision,
                     hipStream_t s, const uint32_t* d_count, const uint32_t* d_perm);
int nrc_train_device(aqua_nrc* nrc, uint32_t n, const float4* d_rec, const float4* d_tgt, uint32_t steps,
                     aqua_nrc_stats* stats, hipStream_t s, hipEvent_t before_update = nullptr);
int nrc_loss_async(aqua_nrc* nrc, float* host_dst, hipStream_t s);
int nrc_presort_rows(aqua_nrc* nrc, uint32_t n_max, const float4* d_rec, const uint32_t* d_count, hipStream_t s);
void nrc_presort_drop(aqua_nrc* nrc);
}  // namespace aqua

namespace {

int grid_for(size_t n) {
  size_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return int(g);
}

CamParams cam_params(const aqua_camera& c) {
  CamParams p;
  memcpy(p.m, c.cam_to_world, sizeof p.m);
  p.tan_half = c.tan_half_fov;
  p.aspect = c.aspect;
  p.width = c.width;
  p.height = c.height;
  return p;
}

int check_camera(const aqua_camera* c) {
  if (c->width == 0 || c->height == 0 || c->width > 16384 || c->height > 16384)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "camera resolution %ux%u out of range", c->width, c->height);
  for (float v : c->cam_to_world)
    if (!std::isfinite(v)) return aqua::fail(AQUA_ERR_NONFINITE, "camera matrix is not finite");
  if (!std::isfinite(c->tan_half_fov) || !(c->tan_half_fov > 0.0f) || !std::isfinite(c->aspect) ||
      !(c->aspect > 0.0f))
    return aqua::fail(AQUA_ERR_INVALID_ARG, "camera fov / aspect out of range");
  return AQUA_OK;
}

int select_device(int device, int* cu_count = nullptr) {
  if (device < 0)
    return aqua::fail(AQUA_ERR_INVALID_ARG,
                      "device %d: the renderer runs on a GPU only (there is no CPU path)", device);
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return aqua::fail(AQUA_ERR_NO_DEVICE, "no HIP device available");
  if (device >= count)
    return aqua::fail(AQUA_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, count);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return aqua::fail(AQUA_ERR_NO_DEVICE, "device %d is %s; this build targets gfx950", device,
                      prop.gcnArchName);
  HIPCHK(hipSetDevice(device));
  return AQUA_OK;
}

// Device address of `scratch` on `device` (which must be current): looked up
// once per device, then kept. Each GPU holds its own copy of the array.
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[64] = {};

int noalloc_scratch(int device, uint32_t** out) {
  if (device < 0 || device >= 64) return aqua::fail(AQUA_ERR_NO_DEVICE, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return aqua::fail(AQUA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[device];
  return AQUA_OK;
}

// Ray counts of one pass from its counters.
void add_pass_counts(aqua_frame_stats* st, const uint32_t* hc, const PassParams& pp) {
  if (pp.nrc) st->nrc_queries += hc[kCntQuery];
  st->camera_rays += hc[cnt_rays(0)];
  for (uint32_t b = 1; b <= pp.max_depth; ++b) st->extension_rays += hc[cnt_rays(b)];
  for (uint32_t b = 0; b < pp.max_depth; ++b) st->shadow_rays += hc[cnt_shadow(b)];
  st->n_intersect_launches += pp.max_depth + 1;
}

// One pass on lane li: S samples of every pixel through all bounces, then
// the film (after film_wait, if given: the previous pass's film on the other
// lane). With counters_out the pass's counters are copied there
// asynchronously and the caller adds them to the stats after the frame;
// otherwise the pass is waited for and its counts are added here.
int run_pass(aqua_renderer* r, int li, const PassParams& pp, bool fp64, aqua_frame_stats* st, bool count,
             hipEvent_t film_wait = nullptr, uint32_t* counters_out = nullptr) {
  Lane& lane = r->lane[li];
  hipStream_t s = lane.s;
  uint32_t n = pp.npix * pp.spp_in_pass;
  DevScene sc = r->dev_scene();
  sc.bvh.spill = lane.spill.p;
  Queues q = r->queues(li);
  const int g_pass = grid_for(n);
  const uint32_t n_slots = pp.nrc ? pp.n_train * (pp.max_depth + 1u) : 0u;
  // the lane's previous film may still be reading its queues on s2
  HIPCHK(hipStreamWaitEvent(s, lane.film_done, 0));
  k_reset<<<64, kBlock, 0, s>>>(q.counters, n, q.cursors, cursor_set(pp.max_depth + 1, 0));
  if (n_slots) k_clear_taux<<<int((n_slots + kBlock - 1) / kBlock), kBlock, 0, s>>>(q.taux, n_slots);
  k_camera<<<g_pass, kBlock, 0, s>>>(pp, r->cfg.sampler, q);
  bool timing = r->opts.timing != 0;
  const bool sort_mats = r->opts.sort_materials != 0;
  if (sort_mats)
    if (int e = lane.ensure_order(n, s)) return e;
  // Without per-kernel timing, the shadow rays of bounce b (stream2) overlap
  // the closest-hit rays of bounce b + 1; shading b + 1 waits for them (they
  // add to the radiance it moves, and it rewrites the shadow queue).
  const bool overlap = !timing;
  DevBvh shadow_bvh = sc.bvh;
  shadow_bvh.spill = lane.spill2.p;
  hipStream_t s2 = lane.s2;
  // Cache queries: sort by position, infer, then add throughput x prediction
  // to each path; the query count stays on the device (no host round trip).
  auto cache_queries = [&](hipStream_t qs) -> int {
    const uint32_t* perm = nullptr;
    size_t need = 0;
    HIPCHK(aqua::query_sort_storage(n, &need));
    if (2 * size_t(n) > lane.skey.n || 2 * size_t(n) > lane.sidx.n)
      return aqua::fail(AQUA_ERR_HIP, "query sort buffers too small for %u queries", n);
    if (int e = lane.stemp.ensure(need)) return e;
    // k_shade wrote the queries' keys; the sort orders the first *count slots
    HIPCHK(aqua::query_sort(lane.stemp.p, lane.stemp.n, nullptr, q.counters + kCntQuery, n, lane.skey.p,
                            lane.sidx.p, qs, &perm));
    if (r->train_launched) HIPCHK(hipStreamWaitEvent(qs, r->train_done, 0));  // this frame's update
    if (int e = aqua::nrc_infer_device(r->nrc, n, lane.qrec.p, lane.qout.p, r->cfg.nrc_precision, qs,
                                       q.counters + kCntQuery, perm))
      return e;
    k_nrc_apply<<<grid_for(n), kBlock, 0, qs>>>(lane.qaux.p, lane.qout.p, q.counters + kCntQuery, q.Lbuf);
    return AQUA_OK;
  };
  // Without timing, the queries are final once the query vertex is shaded:
  // they run on s2 while s traces the training paths' remaining bounces
  // (whose shadow rays then stay on s). The apply touches only query paths'
  // radiance, the training paths only their own.
  const bool qfork = overlap && pp.nrc && pp.query_depth <= pp.max_depth;
  auto shadow_on_s2 = [&](uint32_t b) { return overlap && !(qfork && b >= pp.query_depth); };
  bool queries_forked = false;
  // Past the query vertex only the training paths are alive (query paths
  // end there and shoot no shadow ray): those bounces launch grids sized for
  // them, not for the whole pass (the kernels loop over the device counts, so
  // a grid's size changes no result).
  auto alive_bound = [&](uint32_t b, bool shadow) -> uint32_t {
    const bool tail = pp.nrc && (shadow ? b >= pp.query_depth : b > pp.query_depth);
    return tail ? std::min(n, pp.n_train) : n;
  };
  for (uint32_t b = 0; b <= pp.max_depth; ++b) {
    uint32_t c = b & 1u;
    const int g = grid_for(alive_bound(b, false));
    const int gp = r->trav_grid(alive_bound(b, false));
    const int gps = r->trav_grid(alive_bound(b, true));
    const uint32_t* cnt = q.counters + cnt_rays(b);
    uint32_t* cur = q.cursors + cursor_set(b, 0);
    hipEvent_t* ev = &r->kev[4 * b];
    if (timing) HIPCHK(hipEventRecord(ev[0], s));
    if (count)
      k_intersect<true><<<gp, kBlock, 0, s>>>(sc.bvh, QueueIO{q.ro[c], q.rd[c], q.hit}, cnt, cur, q.stats);
    else
      k_intersect<false><<<gp, kBlock, 0, s>>>(sc.bvh, QueueIO{q.ro[c], q.rd[c], q.hit}, cnt, cur, q.stats);
    if (timing) HIPCHK(hipEventRecord(ev[1], s));
    Queues qs = q;
    if (sort_mats) {
      k_mat_hist<<<std::min(g, 1024), kBlock, 0, s>>>(sc.bvh, q.hit, cnt, lane.mbins.p);
      k_mat_scan<<<1, kMatBins, 0, s>>>(lane.mbins.p, lane.mstart.p);
      k_mat_place<<<g, kBlock, 0, s>>>(sc.bvh, q.hit, cnt, lane.mstart.p, lane.order.p);
      qs.order = lane.order.p;
    }
    if (b > 0 && shadow_on_s2(b - 1)) HIPCHK(hipStreamWaitEvent(s, lane.join_ev[b - 1], 0));
    if (pp.nrc)
      k_shade<true><<<g, kBlock, 0, s>>>(sc, pp, r->cfg.sampler, b, qs);
    else
      k_shade<false><<<g, kBlock, 0, s>>>(sc, pp, r->cfg.sampler, b, qs);
    if (timing) HIPCHK(hipEventRecord(ev[2], s));
    if (qfork && b == pp.query_depth) {
      HIPCHK(hipEventRecord(lane.q_fork, s));
      HIPCHK(hipStreamWaitEvent(s2, lane.q_fork, 0));
      if (int e = cache_queries(s2)) return e;
      HIPCHK(hipEventRecord(lane.q_done, s2));
      queries_forked = true;
    }
    if (b < pp.max_depth) {
      hipStream_t ss = s;
      if (shadow_on_s2(b)) {
        HIPCHK(hipEventRecord(lane.fork_ev[b], s));
        HIPCHK(hipStreamWaitEvent(s2, lane.fork_ev[b], 0));
        ss = s2;
      }
      if (count)
        k_shadow<true><<<gps, kBlock, 0, ss>>>(shadow_bvh, ShadowIO{q, (b & 1u) ^ 1u}, q.counters + cnt_shadow(b),
                                               q.cursors + cursor_set(b, 1), q.stats);
      else
        k_shadow<false><<<gps, kBlock, 0, ss>>>(shadow_bvh, ShadowIO{q, (b & 1u) ^ 1u}, q.counters + cnt_shadow(b),
                                                q.cursors + cursor_set(b, 1), q.stats);
      if (shadow_on_s2(b)) HIPCHK(hipEventRecord(lane.join_ev[b], s2));
    }
    if (timing) HIPCHK(hipEventRecord(ev[3], s));
  }
  if (pp.max_depth > 0 && shadow_on_s2(pp.max_depth - 1))
    HIPCHK(hipStreamWaitEvent(s, lane.join_ev[pp.max_depth - 1], 0));
  HIPCHK(hipGetLastError());
  float ms_nrc = 0.0f;
  uint32_t n_query = 0;
  hipStream_t fs = s;  // the film's stream
  if (queries_forked) {
    // the film follows the apply on s2 once the training paths are traced;
    // s goes on to their targets, and the frame's training step can start
    // while the queries are still inferred (it only reads the weights; its
    // update waits for the film)
    HIPCHK(hipEventRecord(lane.tail_done, s));
    HIPCHK(hipStreamWaitEvent(s2, lane.tail_done, 0));
    fs = s2;
  } else if (pp.nrc) {
    if (timing) HIPCHK(hipEventRecord(r->nev[0], s));
    if (int e = cache_queries(s)) return e;
    if (timing) HIPCHK(hipEventRecord(r->nev[1], s));
  }
  int gf = int((pp.npix + kBlock - 1) / kBlock);
  if (film_wait) HIPCHK(hipStreamWaitEvent(fs, film_wait, 0));
  if (fp64)
    k_film64<<<gf, kBlock, 0, fs>>>(pp, q.Lbuf, r->film64.p, q.stats);
  else
    k_film32<<<gf, kBlock, 0, fs>>>(pp, q.Lbuf, r->film.p, q.stats);
  HIPCHK(hipEventRecord(lane.film_done, fs));
  if (pp.nrc && n_slots) {
    // training targets of this pass's suffix paths (into the batch); the
    // Adam steps run once the frame's passes are done (nrc_train_frame)
    k_nrc_targets<<<std::min(grid_for(n_slots), 4096), kBlock, 0, s>>>(
        q.trec, q.taux, n_slots, pp.max_depth + 1u, pp.train_stride, pp.train_offset, q.Lbuf, r->brec.p,
        r->btgt.p, q.counters + kCntTrain);
    // the frame's single Adam step takes all these rows: their sort goes in
    // stream order now, ahead of the count read below
    if (r->presort)
      if (int e = aqua::nrc_presort_rows(r->nrc, n_slots, r->brec.p, q.counters + kCntTrain, s)) return e;
  }
  HIPCHK(hipGetLastError());
  if (counters_out) {
    HIPCHK(hipMemcpyAsync(counters_out, q.counters, kCntWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(lane.done, s));
    return AQUA_OK;
  }
  // Per-pass counters (ray counts) are read back once the pass is queued;
  // the copy waits on the stream, so passes stay ordered.
  uint32_t hc[kCntWords];
  HIPCHK(hipMemcpyAsync(hc, q.counters, sizeof hc, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (pp.nrc) {
    n_query = hc[kCntQuery];
    if (n_slots) r->train_nt = hc[kCntTrain];
    if (timing) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, r->nev[0], r->nev[1]));
      ms_nrc += ms;
    }
  }
  if (st) {
    (void)n_query;
    st->ms_nrc += ms_nrc;
    add_pass_counts(st, hc, pp);
    if (timing) {
      for (uint32_t b = 0; b <= pp.max_depth; ++b) {
        const hipEvent_t* ev = &r->kev[4 * b];
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        st->ms_intersect += ms;
        HIPCHK(hipEventElapsedTime(&ms, ev[1], ev[2]));
        st->ms_shade += ms;
        HIPCHK(hipEventElapsedTime(&ms, ev[2], ev[3]));
        st->ms_shadow += ms;
      }
    }
  }
  return AQUA_OK;
}

// The frame's Adam steps on the batch its first pass collected (nt samples),
// on the renderer's stream. The first step's forward and backward passes may
// run while queries are still being inferred (they only read the weights);
// its update waits for before_update, the frame's last film, which follows
// every inference (the update rewrites the weights). While the cache is young (its
// first 128 steps) four steps per frame, each on a quarter of the samples,
// so it converges within the first frames; after that one per frame on all
// of them (measured: the same image error as four, at a third of the cost).
// No step waits: the last step's loss is copied to pinned host memory on the
// stream, and the frame reads it once its end event has passed. A
// non-finite loss switches the cache off: later frames render as the plain
// tracer.
// Adam steps the next frame takes: four while the cache is young (its first
// 128), then one, never past training_iters (0: no limit).
uint32_t frame_steps(const aqua_renderer* r) {
  uint32_t steps = r->nrc_steps < 128 ? 4u : 1u;
  if (r->cfg.training_iters > 0)
    steps = uint32_t(std::min<uint64_t>(steps, r->cfg.training_iters > r->nrc_steps
                                                   ? r->cfg.training_iters - r->nrc_steps : 0u));
  return steps;
}

// The frame's Adam steps (frame_steps) on nt rows of (rec, tgt), on stream
// s; the last step's loss is copied to loss_host. Returns the steps' count.
int train_steps(aqua_renderer* r, uint32_t nt, const float4* rec, const float4* tgt, hipStream_t s,
                hipEvent_t before_update, uint32_t* taken) {
  *taken = 0;
  const uint32_t steps = frame_steps(r);
  if (steps == 0 || nt == 0) return AQUA_OK;
  if (!r->loss_host)
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&r->loss_host), sizeof(float), hipHostMallocDefault));
  uint32_t per = (nt + steps - 1) / steps;
  uint32_t last = steps;
  for (uint32_t k = 0; k < steps && per > 0; ++k) {
    uint32_t lo = k * per, cnt = std::min(per, nt - std::min(nt, lo));
    if (cnt >= 32) last = k;
  }
  for (uint32_t k = 0; k <= last && last < steps; ++k) {
    uint32_t lo = k * per, cnt = std::min(per, nt - std::min(nt, lo));
    if (cnt < 32) continue;
    if (int e = aqua::nrc_train_device(r->nrc, cnt, rec + 4 * size_t(lo), tgt + lo, 1, nullptr, s, before_update))
      return e;
    before_update = nullptr;  // later steps follow the first update on s
    if (k == last) {
      if (int e = aqua::nrc_loss_async(r->nrc, r->loss_host, s)) return e;
      r->loss_rows = cnt;
    }
    ++r->nrc_steps;
    ++*taken;
  }
  return AQUA_OK;
}

// The frame's Adam steps on the batch its first pass collected (nt samples),
// on the renderer's stream. The first step's forward and backward passes may
// run while queries are still being inferred (they only read the weights);
// its update waits for before_update, the frame's last film, which follows
// every inference (the update rewrites the weights). While the cache is young (its
// first 128 steps) four steps per frame, each on a quarter of the samples,
// so it converges within the first frames; after that one per frame on all
// of them (measured: the same image error as four, at a third of the cost).
// No step waits: the last step's loss is copied to pinned host memory on the
// stream, and the frame reads it once its end event has passed. A
// non-finite loss switches the cache off: later frames render as the plain
// tracer.
int nrc_train_frame(aqua_renderer* r, uint32_t nt, aqua_frame_stats* st, hipEvent_t before_update) {
  hipStream_t s = r->stream;
  const bool timing = r->opts.timing != 0;
  r->batch_n = nt;
  if (timing) HIPCHK(hipEventRecord(r->nev[2], s));
  uint32_t taken = 0;
  if (int e = train_steps(r, nt, r->brec.p, r->btgt.p, s, before_update, &taken)) return e;
  if (taken == 0) return AQUA_OK;
  if (timing) {
    HIPCHK(hipEventRecord(r->nev[3], s));
    HIPCHK(hipEventSynchronize(r->nev[3]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, r->nev[2], r->nev[3]));
    st->ms_nrc += ms;
  }
  aqua::nrc_presort_drop(r->nrc);  // used by the step above, or not at all
  st->nrc_train_records += nt;
  return AQUA_OK;
}

// opts.train_ahead, at the start of a frame: the previous frame's batch is
// trained on tstream while the frame's first bounces run; the frame's
// queries wait for train_done.
int train_pending(aqua_renderer* r) {
  r->train_launched = false;
  const uint32_t nt = r->pending_nt;
  r->pending_nt = 0;
  if (nt == 0 || !r->nrc_enabled) return AQUA_OK;
  if (!r->tstream) {
    HIPCHK(hipStreamCreateWithFlags(&r->tstream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&r->train_go, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&r->train_done, hipEventDisableTiming));
  }
  HIPCHK(hipEventRecord(r->train_go, r->stream));  // after the frame's set-up
  HIPCHK(hipStreamWaitEvent(r->tstream, r->train_go, 0));
  uint32_t taken = 0;
  if (int e = train_steps(r, nt, r->pbrec.p, r->pbtgt.p, r->tstream, nullptr, &taken)) return e;
  if (taken == 0) return AQUA_OK;
  HIPCHK(hipEventRecord(r->train_done, r->tstream));
  r->train_launched = true;
  return AQUA_OK;
}

int prepare_frame(aqua_renderer* r, const aqua_camera* cam_in, uint32_t spp, aqua_camera* cam,
                  PassParams* pp, uint32_t* S) {
  *cam = cam_in ? *cam_in : r->camera;
  if (int e = check_camera(cam)) return e;
  if (spp == 0 || spp > (1u << 20)) return aqua::fail(AQUA_ERR_INVALID_ARG, "spp %u out of range", spp);
  uint64_t npix = uint64_t(cam->width) * cam->height;
  uint32_t s_pass = r->opts.samples_per_pass;
  if (s_pass == 0) {
    // About 128 M paths per pass (64 spp at 1080p): bounce launches are then
    // long enough that their ramp-up and tails stop mattering, and a frame
    // has few passes, each with its own tail of short bounces (1080p, 64
    // spp: 8 M paths per pass ran 16 % slower, 32 M 5 %). A pass holds ~230
    // bytes per path in queues, ~340 with cache queries: 30-44 GB at this
    // size, a sixth of the GPU's 288 GB.
    uint64_t target = 128ull << 20;
    s_pass = uint32_t(std::max<uint64_t>(1, target / npix));
  }
  s_pass = std::min(s_pass, spp);
  if (npix * s_pass >= 0x7FFFFFFFull)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "too many paths in one pass");
  if (int e = r->ensure_queues(size_t(npix) * s_pass)) return e;
  *S = s_pass;
  memset(pp, 0, sizeof *pp);
  pp->cam = cam_params(*cam);
  pp->npix = uint32_t(npix);
  pp->seed = r->cfg.seed;
  pp->max_depth = r->cfg.max_depth;
  pp->rr_depth = r->cfg.rr_depth;
  pp->dims = 8u * (r->cfg.max_depth + 2u);
  return AQUA_OK;
}

// Cache parameters of one pass of `n` paths: every stride-th path (from a
// per-pass offset) is a training path, the rest query at the query vertex.
int nrc_pass_params(aqua_renderer* r, uint32_t n, uint32_t frame, uint32_t sample0, PassParams* pp) {
  // The cache trains once per frame, on the frame's first pass; later
  // passes of the same frame only query.
  uint32_t want = std::max(1u, r->cfg.nrc_train_paths);
  uint32_t stride = std::max(1u, n / want);
  uint32_t offset = uint32_t((uint64_t(frame) * 2654435761ull + uint64_t(sample0) * 40503ull) % stride);
  // once the cache has taken training_iters Adam steps (0: no limit) it only
  // answers queries: no training paths, no more steps
  const bool trained = r->cfg.training_iters > 0 && r->nrc_steps >= r->cfg.training_iters;
  uint32_t n_train = sample0 == 0 && !trained ? std::min(want, (n - offset + stride - 1) / stride) : 0u;
  pp->nrc = 1;
  // visualize_cache shows the cache itself: paths end in a query at their
  // first hit, so the image is the emission there plus the cache's
  // prediction (training goes on as configured)
  pp->query_depth = r->cfg.visualize_cache ? 0u : (r->cfg.nrc_query_vertex >= 1 ? r->cfg.nrc_query_vertex - 1 : 0u);
  pp->train_stride = stride;
  pp->train_offset = offset;
  pp->n_train = n_train;
  for (int k = 0; k < 3; ++k) {
    pp->bmin[k] = r->bmin[k];
    pp->bscale[k] = r->bscale[k];
  }
  size_t slots = size_t(n_train) * (r->cfg.max_depth + 1u);
  if (int e = r->trec.ensure(4 * slots)) return e;
  if (int e = r->taux.ensure(2 * slots)) return e;
  if (int e = r->brec.ensure(4 * slots)) return e;
  if (int e = r->btgt.ensure(slots)) return e;
  return AQUA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int aqua_device_count(int* count) {
  if (!count) return aqua::fail(AQUA_ERR_INVALID_ARG, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = e == hipSuccess ? c : 0;
  return AQUA_OK;
}

int aqua_renderer_create(const aqua_scene* scene, const aqua_integrator_cfg* cfg, int device,
                         aqua_renderer** out) {
  if (!scene || !cfg || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_renderer_create: null argument");
  if (cfg->sampler > AQUA_SAMPLER_STRATIFIED)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "unknown sampler %u", cfg->sampler);
  *out = nullptr;
  if (cfg->max_depth >= kMaxBounces)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "max_depth %u exceeds %u", cfg->max_depth, kMaxBounces - 1);
  int cus = 0;
  if (int e = select_device(device, &cus)) return e;
  auto r = std::make_unique<aqua_renderer>();
  r->device = device;
  if (int e = noalloc_scratch(device, &r->na_scratch)) return e;
  // Persistent traversal grids: as many blocks as are resident at once for
  // both traversal kernels (their register / LDS footprint), so no block
  // waits for another to exit; a multiple of 8 (one work range per XCD).
  {
    int a = 0, b = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_intersect<false, QueueIO>, kBlock, 0));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow<false, ShadowIO>, kBlock, 0));
    int per_cu = std::max(1, std::min(a, b));
    r->persistent_blocks = std::max(8, (std::max(1, cus) * per_cu) & ~7);
  }
  r->cfg = *cfg;
  r->camera = scene->camera;
  r->opts.timing = 1;
  // lane 0: streams, counters, cursors and the traversal spill stacks of
  // every thread of the (capped) traversal grids; lane 1 on first use
  if (int e = r->lane[0].init(r->persistent_blocks, kMaxBounces)) return e;
  r->stream = r->lane[0].s;
  HIPCHK(hipEventCreate(&r->ev0));
  HIPCHK(hipEventCreate(&r->ev1));
  r->kev.resize(4 * (kMaxBounces + 1));
  for (hipEvent_t& e : r->kev) HIPCHK(hipEventCreate(&e));
  for (hipEvent_t& e : r->nev) HIPCHK(hipEventCreate(&e));

  // Geometry in BVH leaf order.
  const uint32_t n = uint32_t(scene->mat_ids.size());
  aqua::Bvh bvh;
  aqua::BvhBuildOpts bo;
  aqua::build_bvh(scene->verts.data(), n, bo, &bvh);
  aqua::Bvh4 bvh4;
  aqua::collapse_bvh4(bvh, &bvh4);
  // Inner child references are node indices below the leaf flag (2^31), and
  // the traversal addresses triangle slot i at 3 i float4s in 32 bits.
  if (n >= (1u << 29) || bvh4.nodes.size() >= (size_t(1) << 31) || bo.max_leaf > 7)
    return aqua::fail(AQUA_ERR_UNSUPPORTED, "scene too large: %u triangles (at most 2^29)", n);
  // At most three entries are pushed per visited node.
  if (3 * (bvh4.max_inner_depth + 1) > uint32_t(kLdsStack + kSpillStack))
    return aqua::fail(AQUA_ERR_UNSUPPORTED, "BVH depth %u exceeds the traversal stack",
                      bvh4.max_inner_depth);
  // Traversal tests a node's hit leaves as a bit mask over one run of at
  // most 4 x 7 slots starting at its first leaf child.
  if (bvh4.tri_order.size() != n)
    return aqua::fail(AQUA_ERR_UNSUPPORTED, "BVH4 covers %zu of %u triangles", bvh4.tri_order.size(), n);
  for (const aqua::Bvh4Node& nd : bvh4.nodes) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0, sum = 0;
    for (int k = 0; k < 4; ++k)
      if (nd.child[k] >= 0 && nd.count[k] > 0) {
        lo = std::min(lo, uint32_t(nd.child[k]));
        hi = std::max(hi, uint32_t(nd.child[k] + nd.count[k]));
        sum += uint32_t(nd.count[k]);
      }
    if (sum && (hi - lo != sum || hi > n))
      return aqua::fail(AQUA_ERR_UNSUPPORTED, "BVH4 leaf children of a node are not one run of slots");
  }
  std::vector<float4> tris(3 * size_t(n));
  for (uint32_t s = 0; s < n; ++s) {
    uint32_t id = bvh4.tri_order[s];
    const float* v = &scene->verts[9 * size_t(id)];
    uint32_t m = scene->mat_ids[id];
    float idf, mf;
    memcpy(&idf, &id, 4);
    memcpy(&mf, &m, 4);
    tris[3 * size_t(s) + 0] = make_float4(v[0], v[1], v[2], idf);
    tris[3 * size_t(s) + 1] = make_float4(v[3], v[4], v[5], mf);
    tris[3 * size_t(s) + 2] = make_float4(v[6], v[7], v[8], 0.0f);
  }
  r->n_tris = n;
  r->tri_order = bvh4.tri_order;
  std::vector<uint32_t> qnodes;
  if (!aqua::quantize_bvh4(bvh4, &qnodes))
    return aqua::fail(AQUA_ERR_UNSUPPORTED, "BVH4 leaf runs do not fit the node format");
  std::vector<float4> nodes(qnodes.size() / 4);
  memcpy(nodes.data(), qnodes.data(), qnodes.size() * sizeof(uint32_t));
  // Culling slack: 2^-16 of the largest coordinate magnitude.
  float mx = 0.0f;
  for (int k = 0; k < 3; ++k) {
    mx = std::max(mx, std::fabs(scene->info.bounds_lo[k]));
    mx = std::max(mx, std::fabs(scene->info.bounds_hi[k]));
  }
  r->delta = std::ldexp(mx > 0.0f ? mx : 1.0f, -16);

  aqua::LightTable lt;
  aqua::build_light_table(scene, &lt);
  std::vector<float4> mats(3 * scene->materials.size());
  for (size_t m = 0; m < scene->materials.size(); ++m) {
    const aqua_material& a = scene->materials[m];
    mats[3 * m + 0] = make_float4(a.base_color[0], a.base_color[1], a.base_color[2], a.metallic);
    mats[3 * m + 1] = make_float4(a.emission[0], a.emission[1], a.emission[2], a.roughness);
    float texf;
    memcpy(&texf, &a.base_color_tex, 4);
    mats[3 * m + 2] = make_float4(a.specular, lt.mat_pdf[m], texf, 0.0f);
  }
  r->n_area = uint32_t(lt.tri.size());
  r->n_lights = uint32_t(lt.pmf.size());
  if (!(lt.total_power > 0.0)) r->n_lights = 0;
  std::vector<float4> lights(4 * std::max<size_t>(1, lt.pmf.size()), make_float4(0, 0, 0, 0));
  for (uint32_t i = 0; i < r->n_area; ++i) {
    const float* v = &scene->verts[9 * size_t(lt.tri[i])];
    const aqua_material& a = scene->materials[scene->mat_ids[lt.tri[i]]];
    lights[4 * i + 0] = make_float4(v[0], v[1], v[2], lt.area[i]);
    lights[4 * i + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    lights[4 * i + 2] = make_float4(v[6], v[7], v[8], 0.0f);
    lights[4 * i + 3] = make_float4(a.emission[0], a.emission[1], a.emission[2], 0.0f);
  }
  for (size_t j = 0; j < scene->point_lights.size(); ++j) {
    const aqua_point_light& p = scene->point_lights[j];
    size_t i = r->n_area + j;
    lights[4 * i + 0] = make_float4(p.pos[0], p.pos[1], p.pos[2], 0.0f);
    lights[4 * i + 3] = make_float4(p.intensity[0], p.intensity[1], p.intensity[2], 0.0f);
  }
  std::vector<float> cdf = lt.cdf, pmf = lt.pmf;
  if (cdf.empty()) {
    cdf.push_back(1.0f);
    pmf.push_back(0.0f);
  }
  if (n == 0) {
    // An empty scene still needs a valid root: one node, all children empty.
    nodes.assign(4, make_float4(0, 0, 0, 0));
    uint32_t e4[4] = {aqua::kNoChild, aqua::kNoChild, 0u, 0u};
    memcpy(&nodes[2].z, e4, 8);
    memcpy(&nodes[3], e4, 8);
    tris.assign(3, make_float4(0, 0, 0, 0));
  }
  hipStream_t s = r->stream;
  if (int e = r->tris.upload(tris.data(), tris.size(), s)) return e;
  if (int e = r->nodes.upload(nodes.data(), nodes.size(), s)) return e;
  if (int e = r->mats.upload(mats.data(), mats.size(), s)) return e;
  if (int e = r->lights.upload(lights.data(), lights.size(), s)) return e;
  if (int e = r->cdf.upload(cdf.data(), cdf.size(), s)) return e;
  if (int e = r->pmf.upload(pmf.data(), pmf.size(), s)) return e;
  // Image textures of base colours.
  bool textured = false;
  for (const aqua_material& a : scene->materials) {
    if (!a.base_color_tex) continue;
    if (a.base_color_tex - 1u >= scene->images.size())
      return aqua::fail(AQUA_ERR_INVALID_ARG, "material names image %u of %zu", a.base_color_tex - 1u,
                        scene->images.size());
    textured = true;
  }
  if (textured) {
    if (scene->uvs.size() != 6 * size_t(n))
      return aqua::fail(AQUA_ERR_INVALID_ARG, "textured materials need texcoords (6 floats per triangle)");
    std::vector<uint4> info;
    size_t total = 0;
    for (const AquaImage& im : scene->images) {
      if (total + size_t(im.w) * im.h >= (size_t(1) << 32))
        return aqua::fail(AQUA_ERR_UNSUPPORTED, "image textures exceed 2^32 texels");
      info.push_back(make_uint4(uint32_t(total), im.w, im.h, 0u));
      total += size_t(im.w) * im.h;
    }
    std::vector<uint32_t> px(std::max<size_t>(total, 1));
    for (size_t k = 0; k < scene->images.size(); ++k) {
      const AquaImage& im = scene->images[k];
      for (size_t i = 0; i < size_t(im.w) * im.h; ++i) {
        const uint8_t* c = &im.rgba[4 * i];
        px[info[k].x + i] = uint32_t(c[0]) | uint32_t(c[1]) << 8 | uint32_t(c[2]) << 16 | uint32_t(c[3]) << 24;
      }
    }
    std::vector<float4> uv(2 * size_t(std::max(n, 1u)), make_float4(0, 0, 0, 0));
    for (uint32_t sl = 0; sl < n; ++sl) {
      const float* u = &scene->uvs[6 * size_t(r->tri_order[sl])];
      uv[2 * size_t(sl)] = make_float4(u[0], u[1], u[2], u[3]);
      uv[2 * size_t(sl) + 1] = make_float4(u[4], u[5], 0.0f, 0.0f);
    }
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = float(aqua::srgb_to_linear(i / 255.0));
    if (int e = r->texels.upload(px.data(), px.size(), s)) return e;
    if (int e = r->tex_info.upload(info.data(), info.size(), s)) return e;
    if (int e = r->tuv.upload(uv.data(), uv.size(), s)) return e;
    if (int e = r->tex_lut.upload(lut, 256, s)) return e;
  }
  if (int e = r->stats.ensure(kStWords)) return e;
  if (cfg->type == AQUA_INTEGRATOR_NRC) {
    if (cfg->nrc_precision > AQUA_NRC_FP8)
      return aqua::fail(AQUA_ERR_INVALID_ARG, "unknown nrc_precision %u", cfg->nrc_precision);
    if (cfg->nrc_query_vertex < 1) return aqua::fail(AQUA_ERR_INVALID_ARG, "nrc_query_vertex must be >= 1");
    aqua_nrc_cfg nc;
    aqua_nrc_default_cfg(&nc);
    nc.learning_rate = cfg->learning_rate;
    nc.train_precision = cfg->nrc_precision == AQUA_NRC_F16 ? AQUA_NRC_F16 : AQUA_NRC_BF16;
    nc.seed = cfg->seed;
    uint32_t tp = std::max(1u, cfg->nrc_train_paths);
    nc.max_batch = uint32_t(std::min<uint64_t>(uint64_t(tp) * std::max(1u, cfg->max_depth) / 4 + 64, 1u << 26));
    if (int e = aqua_nrc_create(device, &nc, &r->nrc)) return e;
    // positions are normalised to the scene bounds (slightly padded)
    for (int k = 0; k < 3; ++k) {
      float lo = scene->info.bounds_lo[k], hi = scene->info.bounds_hi[k];
      float pad = 1e-3f * std::max(1e-6f, hi - lo);
      r->bmin[k] = lo - pad;
      r->bscale[k] = 1.0f / std::max(1e-6f, (hi - lo) + 2 * pad);
    }
  }
  HIPCHK(hipStreamSynchronize(s));
  *out = r.release();
  return AQUA_OK;
}

int aqua_renderer_set_opts(aqua_renderer* r, const aqua_renderer_opts* o) {
  if (!r || !o) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (!o->train_ahead) r->pending_nt = 0;  // the batch waiting for the next frame is dropped
  r->opts = *o;
  return AQUA_OK;
}

void aqua_renderer_destroy(aqua_renderer* r) { delete r; }

static int render_samples(aqua_renderer* r, const aqua_camera* cam_in, uint32_t frame, uint32_t spp0,
                          uint32_t spp, bool fp64, aqua_frame_stats* stats) {
  HIPCHK(hipSetDevice(r->device));
  aqua_camera cam;
  PassParams pp;
  uint32_t S;
  if (int e = prepare_frame(r, cam_in, spp, &cam, &pp, &S)) return e;
  uint64_t npix = pp.npix;
  bool fresh = fp64 ? spp0 == 0 : true;
  if (fp64) {
    if (int e = r->film64.ensure(3 * npix)) return e;
    if (!fresh && (cam.width != r->film_w || cam.height != r->film_h))
      return aqua::fail(AQUA_ERR_INVALID_ARG, "oracle resume with a different resolution");
    if (fresh) {
      k_zero_f64<<<int((3 * npix + kBlock - 1) / kBlock), kBlock, 0, r->stream>>>(r->film64.p, 3 * npix);
      r->oracle_spp = 0;
    }
  } else {
    if (int e = r->film.ensure(npix)) return e;
    k_zero_f4<<<int((npix + kBlock - 1) / kBlock), kBlock, 0, r->stream>>>(r->film.p, npix);
  }
  aqua_frame_stats st;
  memset(&st, 0, sizeof st);
  bool count = r->opts.count_traversal != 0;
  k_zero_stats<<<1, 64, 0, r->stream>>>(r->stats.p);
  r->loss_rows = 0;
  HIPCHK(hipEventRecord(r->ev0, r->stream));
  // Two lanes: consecutive passes run concurrently on their own queues and
  // streams (cache queries included), their films chained in pass order;
  // the pass counters are read once the frame is done. Only for passes of
  // fewer than kLanePaths paths, whose launches leave the GPU room for a
  // second pass (1080p, 4 spp per pass: +12 %); passes of the default size
  // fill it alone, and a second one only competes for it (16 spp per pass:
  // -4 %). Not with per-kernel timing or traversal counters, which want the
  // passes one at a time.
  constexpr uint64_t kLanePaths = 16ull << 20;
  const uint32_t n_pass = (spp + S - 1) / S;
  const bool two = n_pass > 1 && uint64_t(npix) * S < kLanePaths && !r->opts.timing && !count;
  const bool cache = !fp64 && r->nrc && r->nrc_enabled;
  const bool ahead = cache && r->opts.train_ahead != 0;
  if (ahead) {
    if (int e = train_pending(r)) return e;
  } else if (!fp64) {
    r->pending_nt = 0;  // a batch left from train_ahead frames is not trained
    r->train_launched = false;
  }
  if (two) {
    if (int e = r->lane[1].init(r->persistent_blocks, kMaxBounces)) return e;
    if (int e = r->ensure_queues(size_t(npix) * S, 1)) return e;
    if (!r->pass0_done) HIPCHK(hipEventCreateWithFlags(&r->pass0_done, hipEventDisableTiming));
    if (n_pass * size_t(kCntWords) > r->pass_counters_cap) {
      if (r->pass_counters) HIPCHK(hipHostFree(r->pass_counters));
      r->pass_counters = nullptr;
      r->pass_counters_cap = 0;
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&r->pass_counters), n_pass * size_t(kCntWords) * 4,
                           hipHostMallocDefault));
      r->pass_counters_cap = n_pass * size_t(kCntWords);
    }
    // lane 1 starts after the frame's set-up on lane 0 (film and stats cleared)
    HIPCHK(hipEventRecord(r->lane[0].done, r->stream));
    HIPCHK(hipStreamWaitEvent(r->lane[1].s, r->lane[0].done, 0));
  }
  std::vector<PassParams> passes;
  bool train = false;  // a pass collected training samples
  if (r->nrc) aqua::nrc_presort_drop(r->nrc);
  r->presort = cache && !ahead && n_pass == 1 && !r->opts.timing && frame_steps(r) == 1;
  for (uint32_t k = 0, s0 = 0; s0 < spp; s0 += S, ++k) {
    pp.sample0 = spp0 + s0;
    pp.spp_in_pass = std::min(S, spp - s0);
    pp.frame = frame;
    pp.nrc = 0;
    if (cache) {
      if (int e = nrc_pass_params(r, pp.npix * pp.spp_in_pass, frame, pp.sample0, &pp)) return e;
      train = train || pp.n_train > 0;
    }
    if (two) {
      const int li = int(k & 1u);
      hipEvent_t wait = k > 0 ? r->lane[li ^ 1].film_done : nullptr;
      if (int e = run_pass(r, li, pp, fp64, &st, count, wait, r->pass_counters + size_t(k) * kCntWords)) return e;
      if (k == 0) HIPCHK(hipEventRecord(r->pass0_done, r->lane[0].s));
      passes.push_back(pp);
    } else if (int e = run_pass(r, 0, pp, fp64, &st, count)) {
      return e;
    }
    ++st.n_passes;
  }
  if (two) HIPCHK(hipStreamWaitEvent(r->stream, r->lane[1].done, 0));
  // the last pass's film (possibly still running on its lane's s2) follows
  // every earlier film and every inference of the frame
  const hipEvent_t last_film = r->lane[two ? (n_pass - 1u) & 1u : 0u].film_done;
  if (train) {
    // the first pass's training-sample count sizes the frame's Adam steps
    uint32_t nt = r->train_nt;
    if (two) {
      HIPCHK(hipEventSynchronize(r->pass0_done));
      nt = r->pass_counters[kCntTrain];
    }
    if (ahead) {
      // the batch waits for the next frame: it becomes the pending one, and
      // the buffer just trained on takes the next frame's records
      std::swap(r->brec.p, r->pbrec.p);
      std::swap(r->brec.n, r->pbrec.n);
      std::swap(r->btgt.p, r->pbtgt.p);
      std::swap(r->btgt.n, r->pbtgt.n);
      r->pending_nt = nt;
      r->batch_n = nt;
      st.nrc_train_records += nt;
    } else if (int e = nrc_train_frame(r, nt, &st, last_film)) {
      return e;
    }
  }
  r->presort = false;
  if (r->train_launched) HIPCHK(hipStreamWaitEvent(r->stream, r->train_done, 0));
  if (r->nrc) aqua::nrc_presort_drop(r->nrc);
  HIPCHK(hipStreamWaitEvent(r->stream, last_film, 0));
  HIPCHK(hipEventRecord(r->ev1, r->stream));
  HIPCHK(hipEventSynchronize(r->ev1));
  if (r->loss_rows) {
    // the frame's last Adam step: its loss arrived before ev1
    const float loss = *r->loss_host / float(3.0 * double(r->loss_rows));
    r->loss_rows = 0;
    st.nrc_loss = loss;
    if (!std::isfinite(loss)) {  // a diverged cache is not used again
      r->nrc_enabled = false;
      r->pending_nt = 0;
    }
  }
  r->train_launched = false;
  for (size_t k = 0; k < passes.size(); ++k) add_pass_counts(&st, r->pass_counters + k * kCntWords, passes[k]);
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
  st.ms_frame = ms;
  unsigned long long hs[kStWords];
  HIPCHK(hipMemcpy(hs, r->stats.p, sizeof hs, hipMemcpyDeviceToHost));
  st.nodes_visited = hs[kStNodes];
  st.tris_tested = hs[kStTris];
  st.shadow_nodes_visited = hs[kStShNodes];
  st.shadow_tris_tested = hs[kStShTris];
  st.nonfinite_samples = hs[kStNonfinite];
  st.spill_node_visits = hs[kStSpill];
  r->film_w = cam.width;
  r->film_h = cam.height;
  if (fp64)
    r->oracle_spp += spp;
  else
    r->film_spp = spp;
  if (stats) *stats = st;
  return AQUA_OK;
}

int aqua_render_frame(aqua_renderer* r, const aqua_camera* camera, uint32_t frame, uint32_t spp,
                      aqua_frame_stats* stats) {
  if (!r) return aqua::fail(AQUA_ERR_INVALID_ARG, "null renderer");
  if (frame >= (1u << 24)) return aqua::fail(AQUA_ERR_INVALID_ARG, "frame index too large");
  return render_samples(r, camera, frame, 0, spp, false, stats);
}

int aqua_render_oracle(aqua_renderer* r, const aqua_camera* camera, uint32_t frame, uint32_t spp_begin,
                       uint32_t spp_count) {
  if (!r) return aqua::fail(AQUA_ERR_INVALID_ARG, "null renderer");
  if (frame >= (1u << 24)) return aqua::fail(AQUA_ERR_INVALID_ARG, "frame index too large");
  if (spp_begin != 0 && spp_begin != r->oracle_spp)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "oracle resume at %u but the film holds %u samples", spp_begin,
                      r->oracle_spp);
  if (uint64_t(spp_begin) + spp_count > (1u << 20))
    return aqua::fail(AQUA_ERR_INVALID_ARG, "oracle sample range too large");
  return render_samples(r, camera, frame, spp_begin, spp_count, true, nullptr);
}

int aqua_film_read(aqua_renderer* r, float* rgb, uint32_t* spp) {
  if (!r || !rgb) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (r->film_spp == 0) return aqua::fail(AQUA_ERR_INVALID_ARG, "no frame rendered yet");
  HIPCHK(hipSetDevice(r->device));
  size_t npix = size_t(r->film_w) * r->film_h;
  std::vector<float4> h(npix);
  HIPCHK(hipMemcpy(h.data(), r->film.p, npix * sizeof(float4), hipMemcpyDeviceToHost));
  float inv = 1.0f / float(r->film_spp);
  for (size_t p = 0; p < npix; ++p) {
    rgb[3 * p] = h[p].x * inv;
    rgb[3 * p + 1] = h[p].y * inv;
    rgb[3 * p + 2] = h[p].z * inv;
  }
  if (spp) *spp = r->film_spp;
  return AQUA_OK;
}

int aqua_film_tonemap(aqua_renderer* r, float exposure, uint8_t* rgba) {
  if (!r || !rgba) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (r->film_spp == 0) return aqua::fail(AQUA_ERR_INVALID_ARG, "no frame rendered yet");
  if (!std::isfinite(exposure) || std::fabs(exposure) > 64.0f)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "exposure %g out of range", double(exposure));
  HIPCHK(hipSetDevice(r->device));
  ToneLut lut;
  for (int k = 0; k < 255; ++k) {
    // linear value whose sRGB encoding is (k + 0.5) / 255 (IEC 61966-2-1
    // inverse), then the smallest float at or above it
    double e = (k + 0.5) / 255.0;
    double x = aqua::srgb_to_linear(e);
    float xf = float(x);
    if (double(xf) < x) xf = std::nextafter(xf, INFINITY);
    lut.t[k] = xf;
  }
  lut.t[255] = INFINITY;  // never reached: codes stop at 255
  const uint32_t npix = r->film_w * r->film_h;
  DevBuf<uint32_t> out;
  if (int e = out.ensure(npix)) return e;
  k_tonemap<<<int((npix + kBlock - 1) / kBlock), kBlock, 0, r->stream>>>(
      r->film.p, 1.0f / float(r->film_spp), std::exp2(exposure), npix, lut, out.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rgba, out.p, size_t(npix) * 4, hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return AQUA_OK;
}

int aqua_oracle_read(aqua_renderer* r, double* rgb, uint32_t* spp) {
  if (!r || !rgb) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (r->oracle_spp == 0) return aqua::fail(AQUA_ERR_INVALID_ARG, "no oracle samples yet");
  HIPCHK(hipSetDevice(r->device));
  size_t n = 3 * size_t(r->film_w) * r->film_h;
  HIPCHK(hipMemcpy(rgb, r->film64.p, n * sizeof(double), hipMemcpyDeviceToHost));
  double inv = 1.0 / double(r->oracle_spp);
  for (size_t i = 0; i < n; ++i) rgb[i] *= inv;
  if (spp) *spp = r->oracle_spp;
  return AQUA_OK;
}

int aqua_film_error(aqua_renderer* r, double* relmse, double* relrmse) {
  if (!r || !relmse || !relrmse) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (r->film_spp == 0 || r->oracle_spp == 0)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "need both a frame and oracle samples");
  HIPCHK(hipSetDevice(r->device));
  uint32_t npix = r->film_w * r->film_h;
  int nb = std::min(1024, int((npix + kBlock - 1) / kBlock));
  if (int e = r->err_partial.ensure(3 * size_t(nb) + 3)) return e;
  k_err_partial<<<nb, kBlock, 0, r->stream>>>(r->film.p, 1.0f / float(r->film_spp), r->film64.p,
                                               1.0 / double(r->oracle_spp), npix, r->err_partial.p);
  k_err_final<<<1, 1, 0, r->stream>>>(r->err_partial.p, uint32_t(nb), r->err_partial.p + 3 * nb);
  HIPCHK(hipGetLastError());
  double h[3];
  HIPCHK(hipMemcpyAsync(h, r->err_partial.p + 3 * nb, sizeof h, hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  *relmse = h[0] / (3.0 * npix);
  *relrmse = h[2] > 0 ? std::sqrt(h[1] / h[2]) : 0.0;
  return AQUA_OK;
}

int aqua_primary_hits(aqua_renderer* r, const aqua_camera* cam_in, uint32_t frame, uint32_t sample,
                      uint32_t* prim_ids, float* t_hit) {
  if (!r || !prim_ids) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (frame >= (1u << 24) || sample >= (1u << 20))
    return aqua::fail(AQUA_ERR_INVALID_ARG, "frame / sample out of range");
  HIPCHK(hipSetDevice(r->device));
  aqua_camera cam;
  PassParams pp;
  uint32_t S;
  if (int e = prepare_frame(r, cam_in, 1, &cam, &pp, &S)) return e;
  pp.spp_in_pass = 1;
  pp.sample0 = sample;
  pp.frame = frame;
  uint32_t n = pp.npix;
  Queues q = r->queues();
  DevScene sc = r->dev_scene();
  DevBuf<uint32_t> ids;
  DevBuf<float> ts;
  if (int e = ids.ensure(n)) return e;
  if (int e = ts.ensure(n)) return e;
  hipStream_t s = r->stream;
  k_reset<<<8, kBlock, 0, s>>>(q.counters, n, q.cursors, kCursorSet);
  k_camera<<<grid_for(n), kBlock, 0, s>>>(pp, r->cfg.sampler, q);
  k_intersect<false><<<r->trav_grid(n), kBlock, 0, s>>>(sc.bvh, QueueIO{q.ro[0], q.rd[0], q.hit}, q.counters,
                                                         q.cursors, q.stats);
  k_hits_out<<<int((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(q.hit, sc.bvh.tris, n, ids.p, ts.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(prim_ids, ids.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (t_hit) HIPCHK(hipMemcpyAsync(t_hit, ts.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

static int upload_rays(aqua_renderer* r, uint32_t n, const float* rays) {
  std::vector<float4> o(n), d(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* x = rays + 8 * size_t(i);
    for (int k = 0; k < 8; ++k)
      if (k != 7 && !std::isfinite(x[k]) && !(k == 3 && x[k] == INFINITY))
        return aqua::fail(AQUA_ERR_NONFINITE, "ray %u has a non-finite component", i);
    if (x[4] == 0.0f && x[5] == 0.0f && x[6] == 0.0f)
      return aqua::fail(AQUA_ERR_INVALID_ARG, "ray %u has a zero direction", i);
    o[i] = make_float4(x[0], x[1], x[2], x[3]);
    d[i] = make_float4(x[4], x[5], x[6], 0.0f);
  }
  if (int e = r->ensure_queues(n)) return e;
  HIPCHK(hipMemcpyAsync(r->lane[0].ro[0].p, o.data(), n * sizeof(float4), hipMemcpyHostToDevice, r->stream));
  HIPCHK(hipMemcpyAsync(r->lane[0].rd[0].p, d.data(), n * sizeof(float4), hipMemcpyHostToDevice, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return AQUA_OK;
}

int aqua_trace_closest(aqua_renderer* r, uint32_t n, const float* rays, uint32_t* prim_ids, float* t_hit) {
  if (!r || (n && (!rays || !prim_ids))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (n == 0) return AQUA_OK;
  HIPCHK(hipSetDevice(r->device));
  if (int e = upload_rays(r, n, rays)) return e;
  Queues q = r->queues();
  DevScene sc = r->dev_scene();
  DevBuf<uint32_t> ids;
  DevBuf<float> ts;
  if (int e = ids.ensure(n)) return e;
  if (int e = ts.ensure(n)) return e;
  hipStream_t s = r->stream;
  k_reset<<<8, kBlock, 0, s>>>(q.counters, n, q.cursors, kCursorSet);
  k_intersect<false><<<r->trav_grid(n), kBlock, 0, s>>>(sc.bvh, QueueIO{q.ro[0], q.rd[0], q.hit}, q.counters,
                                                         q.cursors, q.stats);
  k_hits_out<<<int((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(q.hit, sc.bvh.tris, n, ids.p, ts.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(prim_ids, ids.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (t_hit) HIPCHK(hipMemcpyAsync(t_hit, ts.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

int aqua_trace_closest_noalloc(aqua_renderer* r, uint32_t n, const float* d_rays, uint32_t* d_prim_ids,
                               float* d_t_hit) {
  if (!r || (n && (!d_rays || !d_prim_ids || !d_t_hit))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (n > kNaMaxRays) return aqua::fail(AQUA_ERR_INVALID_ARG, "%u rays exceed the limit of %u per call", n, kNaMaxRays);
  if (reinterpret_cast<uintptr_t>(d_rays) % 16 || reinterpret_cast<uintptr_t>(d_prim_ids) % 4 ||
      reinterpret_cast<uintptr_t>(d_t_hit) % 4)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "misaligned device array (rays need 16 bytes, outputs 4)");
  if (n == 0) return AQUA_OK;
  if (!r->na_scratch) return aqua::fail(AQUA_ERR_HIP, "no scratch address for device %d", r->device);
  HIPCHK(hipSetDevice(r->device));
  hipStream_t s = r->stream;
  uint32_t* counters = r->na_scratch + kNaCounterWord;
  uint32_t* cursors = r->na_scratch + kNaCursorWord;
  k_reset<<<8, kBlock, 0, s>>>(counters, n, cursors, kCursorSet);
  DevBvh bvh = r->dev_scene().bvh;
  bvh.spill = r->na_scratch + kNaSpillWord;
  k_intersect<false><<<kNaGrid, kBlock, 0, s>>>(bvh, RayIO{reinterpret_cast<const float4*>(d_rays), d_prim_ids, d_t_hit},
                                                counters + cnt_rays(0), cursors, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

int aqua_trace_occluded(aqua_renderer* r, uint32_t n, const float* rays, uint8_t* occluded) {
  if (!r || (n && (!rays || !occluded))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (n == 0) return AQUA_OK;
  HIPCHK(hipSetDevice(r->device));
  if (int e = upload_rays(r, n, rays)) return e;
  Queues q = r->queues();
  DevScene sc = r->dev_scene();
  DevBuf<uint8_t> out;
  if (int e = out.ensure(n)) return e;
  hipStream_t s = r->stream;
  k_reset<<<8, kBlock, 0, s>>>(q.counters, n, q.cursors, kCursorSet);
  k_shadow<false><<<r->trav_grid(n), kBlock, 0, s>>>(sc.bvh, OcclusionIO{q.ro[0], q.rd[0], out.p},
                                                     q.counters + cnt_rays(0), q.cursors, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(occluded, out.p, n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

aqua_nrc* aqua_renderer_nrc(aqua_renderer* r) { return r ? r->nrc : nullptr; }

int aqua_renderer_set_nrc_enabled(aqua_renderer* r, int enabled) {
  if (!r) return aqua::fail(AQUA_ERR_INVALID_ARG, "null renderer");
  if (!r->nrc && enabled) return aqua::fail(AQUA_ERR_INVALID_ARG, "the renderer has no cache (type is not nrc)");
  r->nrc_enabled = enabled != 0;
  return AQUA_OK;
}

int aqua_renderer_set_nrc_precision(aqua_renderer* r, uint32_t precision) {
  if (!r) return aqua::fail(AQUA_ERR_INVALID_ARG, "null renderer");
  if (!r->nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "the renderer has no cache (type is not nrc)");
  if (precision > AQUA_NRC_FP8) return aqua::fail(AQUA_ERR_INVALID_ARG, "unknown nrc_precision %u", precision);
  r->cfg.nrc_precision = precision;
  return AQUA_OK;
}

int aqua_renderer_nrc_checkpoint(aqua_renderer* r, int restore) {
  if (!r) return aqua::fail(AQUA_ERR_INVALID_ARG, "null renderer");
  if (!r->nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "the renderer has no cache (type is not nrc)");
  HIPCHK(hipSetDevice(r->device));
  if (int e = aqua_nrc_checkpoint(r->nrc, restore)) return e;
  // with train_ahead the batch waiting for the next frame is part of the state
  DevBuf<float4>& rec = restore ? r->pbrec : r->ck_brec;
  DevBuf<float4>& tgt = restore ? r->pbtgt : r->ck_btgt;
  const DevBuf<float4>& rec_src = restore ? r->ck_brec : r->pbrec;
  const DevBuf<float4>& tgt_src = restore ? r->ck_btgt : r->pbtgt;
  const uint32_t np = restore ? r->ck_pending_nt : r->pending_nt;
  if (np) {
    if (int e = rec.ensure(4 * size_t(np))) return e;
    if (int e = tgt.ensure(np)) return e;
    HIPCHK(hipMemcpyAsync(rec.p, rec_src.p, 4 * size_t(np) * sizeof(float4), hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(tgt.p, tgt_src.p, size_t(np) * sizeof(float4), hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
  }
  if (restore) {
    r->nrc_steps = r->ck_nrc_steps;
    r->nrc_enabled = r->ck_nrc_enabled;
    r->pending_nt = r->ck_pending_nt;
  } else {
    r->ck_nrc_steps = r->nrc_steps;
    r->ck_nrc_enabled = r->nrc_enabled;
    r->ck_pending_nt = r->pending_nt;
  }
  return AQUA_OK;
}

int aqua_renderer_nrc_batch(aqua_renderer* r, uint32_t* n, const void** d_records, const void** d_targets) {
  if (!r || !n || !d_records || !d_targets) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (!r->nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "the renderer has no cache (type is not nrc)");
  *n = r->batch_n;
  const bool pending = r->opts.train_ahead != 0;  // the newest batch waits in the pending buffers
  *d_records = pending ? r->pbrec.p : r->brec.p;
  *d_targets = pending ? r->pbtgt.p : r->btgt.p;
  return AQUA_OK;
}

int aqua_fp_probe(uint32_t n, const float* a, const float* b, const float* c, const float* u, float* out) {
  if (n == 0) return AQUA_OK;
  if (!a || !b || !c || !u || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (int e = select_device(dev)) return e;
  DevBuf<float> da, db, dc, du, dout;
  hipStream_t s = nullptr;
  if (int e = da.upload(a, n, s)) return e;
  if (int e = db.upload(b, n, s)) return e;
  if (int e = dc.upload(c, n, s)) return e;
  if (int e = du.upload(u, n, s)) return e;
  if (int e = dout.ensure(5 * size_t(n))) return e;
  k_fp_probe<<<int((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(da.p, db.p, dc.p, du.p, n, dout.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, dout.p, 5 * size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
  return AQUA_OK;
}

}  // extern "C"
