// Neural radiance cache on gfx950: hash-grid + SH encoding and a 64-wide MLP
// on MFMA, fused forward / backward / hash-grid scatter, MFMA weight
// gradients, Adam.
//
// Methods: multiresolution hash encoding (Mueller et al. 2022, "Instant
// Neural Graphics Primitives"); relative L2 loss and online training
// (Mueller et al. 2021, "Real-time Neural Radiance Caching for Path
// Tracing"); Adam (Kingma & Ba 2015). The reference only names the
// integrator (scenes/integrator.json:2 `"type": "nrc"`, :7 learning_rate).
//
// MFMA mapping (v_mfma_f32_16x16x32_{f16,bf16,fp8}, wave64). A wave owns 16
// batch rows. Every layer is computed transposed, out^T = W^T x in^T, so the
// batch row stays in lane%16 and a lane group g = lane/16 holds, after each
// layer, the 16 outputs {16*mt + 4*g + j}. The K (input) dimension of every
// layer is ordered by
//     kappa(g, ks, i) = 16 * (2*ks + i/4) + 4*g + i%4
// which is exactly the set a lane group holds, so the accumulator registers
// of one layer are the B operand of the next with no data movement; the
// permutation lives in the packed weights. The backward pass multiplies by
// W (same packing of W^T) and lands in the same layout, so ReLU masks and
// the hash-feature gradients are lane-local as well.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "aqua/aqua.h"
#include "device/common.h"
#include "host/internal.h"

namespace {

constexpr int kLevels = 16;
// Levels kFineLevel0.. (the finest, hashed with the default 2^19 table)
// accumulate their table gradient as BF16 pairs: one packed atomic per entry
// instead of two FP32 ones. BF16 rather than FP16 because the relative L2
// gradient 2 (y - t) / (y^2 + 0.01) reaches 200 t where the prediction is
// near zero, and radiance targets reach 10^3 and more: FP16 sums overflow
// there. An 8-bit mantissa stops adding once a sum is ~256 addends large, so
// the coarser levels 0-7, whose entries collect many rows each (thousands
// when a batch clusters in a small region: levels 4-7 lost 30-80 % of their
// gradient in BF16 there), accumulate FP32 pairs; k_nrc_scatter merges
// equal entries across a wave of spatially sorted rows first, so their
// atomics are few.
constexpr int kFineLevel0 = 8;
static_assert(kFineLevel0 == 8, "k_nrc_scatter merges levels 0..7 and adds levels 8..15 as BF16 pairs");
// Training stages the forward and backward weight operands in LDS per block
// (72 KB in 16-bit) instead of reading them through L1 per layer and tile.
constexpr bool kLdsTrainWeights = true;
// Training blocks of 4 waves: two fit a CU beside their staged operands
// (8 and 12 waves per block, one block per CU, measured slower).
constexpr int kTrainBlock = 256;
constexpr int kHidden = 4;
constexpr int kMlp = 4 * 64 * 64 + 64 * 16;    // 17408 floats
static_assert(kMlp % 64 == 0, "partial sums run 64 weights per block");
constexpr int kFwdPack = kHidden * 8 * 64 + 2 * 64;  // A operands, forward
constexpr int kBwdPack = kHidden * 8 * 64 + 4 * 64;  // A operands, backward
constexpr int kBlock = 256;
// Rows per weight-gradient partial: about 256 partials per step, so the
// weight-gradient launch fills the chip, each 256..4096 rows (a multiple of
// the kernel's 32-row step).
inline uint32_t chunk_rows(uint32_t npad) {
  uint32_t c = ((npad / 256u) + 31u) & ~31u;
  return c < 256u ? 256u : (c > 4096u ? 4096u : c);
}
inline uint32_t n_chunks(uint32_t npad) { return (npad + chunk_rows(npad) - 1) / chunk_rows(npad); }
// Upper bound of n_chunks over every row count up to npad (partials buffer).
inline uint32_t max_chunks(uint32_t npad) { return std::max(257u, (npad + 4095u) / 4096u); }

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 g2v __attribute__((ext_vector_type(2)));  // BF16 gradient pair

struct Grid {
  float res[kLevels];
  uint32_t dense_mask;
  uint32_t tmask;
  uint32_t log2T;
};

__device__ __forceinline__ int kappa(int g, int ks, int i) { return 16 * (2 * ks + (i >> 2)) + 4 * g + (i & 3); }

// Table index of vertex (cx, cy, cz) of level lv: dense levels index the
// vertex lattice directly; hashed levels hash 4 x 4 x 2 blocks of vertices
// (the primes of the spatial hash of Mueller et al. 2022) and keep the
// vertex's place in its block in the low 5 bits, so a block is one aligned
// 128-byte run of the table: a cell's 8 corners touch 2.3 lines on average
// instead of 4, and nearby queries share lines. Within the block x takes bits
// 0-1, y's low bit bit 2, z bit 3 and y's high bit bit 4, so each 64-byte half
// is a 4 x 2 x 2 brick: a cell's corners lie in 2.8 such segments on average
// (3.1 with 4 x 4 x 1 halves), which is what the gradient atomics pay for.
__device__ __forceinline__ uint32_t grid_index(const Grid& G, int lv, int cx, int cy, int cz) {
  if ((G.dense_mask >> lv) & 1u) {
    uint32_t n = uint32_t(G.res[lv]) + 1u;
    return uint32_t(cx) + n * (uint32_t(cy) + n * uint32_t(cz));
  }
  const uint32_t b = (uint32_t(cx >> 2) * 1u) ^ (uint32_t(cy >> 2) * 2654435761u) ^ (uint32_t(cz >> 1) * 805459861u);
  return ((b << 5) | (uint32_t(cx) & 3u) | ((uint32_t(cy) & 1u) << 2) | ((uint32_t(cz) & 1u) << 3) |
          ((uint32_t(cy) & 2u) << 3)) &
         G.tmask;
}

struct Cell {
  int c[3];
  float f[3];
};

__device__ __forceinline__ Cell cell_of(const Grid& G, int lv, float px, float py, float pz) {
  float N = G.res[lv];
  float s[3] = {px * N, py * N, pz * N};
  Cell c;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float fl = floorf(s[d]);
    if (fl > N - 1.0f) fl = N - 1.0f;
    c.c[d] = int(fl);
    c.f[d] = s[d] - fl;
  }
  return c;
}

__device__ __forceinline__ float corner_w(const Cell& c, int k) {
  float wx = (k & 1) ? c.f[0] : 1.0f - c.f[0];
  float wy = (k & 2) ? c.f[1] : 1.0f - c.f[1];
  float wz = (k & 4) ? c.f[2] : 1.0f - c.f[2];
  return (wx * wy) * wz;
}

typedef uint32_t u2a __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ void unpack_entry(uint32_t u, float* a, float* b) {
  *a = float(__builtin_bit_cast(_Float16, uint16_t(u & 0xFFFFu)));
  *b = float(__builtin_bit_cast(_Float16, uint16_t(u >> 16)));
}


// SH basis l = 0..3 (16 coefficients) of a unit direction.
__device__ __forceinline__ void sh16(float x, float y, float z, float* o) {
  float xx = x * x, yy = y * y, zz = z * z;
  o[0] = 0.28209479177387814f;
  o[1] = -0.48860251190291987f * y;
  o[2] = 0.48860251190291987f * z;
  o[3] = -0.48860251190291987f * x;
  o[4] = 1.0925484305920792f * x * y;
  o[5] = -1.0925484305920792f * y * z;
  o[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
  o[7] = -1.0925484305920792f * x * z;
  o[8] = 0.54627421529603959f * (xx - yy);
  o[9] = 0.59004358992664352f * y * (-3.0f * xx + yy);
  o[10] = 2.8906114426405538f * x * y * z;
  o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * zz);
  o[12] = 0.3731763325901154f * z * (5.0f * zz - 3.0f);
  o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * zz);
  o[14] = 1.4453057213202769f * z * (xx - yy);
  o[15] = 0.59004358992664352f * x * (-xx + 3.0f * yy);
}

// ---- MFMA operand helpers -------------------------------------------------

template <int P>
struct Op;

template <>
struct Op<AQUA_NRC_F16> {
  typedef h8 T;
  typedef uint4 Packed;
  static __device__ __forceinline__ T make(const float* v, float) {
    T r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = (_Float16)v[i];
    return r;
  }
  static __device__ __forceinline__ T unpack(uint4 u) { return __builtin_bit_cast(T, u); }
  static __device__ __forceinline__ f4v mma(T a, T b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ uint16_t bits(float v) {
    return __builtin_bit_cast(uint16_t, (_Float16)v);
  }
};

template <>
struct Op<AQUA_NRC_BF16> {
  typedef b8 T;
  typedef uint4 Packed;
  static __device__ __forceinline__ T make(const float* v, float) {
    T r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = (__bf16)v[i];
    return r;
  }
  static __device__ __forceinline__ T unpack(uint4 u) { return __builtin_bit_cast(T, u); }
  static __device__ __forceinline__ f4v mma(T a, T b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ uint16_t bits(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
};

// FP16 training: backward values are clamped to the FP16 range before they
// become 16-bit operands or stored gradients, so an extreme sample (a
// prediction near zero against a bright target makes the relative L2
// gradient about 200 x the target) saturates instead of turning into inf and
// then NaN. BF16 has the FP32 range and needs no clamp.
template <int P>
__device__ __forceinline__ float clamp16(float v) {
  if constexpr (P == AQUA_NRC_F16)
    return __builtin_amdgcn_fmed3f(v, -65504.0f, 65504.0f);
  else
    return v;
}

__device__ __forceinline__ float sat448(float v) { return v > 448.0f ? 448.0f : (v < -448.0f ? -448.0f : v); }

template <>
struct Op<AQUA_NRC_FP8> {
  typedef long T;
  typedef uint2 Packed;
  // OCP e4m3 (gfx950), values pre-scaled and saturated to +-448.
  static __device__ __forceinline__ T make(const float* v, float scale) {
    int lo = 0, hi = 0;
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(sat448(v[0] * scale), sat448(v[1] * scale), lo, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(sat448(v[2] * scale), sat448(v[3] * scale), lo, true);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(sat448(v[4] * scale), sat448(v[5] * scale), hi, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(sat448(v[6] * scale), sat448(v[7] * scale), hi, true);
    return long((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
  }
  static __device__ __forceinline__ T unpack(uint2 u) { return long((uint64_t(u.y) << 32) | u.x); }
  static __device__ __forceinline__ f4v mma(T a, T b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
  }
};

// ---- the fused kernel -------------------------------------------------------

struct FwdArgs {
  Grid G;
  const float4* rec;        // 4 per row
  const float4* tgt;        // 1 per row (training)
  uint32_t n, npad;
  const uint32_t* n_ptr;    // inference: row count read on the device (else n)
  const uint32_t* perm;     // row j is record perm[j] (else j): spatial order
  const uint32_t* table16;  // [level][T] two FP16 features
  const void* pack_fwd;     // A operands
  const void* pack_bwd;
  const float* wscale;      // FP8: per-layer weight scales (device)
  const float* ascale;      // FP8: per-layer activation scales (device)
  float* out;               // 3 per row (inference)
  // training
  float* table_grad;        // [level][T][2] FP32 (levels < kFineLevel0)
  uint32_t* table_grad16;   // [level - kFineLevel0][T] two BF16
  float2* fgrad;            // [level][npad] d loss / d (feature pair) of each row
  uint16_t* HT;             // 5 x [npad][64] layer inputs
  uint16_t* DT;             // 5 x [npad][64] pre-activation grads (last: 16 used)
  float* loss_acc;
  uint32_t* act_amax;       // 5 per-layer input amax (float bits)
};

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Encodes this lane group's 16 input features of one row: f[ks*8 + i] is
// feature kappa(g, ks, i) (hash levels 2g, 2g+1, 8+2g, 9+2g; SH 4g..4g+3;
// extras 4g..4g+3). Each lane sums all eight corners of its four levels.
__device__ __forceinline__ void encode16(const FwdArgs& a, uint32_t T_, int g, bool valid, float px, float py,
                                         float pz, float4 r0, float4 r1, float4 r2, float4 r3, float f[16]) {
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    int lv = q < 2 ? 2 * g + q : 8 + 2 * g + (q - 2);
    Cell c = cell_of(a.G, lv, px, py, pz);
    float s0 = 0.0f, s1 = 0.0f;
    const uint32_t* tl = a.table16 + size_t(lv) * T_;
    // Corners x and x + 1 are adjacent words (dense levels always; hashed
    // levels unless x + 1 starts the next block): one 8-byte load fetches
    // both, so a gather instruction touches half as many lines. Otherwise
    // the second word is fetched on its own (the table has 2 spare words,
    // so the wide load stays in bounds).
    const bool pair = ((a.G.dense_mask >> lv) & 1u) || (c.c[0] & 3) != 3;
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
      const int cy = c.c[1] + ((k >> 1) & 1), cz = c.c[2] + (k >> 2);
      u2a u = *reinterpret_cast<const u2a*>(tl + grid_index(a.G, lv, c.c[0], cy, cz));
      if (!pair) u.y = tl[grid_index(a.G, lv, c.c[0] + 1, cy, cz)];
      float e0, e1;
      unpack_entry(u.x, &e0, &e1);
      float w = corner_w(c, k);
      s0 = __builtin_fmaf(w, e0, s0);
      s1 = __builtin_fmaf(w, e1, s1);
      unpack_entry(u.y, &e0, &e1);
      w = corner_w(c, k + 1);
      s0 = __builtin_fmaf(w, e0, s0);
      s1 = __builtin_fmaf(w, e1, s1);
    }
    f[2 * q] = valid ? s0 : 0.0f;
    f[2 * q + 1] = valid ? s1 : 0.0f;
  }
  float sh[16];
  sh16(r1.x, r1.y, r1.z, sh);
  float ex[4];
  if (g == 0) {
    ex[0] = r2.x; ex[1] = r2.y; ex[2] = r2.z; ex[3] = r0.w;
  } else if (g == 1) {
    ex[0] = r3.x; ex[1] = r3.y; ex[2] = r3.z; ex[3] = 1.0f;
  } else {
    ex[0] = ex[1] = ex[2] = ex[3] = 0.0f;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float s = g == 0 ? sh[i] : (g == 1 ? sh[4 + i] : (g == 2 ? sh[8 + i] : sh[12 + i]));
    f[8 + i] = valid ? s : 0.0f;
    f[12 + i] = valid ? ex[i] : 0.0f;
  }
}

// Row-major [row][64] activation / gradient buffers for the weight
// gradients: a lane stores its features 16 mt + 4 g + {0..3} as one 8-byte
// write per mt (p points at feature 4 g of the row).
template <class O>
__device__ __forceinline__ void store4(uint16_t* p, const float v[4]) {
  uint2 u;
  u.x = uint32_t(O::bits(v[0])) | (uint32_t(O::bits(v[1])) << 16);
  u.y = uint32_t(O::bits(v[2])) | (uint32_t(O::bits(v[3])) << 16);
  *reinterpret_cast<uint2*>(p) = u;
}
template <class O>
__device__ __forceinline__ void store_row4x4(uint16_t* p, const float h[4][4]) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) store4<O>(p + 16 * mt, h[mt]);
}

// v0 / v1: the two K halves of a layer input held as h[mt][j] (feature
// 16 mt + 4 g + j) -> MFMA B operands.
__device__ __forceinline__ void split_b(const float h[4][4], float v0[8], float v1[8]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v0[i] = h[0][i];
    v0[4 + i] = h[1][i];
    v1[i] = h[2][i];
    v1[4 + i] = h[3][i];
  }
}

template <int P, bool TRAIN>
__global__ void __launch_bounds__(TRAIN ? kTrainBlock : kBlock) k_nrc_fwd(FwdArgs a) {
  typedef Op<P> O;
  typedef typename O::T T;
  typedef typename O::Packed Pk;
  const Pk* pf = reinterpret_cast<const Pk*>(a.pack_fwd);
  // Inference: the block stages the forward weight operands in LDS once
  // (34.8 KB in 16-bit, 17.4 KB in FP8), so each tile reads them at LDS
  // bandwidth instead of re-fetching them through L1. `woff` is an opaque
  // zero that keeps those reads inside the tile loop (hoisted, the 40
  // operands would need ~160 VGPRs).
  extern __shared__ __align__(16) unsigned char nrc_lds[];
  const Pk* wl = reinterpret_cast<const Pk*>(nrc_lds);
  int woff = 0;
  constexpr bool kLdsW = !TRAIN || kLdsTrainWeights;
  if constexpr (kLdsW) {
    Pk* dst = reinterpret_cast<Pk*>(nrc_lds);
    for (int i = int(threadIdx.x); i < kFwdPack; i += int(blockDim.x)) dst[i] = pf[i];
    if constexpr (TRAIN) {
      const Pk* pbg = reinterpret_cast<const Pk*>(a.pack_bwd);
      for (int i = int(threadIdx.x); i < kBwdPack; i += int(blockDim.x)) dst[kFwdPack + i] = pbg[i];
    }
    __syncthreads();
  }
  auto wfwd = [&](int idx) -> Pk {
    if constexpr (kLdsW)
      return wl[idx + woff];
    else
      return pf[idx];
  };
  const int lane = int(__lane_id());
  const int g = lane >> 4, r16 = lane & 15;
  const uint32_t T_ = 1u << a.G.log2T;
  const uint32_t n = (!TRAIN && a.n_ptr) ? *a.n_ptr : a.n;
  const uint32_t tiles = TRAIN ? a.npad / 16u : (n + 15u) / 16u;
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so
  // XCD x (blocks x, x + 8, ...) takes the x-th contiguous eighth of the
  // tiles; with spatially sorted rows each XCD's L2 then serves one region
  // of the grid. The grid is a multiple of 8 blocks.
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t waves_x = (gridDim.x >> 3) * (blockDim.x >> 6);
  const uint32_t per_x = (tiles + 7u) >> 3;
  const uint32_t t_end = min(tiles, (xcd + 1u) * per_x);
  float loss_lane = 0.0f;
  float amax_lane[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  for (uint32_t tile = xcd * per_x + (blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < t_end;
       tile += waves_x) {
    // The weight operands are loop-invariant; without this barrier the
    // compiler hoists all 40 of them out of the tile loop into ~160 VGPRs.
    // Re-reading them per tile hits L1 and keeps occupancy up.
    if constexpr (kLdsW)
      asm volatile("" : "+s"(woff));
    else
      asm volatile("" : "+s"(pf));
    const uint32_t row = tile * 16u + uint32_t(r16);
    const bool valid = row < n;
    const uint32_t src = (a.perm && valid) ? a.perm[row] : row;
    float4 r0 = make_float4(0, 0, 0, 0), r1 = r0, r2 = r0, r3 = r0;
    if (valid) {
      r0 = a.rec[4 * size_t(src)];
      r1 = a.rec[4 * size_t(src) + 1];
      r2 = a.rec[4 * size_t(src) + 2];
      r3 = a.rec[4 * size_t(src) + 3];
    }
    float px = fminf(fmaxf(r0.x, 0.0f), 1.0f), py = fminf(fmaxf(r0.y, 0.0f), 1.0f),
          pz = fminf(fmaxf(r0.z, 0.0f), 1.0f);
    float cur[4][4];  // current layer input, feature 16 mt + 4 g + j
    {
      float f[16];
      encode16(a, T_, g, valid, px, py, pz, r0, r1, r2, r3, f);
#pragma unroll
      for (int i = 0; i < 16; ++i) cur[2 * (i >> 3) + ((i & 7) >> 2)][i & 3] = f[i];
    }
    uint32_t mask[kHidden + 1];  // training: bit (4 mt + j) = layer input > 0
    float y[4];
    // ---- forward
#pragma unroll
    for (int L = 0; L <= kHidden; ++L) {
      // one layer's weights live at a time
      if constexpr (kLdsW)
        asm volatile("" : "+s"(woff));
      else
        asm volatile("" : "+s"(pf));
      float v0[8], v1[8];
      split_b(cur, v0, v1);
      float sa = 1.0f, deq = 1.0f;
      if constexpr (P == AQUA_NRC_FP8) {
        sa = a.ascale[L];
        deq = 1.0f / (sa * a.wscale[L]);
      }
      T b0 = O::make(v0, sa), b1 = O::make(v1, sa);
      if constexpr (TRAIN) {
        float m = 0.0f;
        uint32_t bits = 0;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            m = fmaxf(m, fabsf(cur[mt][j]));
            bits |= (cur[mt][j] > 0.0f ? 1u : 0u) << (4 * mt + j);
          }
        mask[L] = bits;
        amax_lane[L] = fmaxf(amax_lane[L], m);
        store_row4x4<O>(a.HT + (size_t(L) * a.npad + row) * 64 + 4 * g, cur);
      }
      if (L < kHidden) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
          acc = O::mma(O::unpack(wfwd(((L * 4 + mt) * 2 + 0) * 64 + lane)), b0, acc);
          acc = O::mma(O::unpack(wfwd(((L * 4 + mt) * 2 + 1) * 64 + lane)), b1, acc);
#pragma unroll
          for (int j = 0; j < 4; ++j) cur[mt][j] = fmaxf(acc[j] * deq, 0.0f);
        }
      } else {
        f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = O::mma(O::unpack(wfwd(kHidden * 8 * 64 + 0 * 64 + lane)), b0, acc);
        acc = O::mma(O::unpack(wfwd(kHidden * 8 * 64 + 1 * 64 + lane)), b1, acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) y[j] = acc[j] * deq;
      }
    }
    if constexpr (!TRAIN) {
      if (g == 0 && valid) {
        a.out[3 * size_t(src) + 0] = y[0];
        a.out[3 * size_t(src) + 1] = y[1];
        a.out[3 * size_t(src) + 2] = y[2];
      }
    } else {
      // ---- loss gradient (unscaled: the 1 / (3 n) factor is applied in Adam)
      float dy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (g == 0 && valid) {
        float4 t = a.tgt[src];
        float tv[3] = {t.x, t.y, t.z};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          float den = y[j] * y[j] + 0.01f;
          float d = y[j] - tv[j];
          loss_lane += d * d / den;
          dy[j] = clamp16<P>(2.0f * d / den);
        }
      }
      store4<O>(a.DT + (size_t(kHidden) * a.npad + row) * 64 + 4 * g, dy);
      // ---- backward: dz in the accumulator layout, masks from the forward
      const Pk* pb = reinterpret_cast<const Pk*>(a.pack_bwd);
      auto wbwd = [&](int idx) -> Pk {
        if constexpr (kLdsW)
          return wl[kFwdPack + idx + woff];
        else
          return pb[idx];
      };
      // as for the forward operands: keep the loads in the loop
      if constexpr (kLdsW)
        asm volatile("" : "+s"(woff));
      else
        asm volatile("" : "+s"(pb));
      float dz[4][4];
      {
        float v[8] = {dy[0], dy[1], dy[2], dy[3], 0.0f, 0.0f, 0.0f, 0.0f};
        T bd = O::make(v, 1.0f);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          f4v acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
          acc2 = O::mma(O::unpack(wbwd(kHidden * 8 * 64 + mt * 64 + lane)), bd, acc2);
#pragma unroll
          for (int j = 0; j < 4; ++j) dz[mt][j] = ((mask[kHidden] >> (4 * mt + j)) & 1u) ? clamp16<P>(acc2[j]) : 0.0f;
        }
      }
#pragma unroll
      for (int LL = kHidden - 1; LL >= 0; --LL) {
        if constexpr (kLdsW)
          asm volatile("" : "+s"(woff));
        else
          asm volatile("" : "+s"(pb));
        store_row4x4<O>(a.DT + (size_t(LL) * a.npad + row) * 64 + 4 * g, dz);
        float v0b[8], v1b[8];
        split_b(dz, v0b, v1b);
        T bd0 = O::make(v0b, 1.0f), bd1 = O::make(v1b, 1.0f);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          f4v acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
          acc2 = O::mma(O::unpack(wbwd(((LL * 4 + mt) * 2 + 0) * 64 + lane)), bd0, acc2);
          acc2 = O::mma(O::unpack(wbwd(((LL * 4 + mt) * 2 + 1) * 64 + lane)), bd1, acc2);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            dz[mt][j] = (LL == 0 || ((mask[LL] >> (4 * mt + j)) & 1u)) ? clamp16<P>(acc2[j]) : 0.0f;
        }
      }
      // ---- d loss / d (feature pair) of this lane's four levels, for
      // k_nrc_scatter: level q's two features are dz[q < 2 ? 0 : 1][2 (q % 2) + {0,1}]
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int lv = q < 2 ? 2 * g + q : 8 + 2 * g + (q - 2);
        float d0 = q < 2 ? dz[0][2 * (q & 1)] : dz[1][2 * (q & 1)];
        float d1 = q < 2 ? dz[0][2 * (q & 1) + 1] : dz[1][2 * (q & 1) + 1];
        if (!valid) d0 = d1 = 0.0f;
        a.fgrad[size_t(lv) * a.npad + row] = make_float2(d0, d1);
      }
    }
  }
  if constexpr (TRAIN) {
    float ls = wave_sum(loss_lane);
    if (lane == 0 && ls != 0.0f) atomicAdd(a.loss_acc, ls);
#pragma unroll
    for (int L = 0; L < 5; ++L) {
      float m = wave_max(amax_lane[L]);
      if (lane == 0) atomicMax(&a.act_amax[L], __float_as_uint(m));
    }
  }
}

// ---- hash-grid gradient scatter -------------------------------------------------
//
// One wave per 64 consecutive rows of the batch (in the spatial order the
// training step sorts it into), after k_nrc_fwd<P, true> has written every
// row's d loss / d (feature pair) per level to fgrad.
//
// The atomics' cost is set by the 64-byte segments an instruction touches,
// not by how many lanes hit a segment, so each layout keeps a row's corners
// in one instruction:
// * levels 0..7 (FP32 pairs): 16 lanes per row (corner k, component c), four
//   rows per instruction; the two halves of an entry share a segment. Where
//   the four rows (spatial neighbours) hit the same entry, their values are
//   summed first and one lane adds.
// * levels 8..15 (BF16 pairs, one 32-bit atomic per entry): 8 lanes per row,
//   one corner each, eight rows per instruction (a cell's corners lie in
//   about 2.8 segments with the block layout of grid_index).
__global__ void __launch_bounds__(kBlock) k_nrc_scatter(FwdArgs a) {
  const int lane = int(__lane_id());
  const uint32_t T_ = 1u << a.G.log2T;
  const uint32_t waves = (a.n + 63u) / 64u;
  const uint32_t wpb = blockDim.x >> 6;
  const float* fg = reinterpret_cast<const float*>(a.fgrad);
  for (uint32_t w = blockIdx.x * wpb + (threadIdx.x >> 6); w < waves; w += gridDim.x * wpb) {
    const uint32_t row = w * 64u + uint32_t(lane);
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (row < a.n) {
      const float4 r0 = a.rec[4 * size_t(a.perm ? a.perm[row] : row)];
      px = fminf(fmaxf(r0.x, 0.0f), 1.0f);
      py = fminf(fmaxf(r0.y, 0.0f), 1.0f);
      pz = fminf(fmaxf(r0.z, 0.0f), 1.0f);
    }
    {
      const int comp = lane & 1, k = (lane >> 1) & 7;
#pragma unroll 1
      for (int r4 = 0; r4 < 16; ++r4) {
        const int from = 4 * r4 + (lane >> 4);
        const float qx = __shfl(px, from), qy = __shfl(py, from), qz = __shfl(pz, from);
        const uint32_t rq = w * 64u + uint32_t(from);
        const bool vq = rq < a.n;
#pragma unroll 2
        for (int lv = 0; lv < kFineLevel0; ++lv) {
          const Cell c = cell_of(a.G, lv, qx, qy, qz);
          const uint32_t idx =
              grid_index(a.G, lv, c.c[0] + (k & 1), c.c[1] + ((k >> 1) & 1), c.c[2] + (k >> 2));
          float v = vq ? corner_w(c, k) * fg[2 * (size_t(lv) * a.npad + rq) + comp] : 0.0f;
          // lanes l, l ^ 16, l ^ 32, l ^ 48 hold the same (corner, half) of
          // the four rows: fold equal entries pairwise into the lower lane
          bool live = true;
#pragma unroll
          for (int m = 16; m <= 32; m <<= 1) {
            const uint32_t pi = __shfl_xor(idx, m);
            const float pv = __shfl_xor(v, m);
            const bool plive = __shfl_xor(int(live), m) != 0;
            if (pi == idx && live && plive) {
              if (lane & m)
                live = false;
              else
                v += pv;
            }
          }
          if (live && v != 0.0f) atomicAdd(a.table_grad + 2 * (size_t(lv) * T_ + idx) + comp, v);
        }
      }
    }
    {
      const int k = lane & 7;
#pragma unroll 1
      for (int r8 = 0; r8 < 8; ++r8) {
        const int from = 8 * r8 + (lane >> 3);
        const float qx = __shfl(px, from), qy = __shfl(py, from), qz = __shfl(pz, from);
        const uint32_t rq = w * 64u + uint32_t(from);
        if (rq >= a.n) continue;
#pragma unroll 2
        for (int lv = kFineLevel0; lv < kLevels; ++lv) {
          const float2 d = a.fgrad[size_t(lv) * a.npad + rq];
          if (d.x == 0.0f && d.y == 0.0f) continue;
          const Cell c = cell_of(a.G, lv, qx, qy, qz);
          const uint32_t idx =
              grid_index(a.G, lv, c.c[0] + (k & 1), c.c[1] + ((k >> 1) & 1), c.c[2] + (k >> 2));
          const float wk = corner_w(c, k);
          const g2v hv = {(__bf16)(wk * d.x), (__bf16)(wk * d.y)};
          __builtin_amdgcn_global_atomic_fadd_v2bf16(
              reinterpret_cast<g2v*>(a.table_grad16) + size_t(lv - kFineLevel0) * T_ + idx, hv);
        }
      }
    }
  }
}

// ---- weight gradients: dW_L = H_L^T dZ_L, one partial per row chunk ----------
//
// H_L and dZ_L are row-major [row][64] 16-bit buffers. Each 32-row step the
// block stages both tiles in LDS, transposed to [feature][row], so that a
// lane's MFMA operand (8 consecutive rows of one feature) is one 16-byte LDS
// read. Wave w owns input features 16 w .. 16 w + 15.

constexpr int kWgLds = 40;  // LDS row stride in halves (16-byte aligned rows)

template <int P>
__global__ void __launch_bounds__(kBlock) k_nrc_wgrad(const uint16_t* HT, const uint16_t* DT, uint32_t npad,
                                                      uint32_t rows_per_chunk, float* partial) {
  typedef Op<P> O;
  __shared__ __align__(16) uint16_t sH[64 * kWgLds];
  __shared__ __align__(16) uint16_t sD[64 * kWgLds];
  const int L = blockIdx.y;
  const uint32_t chunk = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = int(__lane_id());
  const int mt = t >> 6;
  const int nout = L < kHidden ? 64 : 16;
  const int nts = nout / 16;
  const uint16_t* H = HT + size_t(L) * 64 * npad;
  const uint16_t* D = DT + size_t(L) * 64 * npad;
  const int lr = t >> 3, lc = 8 * (t & 7);  // staging: row, first feature
  const bool dload = lc < nout;
  uint32_t b_begin = chunk * rows_per_chunk, b_end = min(npad, b_begin + rows_per_chunk);
  f4v acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc[nt] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t b0 = b_begin; b0 < b_end; b0 += 32) {
    uint4 h = *reinterpret_cast<const uint4*>(H + size_t(b0 + lr) * 64 + lc);
    uint4 d = make_uint4(0, 0, 0, 0);
    if (dload) d = *reinterpret_cast<const uint4*>(D + size_t(b0 + lr) * 64 + lc);
    uint32_t hw[4] = {h.x, h.y, h.z, h.w}, dw[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      sH[(lc + i) * kWgLds + lr] = uint16_t(hw[i >> 1] >> (16 * (i & 1)));
      if (dload) sD[(lc + i) * kWgLds + lr] = uint16_t(dw[i >> 1] >> (16 * (i & 1)));
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    uint4 ua = *reinterpret_cast<const uint4*>(&sH[(16 * mt + (lane & 15)) * kWgLds + kq]);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      if (nt >= nts) break;
      uint4 ub = *reinterpret_cast<const uint4*>(&sD[(16 * nt + (lane & 15)) * kWgLds + kq]);
      acc[nt] = O::mma(O::unpack(ua), O::unpack(ub), acc[nt]);
    }
    __syncthreads();
  }
  size_t base = size_t(chunk) * kMlp + size_t(L) * 4096;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    if (nt >= nts) break;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int in = 16 * mt + 4 * (lane >> 4) + j, out = 16 * nt + (lane & 15);
      partial[base + size_t(in) * nout + out] = acc[nt][j];
    }
  }
}

// ---- Adam ---------------------------------------------------------------------

struct AdamArgs {
  float lr, beta1, beta2, eps;
  float c1, c2;      // 1 / (1 - beta1^t), 1 / (1 - beta2^t)
  float gscale;      // 1 / (3 n)
};

__device__ __forceinline__ float adam_one(float p, float gr, float* m, float* v, const AdamArgs& A) {
  float mm = A.beta1 * *m + (1.0f - A.beta1) * gr;
  float vv = A.beta2 * *v + (1.0f - A.beta2) * gr * gr;
  *m = mm;
  *v = vv;
  float mh = mm * A.c1, vh = vv * A.c2;
  return p - A.lr * mh / (sqrtf(vh) + A.eps);
}

// Sum of the weight-gradient partials of weight blockIdx.x * 64 + (t % 64):
// the block's 4 waves each add every 4th chunk, then the four sums are
// combined in a fixed order (deterministic). Launch kMlp / 64 blocks of 256.
__device__ __forceinline__ float sum_partials(const float* partial, uint32_t nchunks, uint32_t* p_out) {
  __shared__ float red[4][64];
  const uint32_t t = threadIdx.x, p = blockIdx.x * 64u + (t & 63u);
  float g = 0.0f;
  for (uint32_t c = t >> 6; c < nchunks; c += 4) g += partial[size_t(c) * kMlp + p];
  red[t >> 6][t & 63u] = g;
  __syncthreads();
  *p_out = p;
  return ((red[0][t & 63u] + red[1][t & 63u]) + red[2][t & 63u]) + red[3][t & 63u];
}

__global__ void __launch_bounds__(256) k_nrc_adam_mlp(const float* partial, uint32_t nchunks, float* master,
                                                      float* m, float* v, AdamArgs A) {
  uint32_t p;
  float g = sum_partials(partial, nchunks, &p);
  if (threadIdx.x < 64) master[p] = adam_one(master[p], g * A.gscale, &m[p], &v[p], A);
}

__global__ void __launch_bounds__(256) k_nrc_reduce_partials(const float* partial, uint32_t nchunks, float* out,
                                                             float gscale) {
  uint32_t p;
  float g = sum_partials(partial, nchunks, &p);
  if (threadIdx.x < 64) out[p] = g * gscale;
}

// Entries below `fine0` take their gradient from the FP32 pairs, the rest
// from the BF16 pairs of grad16; both are cleared for the next step.
__global__ void k_nrc_adam_table(float* master, float* grad, uint32_t* grad16, size_t fine0, float* m, float* v,
                                 uint32_t* table16, size_t n_entries, AdamArgs A) {
  size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_entries) return;
  float2 p = reinterpret_cast<float2*>(master)[e];
  float2 gr;
  if (e < fine0) {
    gr = reinterpret_cast<float2*>(grad)[e];
    reinterpret_cast<float2*>(grad)[e] = make_float2(0.0f, 0.0f);
  } else {
    const g2v h = reinterpret_cast<const g2v*>(grad16)[e - fine0];
    gr = make_float2(float(h.x), float(h.y));
    grad16[e - fine0] = 0u;
  }
  float2 mm = reinterpret_cast<float2*>(m)[e];
  float2 vv = reinterpret_cast<float2*>(v)[e];
  p.x = adam_one(p.x, gr.x * A.gscale, &mm.x, &vv.x, A);
  p.y = adam_one(p.y, gr.y * A.gscale, &mm.y, &vv.y, A);
  reinterpret_cast<float2*>(master)[e] = p;
  reinterpret_cast<float2*>(m)[e] = mm;
  reinterpret_cast<float2*>(v)[e] = vv;
  uint16_t lo = __builtin_bit_cast(uint16_t, (_Float16)p.x), hi = __builtin_bit_cast(uint16_t, (_Float16)p.y);
  table16[e] = uint32_t(lo) | (uint32_t(hi) << 16);
}

// ---- packing ------------------------------------------------------------------

// One block per layer. Forward A operand of (mt, ks, lane, i):
//   W[kappa(lane/16, ks, i)][16 mt + lane%16]; backward: W^T likewise, i.e.
//   W[16 mt + lane%16][kappa(lane/16, ks, i)]; the output layer's backward
//   K index is the output 4*(lane/16) + i (i < 4, else 0).
template <typename E>
__device__ __forceinline__ E cvt_e(float v);
template <>
__device__ __forceinline__ _Float16 cvt_e<_Float16>(float v) { return (_Float16)v; }
template <>
__device__ __forceinline__ __bf16 cvt_e<__bf16>(float v) { return (__bf16)v; }

template <typename E>
__global__ void k_nrc_repack(const float* master, E* fwd, E* bwd, uint8_t* fwd8, float* wscale,
                             const uint32_t* act_amax, float* ascale) {
  const int L = blockIdx.x;
  const float* W = master + size_t(L) * 4096;
  const int nout = L < kHidden ? 64 : 16;
  __shared__ float red[kBlock];
  float m = 0.0f;
  for (int i = threadIdx.x; i < 64 * nout; i += blockDim.x) m = fmaxf(m, fabsf(W[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  float sw = red[0] > 0.0f ? 448.0f / red[0] : 1.0f;
  // blocks (L, 0 .. gridDim.y - 1) share layer L's elements
  const int e0 = int(blockIdx.y * blockDim.x + threadIdx.x), es = int(gridDim.y * blockDim.x);
  if (threadIdx.x == 0 && blockIdx.y == 0 && fwd8) {
    wscale[L] = sw;
    float am = __uint_as_float(act_amax[L]);
    ascale[L] = am > 0.0f ? 448.0f / am : 1.0f;
  }
  // forward
  const int n_mt = L < kHidden ? 4 : 1;
  const int fbase = L * 8 * 64 * 8;  // elements (8 per lane per (mt, ks))
  for (int e = e0; e < n_mt * 2 * 64 * 8; e += es) {
    int i = e & 7, lane = (e >> 3) & 63, ks = (e >> 9) & 1, mt = e >> 10;
    int kin = kappa(lane >> 4, ks, i), out = 16 * mt + (lane & 15);
    float w = W[size_t(kin) * nout + out];
    fwd[fbase + e] = cvt_e<E>(w);
    if (fwd8) {
      float q = sat448(w * sw);
      int pk = __builtin_amdgcn_cvt_pk_fp8_f32(q, 0.0f, 0, false);
      fwd8[fbase + e] = uint8_t(pk & 0xFF);
    }
  }
  if (!bwd) return;
  // backward
  if (L < kHidden) {
    for (int e = e0; e < 4 * 2 * 64 * 8; e += es) {
      int i = e & 7, lane = (e >> 3) & 63, ks = (e >> 9) & 1, mt = e >> 10;
      int in = 16 * mt + (lane & 15), kout = kappa(lane >> 4, ks, i);
      bwd[L * 8 * 64 * 8 + e] = cvt_e<E>(W[size_t(in) * 64 + kout]);
    }
  } else {
    for (int e = e0; e < 4 * 64 * 8; e += es) {
      int i = e & 7, lane = (e >> 3) & 63, mt = e >> 9;
      int in = 16 * mt + (lane & 15), out = 4 * (lane >> 4) + i;
      bwd[kHidden * 8 * 64 * 8 + e] = cvt_e<E>(i < 4 ? W[size_t(in) * 16 + out] : 0.0f);
    }
  }
}

__global__ void k_nrc_init_table(float* master, uint32_t* table16, size_t n_entries, uint32_t seed) {
  size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_entries) return;
  aqd::Pcg r = aqd::pcg_seed(seed, e);
  float a = (float(aqd::u01(aqd::pcg_next(r))) * 2.0f - 1.0f) * 1e-4f;
  float b = (float(aqd::u01(aqd::pcg_next(r))) * 2.0f - 1.0f) * 1e-4f;
  master[2 * e] = a;
  master[2 * e + 1] = b;
  uint16_t lo = __builtin_bit_cast(uint16_t, (_Float16)a), hi = __builtin_bit_cast(uint16_t, (_Float16)b);
  table16[e] = uint32_t(lo) | (uint32_t(hi) << 16);
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The BF16 gradient pairs of the fine levels as FP32 pairs (for reading the
// gradient back); clears them.
__global__ void k_grad16_to_f32(uint32_t* g16, size_t n, float* out) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const g2v h = reinterpret_cast<const g2v*>(g16)[i];
  reinterpret_cast<float2*>(out)[i] = make_float2(float(h.x), float(h.y));
  g16[i] = 0u;
}

}  // namespace

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

namespace {

template <class T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (count <= n && p) return AQUA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return aqua::fail(e == hipErrorOutOfMemory ? AQUA_ERR_OOM : AQUA_ERR_HIP, "hipMalloc(%zu): %s",
                        count * sizeof(T), hipGetErrorString(e));
    }
    n = count;
    return AQUA_OK;
  }
};

int blocks_for(size_t n, int per = kBlock) { return int((n + per - 1) / per); }

// Events of a cache: 0-7 time the training step, kEvFork / kEvJoin order its
// two streams.
constexpr int kEvFork = 8, kEvJoin = 9, kEvCount = 10;

}  // namespace

namespace aqua {
hipError_t query_sort_storage(uint32_t n, size_t* bytes);
hipError_t query_sort(void* temp, size_t temp_bytes, const float4* rec, const uint32_t* d_count, uint32_t n,
                      uint32_t* keys, uint32_t* idx, hipStream_t s, const uint32_t** perm);
}  // namespace aqua

struct aqua_nrc {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // the training step's second stream
  aqua_nrc_cfg cfg{};
  Grid grid{};
  size_t entries = 0;  // levels * T
  size_t fine_entries() const { return entries / kLevels * (kLevels - kFineLevel0); }
  uint32_t step = 0;
  int cus = 256;
  int train_bpc = 1;  // resident training blocks per CU
  Buf<float> tmaster, tgrad, tm, tv;
  Buf<uint32_t> tgrad16;  // gradient of levels >= kFineLevel0, BF16 pairs
  Buf<uint32_t> t16;
  Buf<float> mmaster, mm, mv, partial;
  Buf<uint16_t> pack_fwd, pack_bwd;  // FP16 or BF16 bits, per training precision
  Buf<uint16_t> pack_fwd_other;      // the other 16-bit precision, for inference in it
  Buf<uint8_t> pack_fp8;
  Buf<float> wscale, ascale, loss;
  Buf<uint32_t> act_amax;
  Buf<uint16_t> HT, DT;
  Buf<float2> fgrad;         // [level][npad] feature-pair gradients, forward/backward -> scatter
  Buf<uint32_t> skey, sidx;  // training batch sort: Morton keys, row order (double buffers)
  Buf<uint8_t> stemp;
  // a sort queued ahead of the training step by nrc_presort_rows: the rows it
  // ordered (pre_rec) and their order, for the next nrc_train_device call
  const float4* pre_rec = nullptr;
  const uint32_t* pre_perm = nullptr;
  const uint32_t* pre_sidx = nullptr;  // the order buffer it lives in (a regrown one drops it)
  uint32_t pre_cap = 0;
  uint32_t cap_rows = 0;
  Buf<float4> rec_buf, tgt_buf;
  Buf<float> out_buf;
  hipEvent_t ev[kEvCount] = {};
  // checkpoint (aqua_nrc_checkpoint): parameters, Adam moments, FP16 table,
  // activation amax and the step count
  Buf<float> ck_tmaster, ck_tm, ck_tv, ck_mmaster, ck_mm, ck_mv;
  Buf<uint32_t> ck_t16, ck_amax;
  uint32_t ck_step = 0;
  bool ck_valid = false;

  ~aqua_nrc() {
    if (stream || side) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (side) (void)hipStreamDestroy(side);
  }

  // Grows to the next power of two of rows (at least 2^16), so a batch that
  // varies from frame to frame reallocates (a device-wide hipMalloc, about a
  // millisecond) only a few times.
  int ensure_rows(uint32_t n) {
    uint32_t npad = (n + 31u) & ~31u;
    if (npad <= cap_rows) return AQUA_OK;
    npad = std::max<uint32_t>(npad, 1u << 16);
    if (npad & (npad - 1u)) npad = 1u << (32 - __builtin_clz(npad));
    if (int e = HT.alloc(size_t(5) * 64 * npad)) return e;
    if (int e = DT.alloc(size_t(5) * 64 * npad)) return e;
    if (int e = fgrad.alloc(size_t(kLevels) * npad)) return e;
    uint32_t nchunks = max_chunks(npad);
    if (int e = partial.alloc(size_t(nchunks) * kMlp)) return e;
    if (int e = skey.alloc(size_t(2) * npad)) return e;
    if (int e = sidx.alloc(size_t(2) * npad)) return e;
    cap_rows = npad;
    return AQUA_OK;
  }

  // Rebuilds the MFMA operand packs from the FP32 master weights: forward +
  // backward + FP8 in the training precision, forward only in the other one.
  // Stream-ordered on s (the stream the weights were updated on).
  int repack(hipStream_t s) {
    const dim3 g(5, 8);  // layer x 8 slices of its elements
    if (cfg.train_precision == AQUA_NRC_BF16) {
      k_nrc_repack<__bf16><<<g, kBlock, 0, s>>>(mmaster.p, reinterpret_cast<__bf16*>(pack_fwd.p),
                                                reinterpret_cast<__bf16*>(pack_bwd.p), pack_fp8.p, wscale.p,
                                                act_amax.p, ascale.p);
      k_nrc_repack<_Float16><<<g, kBlock, 0, s>>>(mmaster.p, reinterpret_cast<_Float16*>(pack_fwd_other.p),
                                                  nullptr, nullptr, wscale.p, act_amax.p, ascale.p);
    } else {
      k_nrc_repack<_Float16><<<g, kBlock, 0, s>>>(mmaster.p, reinterpret_cast<_Float16*>(pack_fwd.p),
                                                  reinterpret_cast<_Float16*>(pack_bwd.p), pack_fp8.p, wscale.p,
                                                  act_amax.p, ascale.p);
      k_nrc_repack<__bf16><<<g, kBlock, 0, s>>>(mmaster.p, reinterpret_cast<__bf16*>(pack_fwd_other.p), nullptr,
                                                nullptr, wscale.p, act_amax.p, ascale.p);
    }
    if (hipGetLastError() != hipSuccess) return aqua::fail(AQUA_ERR_HIP, "repack launch failed");
    return AQUA_OK;
  }

  FwdArgs args(uint32_t n, const float4* rec, const float4* tgt) {
    FwdArgs a;
    memset(&a, 0, sizeof a);
    a.G = grid;
    a.rec = rec;
    a.tgt = tgt;
    a.n = n;
    a.npad = (n + 31u) & ~31u;
    a.table16 = t16.p;
    a.table_grad = tgrad.p;
    a.table_grad16 = tgrad16.p;
    a.HT = HT.p;
    a.DT = DT.p;
    a.fgrad = fgrad.p;
    a.loss_acc = loss.p;
    a.act_amax = act_amax.p;
    return a;
  }

  // Training forward + backward of a's rows: activations and feature
  // gradients.
  void launch_fwd(const FwdArgs& a, hipStream_t s) const {
    const size_t lds = kLdsTrainWeights ? size_t(kFwdPack + kBwdPack) * 16 : 0;
    if (cfg.train_precision == AQUA_NRC_BF16)
      k_nrc_fwd<AQUA_NRC_BF16, true><<<train_grid(a.npad), kTrainBlock, lds, s>>>(a);
    else
      k_nrc_fwd<AQUA_NRC_F16, true><<<train_grid(a.npad), kTrainBlock, lds, s>>>(a);
  }
  // The feature gradients into the table gradient.
  void launch_scatter(const FwdArgs& a, hipStream_t s) const {
    const int blocks = int(((a.n + 63u) / 64u + 3u) / 4u);
    k_nrc_scatter<<<std::max(1, std::min(blocks, cus * 8)), kBlock, 0, s>>>(a);
  }

  // Blocks of 4 waves, one wave per 16-row tile up to 32 waves per CU; a
  // multiple of 8 blocks (the kernel splits the tiles by XCD).
  int fwd_grid(uint32_t npad) const {
    int tiles = int(npad / 16);
    int waves = std::min(tiles, cus * 32);
    int blocks = (waves + 3) / 4;
    return std::max(8, (blocks + 7) & ~7);
  }
  // Training: resident blocks only (each stages the weight operands once),
  // a multiple of 8.
  int train_grid(uint32_t npad) const {
    const int wpb = kTrainBlock / 64;
    int blocks = std::min((int(npad / 16) + wpb - 1) / wpb, cus * train_bpc);
    return std::max(8, (blocks + 7) & ~7);
  }
};

namespace aqua {

// For the renderer: device-side query of up to `n_max` records already on the
// NRC's device; the actual count is read by the kernel from `d_count` (or is
// n_max when d_count is null). Writes 3 floats per record to d_out. With
// d_perm, row j of the pass is record d_perm[j] (a spatial order of the
// records; output still goes to the record's own slot).
// Stream-ordered on `s`, no host synchronisation.
int nrc_infer_device(aqua_nrc* nrc, uint32_t n_max, const float4* d_rec, float* d_out, uint32_t precision,
                     hipStream_t s, const uint32_t* d_count, const uint32_t* d_perm) {
  if (n_max == 0) return AQUA_OK;
  FwdArgs a = nrc->args(n_max, d_rec, nullptr);
  a.out = d_out;
  a.n_ptr = d_count;
  a.perm = d_perm;
  a.wscale = nrc->wscale.p;
  a.ascale = nrc->ascale.p;
  uint32_t npad = a.npad;
  if (precision == AQUA_NRC_FP8) {
    a.pack_fwd = nrc->pack_fp8.p;
    k_nrc_fwd<AQUA_NRC_FP8, false><<<nrc->fwd_grid(npad), kBlock, kFwdPack * 8, s>>>(a);
  } else {
    bool native = precision == nrc->cfg.train_precision;
    a.pack_fwd = native ? nrc->pack_fwd.p : nrc->pack_fwd_other.p;
    if (precision == AQUA_NRC_BF16)
      k_nrc_fwd<AQUA_NRC_BF16, false><<<nrc->fwd_grid(npad), kBlock, kFwdPack * 16, s>>>(a);
    else
      k_nrc_fwd<AQUA_NRC_F16, false><<<nrc->fwd_grid(npad), kBlock, kFwdPack * 16, s>>>(a);
  }
  HIPCHK(hipGetLastError());
  return AQUA_OK;
}

// Queues the Morton sort of the training rows at d_rec (at most n_max; the
// real count *d_count stays on the device) on stream s, so that it runs right
// after the work that wrote them instead of after the host has read the
// count. The next nrc_train_device call on the same rows (d_rec, n =
// *d_count) uses this order instead of sorting; any other call drops it.
int nrc_presort_rows(aqua_nrc* nrc, uint32_t n_max, const float4* d_rec, const uint32_t* d_count, hipStream_t s) {
  nrc->pre_rec = nullptr;
  if (n_max == 0 || !d_count) return AQUA_OK;
  if (int e = nrc->skey.alloc(size_t(2) * n_max)) return e;
  if (int e = nrc->sidx.alloc(size_t(2) * n_max)) return e;
  size_t need = 0;
  HIPCHK(aqua::query_sort_storage(n_max, &need));
  if (int e = nrc->stemp.alloc(need)) return e;
  HIPCHK(aqua::query_sort(nrc->stemp.p, nrc->stemp.n, d_rec, d_count, n_max, nrc->skey.p, nrc->sidx.p, s,
                          &nrc->pre_perm));
  nrc->pre_rec = d_rec;
  nrc->pre_sidx = nrc->sidx.p;
  nrc->pre_cap = n_max;
  return AQUA_OK;
}

// Forgets a queued-ahead sort (its rows are about to change).
void nrc_presort_drop(aqua_nrc* nrc) { nrc->pre_rec = nullptr; }

// The last training step's loss sum (over its rows x 3 channels) copied to
// host memory on stream s (pinned memory keeps the copy asynchronous).
int nrc_loss_async(aqua_nrc* nrc, float* host_dst, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(host_dst, nrc->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  return AQUA_OK;
}

// `steps` Adam steps on n samples already on the device, stream-ordered on
// `s`. With stats, the call waits for the steps and reports their timing and
// the last loss; without, it returns at once (the loss stays on the device).
// With before_update, the first weight update also waits for that event (the
// weights may be read on another stream until then; the forward and
// backward passes only read them).
int nrc_train_device(aqua_nrc* nrc, uint32_t n, const float4* d_rec, const float4* d_tgt, uint32_t steps,
                     aqua_nrc_stats* stats, hipStream_t s, hipEvent_t before_update) {
  if (n == 0 || steps == 0) {
    nrc->pre_rec = nullptr;
    return AQUA_OK;
  }
  if (int e = nrc->ensure_rows(n)) return e;
  FwdArgs a = nrc->args(n, d_rec, d_tgt);
  a.pack_fwd = nrc->pack_fwd.p;
  a.pack_bwd = nrc->pack_bwd.p;
  // rows in Morton order: the scatter merges rows that share coarse entries
  // (sorted here unless nrc_presort_rows already queued these rows' sort)
  if (nrc->pre_rec == d_rec && n <= nrc->pre_cap && nrc->pre_sidx == nrc->sidx.p) {
    a.perm = nrc->pre_perm;
  } else {
    size_t need = 0;
    HIPCHK(aqua::query_sort_storage(n, &need));
    if (int e = nrc->stemp.alloc(need)) return e;
    HIPCHK(aqua::query_sort(nrc->stemp.p, nrc->stemp.n, d_rec, nullptr, n, nrc->skey.p, nrc->sidx.p, s,
                            &a.perm));
  }
  nrc->pre_rec = nullptr;
  uint32_t npad = a.npad;
  uint32_t nchunks = n_chunks(npad);
  double t_fb = 0, t_wg = 0, t_ad = 0, t_sc = 0;
  float last_loss = 0.0f;
  HIPCHK(hipEventRecord(nrc->ev[0], s));
  for (uint32_t it = 0; it < steps; ++it) {
    ++nrc->step;
    const aqua_nrc_cfg& c = nrc->cfg;
    AdamArgs A;
    A.lr = c.learning_rate;
    A.beta1 = c.beta1;
    A.beta2 = c.beta2;
    A.eps = c.eps;
    A.c1 = float(1.0 / (1.0 - std::pow(double(c.beta1), double(nrc->step))));
    A.c2 = float(1.0 / (1.0 - std::pow(double(c.beta2), double(nrc->step))));
    A.gscale = float(1.0 / (3.0 * double(n)));
    k_fill_f32<<<1, 64, 0, s>>>(nrc->loss.p, 1, 0.0f);
    k_fill_u32<<<1, 64, 0, s>>>(nrc->act_amax.p, 5, 0u);
    HIPCHK(hipEventRecord(nrc->ev[1], s));
    nrc->launch_fwd(a, s);
    // the weight gradients and the MLP's Adam step (side stream) run beside
    // the table-gradient scatter and the table's Adam step (this stream)
    HIPCHK(hipEventRecord(nrc->ev[kEvFork], s));
    hipStream_t s2 = nrc->side;
    HIPCHK(hipStreamWaitEvent(s2, nrc->ev[kEvFork], 0));
    HIPCHK(hipEventRecord(nrc->ev[6], s2));
    dim3 gw(nchunks, 5);
    if (c.train_precision == AQUA_NRC_BF16)
      k_nrc_wgrad<AQUA_NRC_BF16><<<gw, kBlock, 0, s2>>>(nrc->HT.p, nrc->DT.p, npad, chunk_rows(npad),
                                                         nrc->partial.p);
    else
      k_nrc_wgrad<AQUA_NRC_F16><<<gw, kBlock, 0, s2>>>(nrc->HT.p, nrc->DT.p, npad, chunk_rows(npad),
                                                        nrc->partial.p);
    HIPCHK(hipEventRecord(nrc->ev[7], s2));
    // the weights may still be read on another stream until before_update
    if (before_update && it == 0) HIPCHK(hipStreamWaitEvent(s2, before_update, 0));
    k_nrc_adam_mlp<<<kMlp / 64, 256, 0, s2>>>(nrc->partial.p, nchunks, nrc->mmaster.p, nrc->mm.p, nrc->mv.p, A);
    HIPCHK(hipEventRecord(nrc->ev[kEvJoin], s2));
    HIPCHK(hipEventRecord(nrc->ev[3], s));
    nrc->launch_scatter(a, s);
    HIPCHK(hipEventRecord(nrc->ev[2], s));
    if (before_update && it == 0) HIPCHK(hipStreamWaitEvent(s, before_update, 0));
    k_nrc_adam_table<<<blocks_for(nrc->entries), kBlock, 0, s>>>(
        nrc->tmaster.p, nrc->tgrad.p, nrc->tgrad16.p, nrc->entries - nrc->fine_entries(), nrc->tm.p, nrc->tv.p,
        nrc->t16.p, nrc->entries, A);
    HIPCHK(hipStreamWaitEvent(s, nrc->ev[kEvJoin], 0));
    if (int e = nrc->repack(s)) return e;
    HIPCHK(hipEventRecord(nrc->ev[4], s));
    HIPCHK(hipGetLastError());
    if (stats) {
      HIPCHK(hipEventSynchronize(nrc->ev[4]));
      float ms;
      HIPCHK(hipEventElapsedTime(&ms, nrc->ev[1], nrc->ev[2]));
      t_fb += ms;
      HIPCHK(hipEventElapsedTime(&ms, nrc->ev[3], nrc->ev[2]));
      t_sc += ms;
      HIPCHK(hipEventElapsedTime(&ms, nrc->ev[6], nrc->ev[7]));
      t_wg += ms;
      HIPCHK(hipEventElapsedTime(&ms, nrc->ev[2], nrc->ev[4]));
      t_ad += ms;
    }
  }
  HIPCHK(hipEventRecord(nrc->ev[5], s));
  if (stats) {
    HIPCHK(hipMemcpyAsync(&last_loss, nrc->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memset(stats, 0, sizeof *stats);
    float ms;
    HIPCHK(hipEventElapsedTime(&ms, nrc->ev[0], nrc->ev[5]));
    stats->ms_total = ms;
    stats->ms_fwd_bwd = t_fb;
    stats->ms_wgrad = t_wg;
    stats->ms_adam = t_ad;
    stats->ms_scatter = float(t_sc);
    stats->loss = last_loss / float(3.0 * double(n));
    stats->steps = steps;
    stats->rows = n;
  }
  return AQUA_OK;
}

}  // namespace aqua

namespace {

void init_mlp_host(uint32_t seed, std::vector<float>* w) {
  // Xavier / Glorot uniform, PCG32 stream `seed` (initseq 0x6e7263 = "nrc").
  w->assign(kMlp, 0.0f);
  uint64_t state = 0, inc = (0x6e7263ull << 1u) | 1u;
  auto next = [&]() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u), rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  };
  next();
  state += seed;
  next();
  for (int L = 0; L <= kHidden; ++L) {
    int nout = L < kHidden ? 64 : 16, nout_used = L < kHidden ? 64 : 3;
    double s = std::sqrt(6.0 / (64.0 + double(nout_used)));
    for (int i = 0; i < 64; ++i)
      for (int o = 0; o < nout; ++o) {
        double u = double(next() >> 8) * (1.0 / 16777216.0);
        (*w)[size_t(L) * 4096 + size_t(i) * nout + o] = o < nout_used ? float((2.0 * u - 1.0) * s) : 0.0f;
      }
  }
}

}  // namespace

extern "C" {

void aqua_nrc_default_cfg(aqua_nrc_cfg* c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->n_levels = 16;
  c->log2_table = 19;
  c->base_res = 16;
  c->level_scale = 1.5f;
  c->learning_rate = 1e-2f;
  c->beta1 = 0.9f;
  c->beta2 = 0.99f;
  c->eps = 1e-15f;
  c->train_precision = AQUA_NRC_F16;
  c->seed = 0;
  c->max_batch = 1u << 18;
}

int aqua_nrc_create(int device, const aqua_nrc_cfg* cfg, aqua_nrc** out) {
  if (!cfg || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_nrc_create: null argument");
  *out = nullptr;
  if (cfg->n_levels != uint32_t(kLevels))
    return aqua::fail(AQUA_ERR_UNSUPPORTED, "aqua_nrc_create: n_levels must be %d", kLevels);
  if (cfg->log2_table < 10 || cfg->log2_table > 22)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_nrc_create: log2_table out of [10, 22]");
  if (cfg->train_precision > AQUA_NRC_BF16)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_nrc_create: training precision must be FP16 or BF16");
  if (!(cfg->learning_rate > 0.0f) || !(cfg->level_scale >= 1.0f) || cfg->base_res < 2)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_nrc_create: bad hyper-parameters");
  if (device < 0) return aqua::fail(AQUA_ERR_INVALID_ARG, "device %d: the cache runs on a GPU only", device);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return aqua::fail(AQUA_ERR_NO_DEVICE, "no HIP device available");
  if (device >= count) return aqua::fail(AQUA_ERR_NO_DEVICE, "device %d out of range", device);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return aqua::fail(AQUA_ERR_NO_DEVICE, "device %d is %s; this build targets gfx950", device, prop.gcnArchName);
  HIPCHK(hipSetDevice(device));
  auto r = std::make_unique<aqua_nrc>();
  r->device = device;
  r->cfg = *cfg;
  r->cus = prop.multiProcessorCount;
  {
    int bpc = 0;
    const size_t lds = kLdsTrainWeights ? size_t(kFwdPack + kBwdPack) * 16 : 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_nrc_fwd<AQUA_NRC_F16, true>, kTrainBlock, lds));
    r->train_bpc = std::max(1, bpc);
  }
  HIPCHK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
  for (hipEvent_t& e : r->ev) HIPCHK(hipEventCreate(&e));
  const uint32_t T = 1u << cfg->log2_table;
  r->grid.log2T = cfg->log2_table;
  r->grid.tmask = T - 1u;
  r->grid.dense_mask = 0;
  for (int l = 0; l < kLevels; ++l) {
    double res = std::floor(double(cfg->base_res) * std::pow(double(cfg->level_scale), double(l)));
    r->grid.res[l] = float(res);
    double n1 = res + 1.0;
    if (n1 * n1 * n1 <= double(T)) r->grid.dense_mask |= 1u << l;
  }
  r->entries = size_t(kLevels) * T;
  hipStream_t s = r->stream;
  if (int e = r->tmaster.alloc(2 * r->entries)) return e;
  if (int e = r->tgrad.alloc(2 * r->entries)) return e;
  if (int e = r->tgrad16.alloc(r->fine_entries())) return e;
  if (int e = r->tm.alloc(2 * r->entries)) return e;
  if (int e = r->tv.alloc(2 * r->entries)) return e;
  if (int e = r->t16.alloc(r->entries + 2)) return e;  // + 2 spare words: see encode16
  if (int e = r->mmaster.alloc(kMlp)) return e;
  if (int e = r->mm.alloc(kMlp)) return e;
  if (int e = r->mv.alloc(kMlp)) return e;
  if (int e = r->pack_fwd.alloc(size_t(kFwdPack) * 8)) return e;
  if (int e = r->pack_fwd_other.alloc(size_t(kFwdPack) * 8)) return e;
  if (int e = r->pack_bwd.alloc(size_t(kBwdPack) * 8)) return e;
  if (int e = r->pack_fp8.alloc(size_t(kFwdPack) * 8)) return e;
  if (int e = r->wscale.alloc(8)) return e;
  if (int e = r->ascale.alloc(8)) return e;
  if (int e = r->loss.alloc(1)) return e;
  if (int e = r->act_amax.alloc(8)) return e;
  if (int e = r->ensure_rows(std::max(32u, cfg->max_batch))) return e;
  k_nrc_init_table<<<blocks_for(r->entries), kBlock, 0, s>>>(r->tmaster.p, r->t16.p, r->entries, cfg->seed);
  k_fill_u32<<<1, kBlock, 0, s>>>(r->t16.p + r->entries, 2, 0u);
  k_fill_f32<<<blocks_for(2 * r->entries), kBlock, 0, s>>>(r->tgrad.p, 2 * r->entries, 0.0f);
  k_fill_u32<<<blocks_for(r->fine_entries()), kBlock, 0, s>>>(r->tgrad16.p, r->fine_entries(), 0u);
  k_fill_f32<<<blocks_for(2 * r->entries), kBlock, 0, s>>>(r->tm.p, 2 * r->entries, 0.0f);
  k_fill_f32<<<blocks_for(2 * r->entries), kBlock, 0, s>>>(r->tv.p, 2 * r->entries, 0.0f);
  k_fill_f32<<<blocks_for(kMlp), kBlock, 0, s>>>(r->mm.p, kMlp, 0.0f);
  k_fill_f32<<<blocks_for(kMlp), kBlock, 0, s>>>(r->mv.p, kMlp, 0.0f);
  k_fill_u32<<<1, 64, 0, s>>>(r->act_amax.p, 8, 0u);
  std::vector<float> w;
  init_mlp_host(cfg->seed, &w);
  HIPCHK(hipMemcpyAsync(r->mmaster.p, w.data(), kMlp * sizeof(float), hipMemcpyHostToDevice, s));
  if (int e = r->repack(r->stream)) return e;
  HIPCHK(hipStreamSynchronize(s));
  *out = r.release();
  return AQUA_OK;
}

void aqua_nrc_destroy(aqua_nrc* nrc) { delete nrc; }

int aqua_nrc_param_counts(const aqua_nrc* nrc, uint64_t* n_table, uint32_t* n_mlp) {
  if (!nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "null nrc");
  if (n_table) *n_table = 2 * uint64_t(nrc->entries);
  if (n_mlp) *n_mlp = kMlp;
  return AQUA_OK;
}

int aqua_nrc_get_params(aqua_nrc* nrc, float* table, float* mlp) {
  if (!nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "null nrc");
  HIPCHK(hipSetDevice(nrc->device));
  if (table)
    HIPCHK(hipMemcpyAsync(table, nrc->tmaster.p, 2 * nrc->entries * sizeof(float), hipMemcpyDeviceToHost,
                          nrc->stream));
  if (mlp) HIPCHK(hipMemcpyAsync(mlp, nrc->mmaster.p, kMlp * sizeof(float), hipMemcpyDeviceToHost, nrc->stream));
  HIPCHK(hipStreamSynchronize(nrc->stream));
  return AQUA_OK;
}

int aqua_nrc_checkpoint(aqua_nrc* nrc, int restore) {
  if (!nrc) return aqua::fail(AQUA_ERR_INVALID_ARG, "null nrc");
  HIPCHK(hipSetDevice(nrc->device));
  hipStream_t s = nrc->stream;
  HIPCHK(hipDeviceSynchronize());  // the cache may have been used on any stream
  const size_t nt = 2 * nrc->entries;
  struct Pair {
    void* live;
    void* saved;
    size_t bytes;
  };
  if (!restore) {
    if (int e = nrc->ck_tmaster.alloc(nt)) return e;
    if (int e = nrc->ck_tm.alloc(nt)) return e;
    if (int e = nrc->ck_tv.alloc(nt)) return e;
    if (int e = nrc->ck_t16.alloc(nrc->entries)) return e;
    if (int e = nrc->ck_mmaster.alloc(kMlp)) return e;
    if (int e = nrc->ck_mm.alloc(kMlp)) return e;
    if (int e = nrc->ck_mv.alloc(kMlp)) return e;
    if (int e = nrc->ck_amax.alloc(8)) return e;
  } else if (!nrc->ck_valid) {
    return aqua::fail(AQUA_ERR_INVALID_ARG, "no checkpoint to restore");
  }
  const Pair pairs[] = {{nrc->tmaster.p, nrc->ck_tmaster.p, nt * 4},   {nrc->tm.p, nrc->ck_tm.p, nt * 4},
                        {nrc->tv.p, nrc->ck_tv.p, nt * 4},             {nrc->t16.p, nrc->ck_t16.p, nrc->entries * 4},
                        {nrc->mmaster.p, nrc->ck_mmaster.p, kMlp * 4}, {nrc->mm.p, nrc->ck_mm.p, kMlp * 4},
                        {nrc->mv.p, nrc->ck_mv.p, kMlp * 4},           {nrc->act_amax.p, nrc->ck_amax.p, 5 * 4}};
  for (const Pair& p : pairs)
    HIPCHK(hipMemcpyAsync(restore ? p.live : p.saved, restore ? p.saved : p.live, p.bytes, hipMemcpyDeviceToDevice, s));
  if (restore) {
    nrc->step = nrc->ck_step;
    if (int e = nrc->repack(s)) return e;  // operand packs and FP8 scales of the restored weights
  } else {
    nrc->ck_step = nrc->step;
    nrc->ck_valid = true;
  }
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

int aqua_nrc_fp8_scales(aqua_nrc* nrc, float* wscale, float* ascale) {
  if (!nrc || !wscale || !ascale) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipSetDevice(nrc->device));
  HIPCHK(hipMemcpyAsync(wscale, nrc->wscale.p, 5 * sizeof(float), hipMemcpyDeviceToHost, nrc->stream));
  HIPCHK(hipMemcpyAsync(ascale, nrc->ascale.p, 5 * sizeof(float), hipMemcpyDeviceToHost, nrc->stream));
  HIPCHK(hipStreamSynchronize(nrc->stream));
  return AQUA_OK;
}

int aqua_nrc_set_params(aqua_nrc* nrc, const float* table, const float* mlp) {
  if (!nrc || !table || !mlp) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipSetDevice(nrc->device));
  hipStream_t s = nrc->stream;
  for (size_t i = 0; i < 2 * nrc->entries; i += 4096)
    for (size_t k = i; k < std::min(2 * nrc->entries, i + 4096); ++k)
      if (!std::isfinite(table[k])) return aqua::fail(AQUA_ERR_NONFINITE, "non-finite table parameter");
  for (int k = 0; k < kMlp; ++k)
    if (!std::isfinite(mlp[k])) return aqua::fail(AQUA_ERR_NONFINITE, "non-finite MLP parameter");
  HIPCHK(hipMemcpyAsync(nrc->tmaster.p, table, 2 * nrc->entries * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(nrc->mmaster.p, mlp, kMlp * sizeof(float), hipMemcpyHostToDevice, s));
  // FP16 table copy from the master
  {
    std::vector<uint32_t> t16(nrc->entries);
    for (size_t e = 0; e < nrc->entries; ++e) {
      _Float16 a = (_Float16)table[2 * e], b = (_Float16)table[2 * e + 1];
      uint16_t ua, ub;
      memcpy(&ua, &a, 2);
      memcpy(&ub, &b, 2);
      t16[e] = uint32_t(ua) | (uint32_t(ub) << 16);
    }
    HIPCHK(hipMemcpyAsync(nrc->t16.p, t16.data(), nrc->entries * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  k_fill_f32<<<blocks_for(2 * nrc->entries), kBlock, 0, s>>>(nrc->tgrad.p, 2 * nrc->entries, 0.0f);
  k_fill_u32<<<blocks_for(nrc->fine_entries()), kBlock, 0, s>>>(nrc->tgrad16.p, nrc->fine_entries(), 0u);
  k_fill_f32<<<blocks_for(2 * nrc->entries), kBlock, 0, s>>>(nrc->tm.p, 2 * nrc->entries, 0.0f);
  k_fill_f32<<<blocks_for(2 * nrc->entries), kBlock, 0, s>>>(nrc->tv.p, 2 * nrc->entries, 0.0f);
  k_fill_f32<<<blocks_for(kMlp), kBlock, 0, s>>>(nrc->mm.p, kMlp, 0.0f);
  k_fill_f32<<<blocks_for(kMlp), kBlock, 0, s>>>(nrc->mv.p, kMlp, 0.0f);
  nrc->step = 0;
  if (int e = nrc->repack(nrc->stream)) return e;
  HIPCHK(hipStreamSynchronize(s));
  return AQUA_OK;
}

static int upload_records(aqua_nrc* nrc, uint32_t n, const float* records, const float* targets) {
  for (size_t i = 0; i < size_t(n) * 16; ++i)
    if (!std::isfinite(records[i])) return aqua::fail(AQUA_ERR_NONFINITE, "non-finite record value");
  if (int e = nrc->rec_buf.alloc(size_t(n) * 4)) return e;
  HIPCHK(hipMemcpyAsync(nrc->rec_buf.p, records, size_t(n) * 64, hipMemcpyHostToDevice, nrc->stream));
  if (targets) {
    std::vector<float4> t(n);
    for (uint32_t i = 0; i < n; ++i) {
      t[i] = make_float4(targets[3 * size_t(i)], targets[3 * size_t(i) + 1], targets[3 * size_t(i) + 2], 0.0f);
      if (!std::isfinite(t[i].x) || !std::isfinite(t[i].y) || !std::isfinite(t[i].z))
        return aqua::fail(AQUA_ERR_NONFINITE, "non-finite target");
    }
    if (int e = nrc->tgt_buf.alloc(n)) return e;
    HIPCHK(hipMemcpyAsync(nrc->tgt_buf.p, t.data(), size_t(n) * 16, hipMemcpyHostToDevice, nrc->stream));
    HIPCHK(hipStreamSynchronize(nrc->stream));
  }
  return AQUA_OK;
}

int aqua_nrc_infer(aqua_nrc* nrc, uint32_t n, const float* records, uint32_t precision, float* rgb) {
  if (!nrc || (n && (!records || !rgb))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (precision > AQUA_NRC_FP8) return aqua::fail(AQUA_ERR_INVALID_ARG, "unknown precision %u", precision);
  if (n == 0) return AQUA_OK;
  HIPCHK(hipSetDevice(nrc->device));
  if (int e = upload_records(nrc, n, records, nullptr)) return e;
  if (int e = nrc->out_buf.alloc(size_t(n) * 3)) return e;
  if (int e = aqua::nrc_infer_device(nrc, n, nrc->rec_buf.p, nrc->out_buf.p, precision, nrc->stream, nullptr,
                                     nullptr))
    return e;
  HIPCHK(hipMemcpyAsync(rgb, nrc->out_buf.p, size_t(n) * 12, hipMemcpyDeviceToHost, nrc->stream));
  HIPCHK(hipStreamSynchronize(nrc->stream));
  return AQUA_OK;
}

int aqua_nrc_query_order(aqua_nrc* nrc, uint32_t n, const float* records, uint32_t* order) {
  if (!nrc || (n && (!records || !order))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (n == 0) return AQUA_OK;
  HIPCHK(hipSetDevice(nrc->device));
  if (int e = upload_records(nrc, n, records, nullptr)) return e;
  if (int e = nrc->skey.alloc(2 * size_t(n))) return e;
  if (int e = nrc->sidx.alloc(2 * size_t(n))) return e;
  size_t need = 0;
  HIPCHK(aqua::query_sort_storage(n, &need));
  if (int e = nrc->stemp.alloc(need)) return e;
  const uint32_t* perm = nullptr;
  HIPCHK(aqua::query_sort(nrc->stemp.p, nrc->stemp.n, nrc->rec_buf.p, nullptr, n, nrc->skey.p, nrc->sidx.p,
                          nrc->stream, &perm));
  HIPCHK(hipMemcpyAsync(order, perm, size_t(n) * 4, hipMemcpyDeviceToHost, nrc->stream));
  HIPCHK(hipStreamSynchronize(nrc->stream));
  return AQUA_OK;
}

int aqua_nrc_train(aqua_nrc* nrc, uint32_t n, const float* records, const float* targets, uint32_t steps,
                   aqua_nrc_stats* stats) {
  if (!nrc || (n && (!records || !targets))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (n == 0) return AQUA_OK;
  HIPCHK(hipSetDevice(nrc->device));
  if (int e = upload_records(nrc, n, records, targets)) return e;
  aqua_nrc_stats st;
  if (int e = aqua::nrc_train_device(nrc, n, nrc->rec_buf.p, nrc->tgt_buf.p, steps, &st, nrc->stream, nullptr))
    return e;
  if (stats) *stats = st;
  return AQUA_OK;
}

int aqua_nrc_gradients(aqua_nrc* nrc, uint32_t n, const float* records, const float* targets, float* g_table,
                       float* g_mlp, float* loss) {
  if (!nrc || n == 0 || !records || !targets) return aqua::fail(AQUA_ERR_INVALID_ARG, "bad argument");
  HIPCHK(hipSetDevice(nrc->device));
  if (int e = upload_records(nrc, n, records, targets)) return e;
  if (int e = nrc->ensure_rows(n)) return e;
  hipStream_t s = nrc->stream;
  FwdArgs a = nrc->args(n, nrc->rec_buf.p, nrc->tgt_buf.p);
  a.pack_fwd = nrc->pack_fwd.p;
  a.pack_bwd = nrc->pack_bwd.p;
  uint32_t npad = a.npad, nchunks = n_chunks(npad);
  float gscale = float(1.0 / (3.0 * double(n)));
  k_fill_f32<<<1, 64, 0, s>>>(nrc->loss.p, 1, 0.0f);
  k_fill_u32<<<1, 64, 0, s>>>(nrc->act_amax.p, 5, 0u);
  nrc->launch_fwd(a, s);
  nrc->launch_scatter(a, s);
  dim3 gw(nchunks, 5);
  if (nrc->cfg.train_precision == AQUA_NRC_BF16)
    k_nrc_wgrad<AQUA_NRC_BF16><<<gw, kBlock, 0, s>>>(nrc->HT.p, nrc->DT.p, npad, chunk_rows(npad),
                                                        nrc->partial.p);
  else
    k_nrc_wgrad<AQUA_NRC_F16><<<gw, kBlock, 0, s>>>(nrc->HT.p, nrc->DT.p, npad, chunk_rows(npad),
                                                        nrc->partial.p);
  Buf<float> gm;
  if (int e = gm.alloc(kMlp)) return e;
  k_nrc_reduce_partials<<<kMlp / 64, 256, 0, s>>>(nrc->partial.p, nchunks, gm.p, gscale);
  HIPCHK(hipGetLastError());
  float lsum = 0.0f;
  HIPCHK(hipMemcpyAsync(&lsum, nrc->loss.p, sizeof lsum, hipMemcpyDeviceToHost, s));
  if (g_mlp) HIPCHK(hipMemcpyAsync(g_mlp, gm.p, kMlp * sizeof(float), hipMemcpyDeviceToHost, s));
  const size_t fine = nrc->fine_entries(), fine0 = nrc->entries - fine;
  k_grad16_to_f32<<<blocks_for(fine), kBlock, 0, s>>>(nrc->tgrad16.p, fine, nrc->tgrad.p + 2 * fine0);
  if (g_table)
    HIPCHK(hipMemcpyAsync(g_table, nrc->tgrad.p, 2 * nrc->entries * sizeof(float), hipMemcpyDeviceToHost, s));
  k_fill_f32<<<blocks_for(2 * nrc->entries), kBlock, 0, s>>>(nrc->tgrad.p, 2 * nrc->entries, 0.0f);
  HIPCHK(hipStreamSynchronize(s));
  if (g_table)
    for (size_t i = 0; i < 2 * nrc->entries; ++i) g_table[i] *= gscale;
  if (loss) *loss = lsum * gscale;
  return AQUA_OK;
}

int aqua_nrc_train_device(aqua_nrc* nrc, uint32_t n, const void* d_records, const void* d_targets, uint32_t steps,
                          aqua_nrc_stats* stats) {
  if (!nrc || (n && (!d_records || !d_targets))) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipSetDevice(nrc->device));
  if (int e = aqua::nrc_train_device(nrc, n, static_cast<const float4*>(d_records),
                                     static_cast<const float4*>(d_targets), steps, stats, nrc->stream, nullptr))
    return e;
  HIPCHK(hipStreamSynchronize(nrc->stream));
  return AQUA_OK;
}

}  // extern "C"
