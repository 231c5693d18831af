// Device math for the gfx950 path tracer.
//
// Every routine here has a fixed IEEE-754 FP32 operation sequence: the build
// uses -ffp-contract=off, so a*b+c is two roundings unless fmaf() is written
// out, and division and sqrtf are the correctly rounded ones. With the same
// sequence on the host, results are bit-identical, which is what lets a CPU
// restatement check the kernels pixel for pixel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define AQ_DEV __device__ __forceinline__

namespace aqd {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kFltMax = 3.402823466e+38f;

struct v3 {
  float x, y, z;
};

AQ_DEV v3 mk(float x, float y, float z) { return v3{x, y, z}; }
AQ_DEV v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AQ_DEV v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AQ_DEV v3 mul(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AQ_DEV v3 scl(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
AQ_DEV v3 neg(v3 a) { return v3{-a.x, -a.y, -a.z}; }
AQ_DEV float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AQ_DEV v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
AQ_DEV v3 normalize(v3 a) {
  float inv = 1.0f / __builtin_sqrtf(dot(a, a));
  return scl(a, inv);
}
AQ_DEV float fmin2(float a, float b) { return a < b ? a : b; }
AQ_DEV float fmax2(float a, float b) { return a > b ? a : b; }
AQ_DEV v3 xyz(float4 f) { return v3{f.x, f.y, f.z}; }

// ---- PCG32 (O'Neill 2014: PCG-XSH-RR, 64-bit LCG state, 32-bit output) ----
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

struct Pcg {
  uint64_t state, inc;
};

AQ_DEV uint32_t pcg_next(Pcg& r) {
  uint64_t old = r.state;
  r.state = old * kPcgMult + r.inc;
  uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = uint32_t(old >> 59u);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// pcg32_srandom_r(initstate, initseq).
AQ_DEV Pcg pcg_seed(uint64_t initstate, uint64_t initseq) {
  Pcg r;
  r.state = 0u;
  r.inc = (initseq << 1u) | 1u;
  pcg_next(r);
  r.state += initstate;
  pcg_next(r);
  return r;
}

// Jump the LCG ahead by `delta` steps in O(log delta) (Brown 1994,
// "Random number generation with arbitrary strides").
AQ_DEV void pcg_advance(Pcg& r, uint64_t delta) {
  uint64_t cur_mult = kPcgMult, cur_plus = r.inc, acc_mult = 1u, acc_plus = 0u;
  while (delta > 0) {
    if (delta & 1u) {
      acc_mult *= cur_mult;
      acc_plus = acc_plus * cur_mult + cur_plus;
    }
    cur_plus = (cur_mult + 1u) * cur_plus;
    cur_mult *= cur_mult;
    delta >>= 1u;
  }
  r.state = acc_mult * r.state + acc_plus;
}

// Top 24 bits -> [0, 1), exact.
AQ_DEV float u01(uint32_t x) { return float(x >> 8) * 5.9604644775390625e-8f; }

// ---- sin / cos of 2*pi*u, u in [0, 1) -------------------------------------
// Quadrant reduction is exact (u*4 and the integer split are exact in FP32),
// the remaining angle is folded to [0, pi/4] with the exact complement
// 1 - f, and sin / cos of (pi/2)*g come from their Taylor series in g
// (degree 9 / 10 in g; truncation < 3e-9 relative on [0, 1/2]).
AQ_DEV void sincos_2pi(float u, float* s, float* c) {
  float x = u * 4.0f;
  int q = int(x);
  float f = x - float(q);
  bool comp = f > 0.5f;
  float g = comp ? 1.0f - f : f;
  float g2 = g * g;
  float sp = 0.000160441184787359821f;
  sp = sp * g2 + -0.00468175413531868810f;
  sp = sp * g2 + 0.0796926262461670451f;
  sp = sp * g2 + -0.645964097506246254f;
  sp = sp * g2 + 1.57079632679489662f;
  float sg = sp * g;
  float cp = -2.52020423730606054e-05f;
  cp = cp * g2 + 0.000919260274839426580f;
  cp = cp * g2 + -0.0208634807633529609f;
  cp = cp * g2 + 0.253669507901048014f;
  cp = cp * g2 + -1.23370055013616983f;
  float cg = cp * g2 + 1.0f;
  float s0 = comp ? cg : sg;
  float c0 = comp ? sg : cg;
  q &= 3;
  if (q == 0) {
    *s = s0;
    *c = c0;
  } else if (q == 1) {
    *s = c0;
    *c = -s0;
  } else if (q == 2) {
    *s = -s0;
    *c = -c0;
  } else {
    *s = -c0;
    *c = s0;
  }
}

// Orthonormal basis around unit n (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited").
AQ_DEV void onb(v3 n, v3* t, v3* b) {
  float sign = __builtin_copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = (n.x * n.y) * a;
  *t = v3{1.0f + ((sign * n.x) * n.x) * a, sign * bb, -sign * n.x};
  *b = v3{bb, sign + (n.y * n.y) * a, -n.y};
}
AQ_DEV v3 from_local(v3 t, v3 b, v3 n, v3 l) {
  return v3{(t.x * l.x + b.x * l.y) + n.x * l.z, (t.y * l.x + b.y * l.y) + n.y * l.z,
            (t.z * l.x + b.z * l.y) + n.z * l.z};
}

// Ray origin pushed off a surface along n (the side the ray leaves on):
// Waechter & Binder 2019, "A Fast and Robust Method for Avoiding
// Self-Intersection" (Ray Tracing Gems, ch. 6): integer ULP offset, or a
// small float offset near the origin.
AQ_DEV v3 offset_ray(v3 p, v3 n) {
  const float kOrigin = 1.0f / 32.0f, kFloatScale = 1.0f / 65536.0f, kIntScale = 256.0f;
  int ox = int(kIntScale * n.x), oy = int(kIntScale * n.y), oz = int(kIntScale * n.z);
  float px = __int_as_float(__float_as_int(p.x) + (p.x < 0.0f ? -ox : ox));
  float py = __int_as_float(__float_as_int(p.y) + (p.y < 0.0f ? -oy : oy));
  float pz = __int_as_float(__float_as_int(p.z) + (p.z < 0.0f ? -oz : oz));
  return v3{__builtin_fabsf(p.x) < kOrigin ? p.x + kFloatScale * n.x : px,
            __builtin_fabsf(p.y) < kOrigin ? p.y + kFloatScale * n.y : py,
            __builtin_fabsf(p.z) < kOrigin ? p.z + kFloatScale * n.z : pz};
}

}  // namespace aqd
