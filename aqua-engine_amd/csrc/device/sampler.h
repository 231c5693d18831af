// Stratified sample patterns for the first path dimensions.
#pragma once

#include "device/common.h"

namespace aqd {

// ---- stratified sampler ------------------------------------------------
// With the integrator's "stratified" sampler the camera ray's two dimensions
// and the first hit's six come from shuffled, Owen-scrambled 2-D Sobol
// points over the sample index (Burley 2020, "Practical Hash-based Owen
// Scrambling", JCGT 9(4)); the PCG32 stream still supplies all the other
// draws and advances past the replaced ones. Every aligned block of 2^m of a
// pixel's samples is then a (0, m, 2)-net in each of its patterns:
// 0 = camera (u0, u1); first hit: 1 = light point (u1, u2), 2 = BSDF
// direction (u4, u5), 3 = light pick and lobe (u0, u3).

// Laine & Karras 2011 hash (the constants of Burley 2020, listing 2): each
// output bit depends only on the input bits at or below it.
AQ_DEV uint32_t lk_hash(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}

// Nested uniform scramble, most significant bit first.
AQ_DEV uint32_t owen_scramble(uint32_t x, uint32_t seed) {
  return __builtin_bitreverse32(lk_hash(__builtin_bitreverse32(x), seed));
}

// Integer finaliser "lowbias32" (C. Wellons, hash-prospector, public domain).
AQ_DEV uint32_t mix32(uint32_t x) {
  x = (x ^ (x >> 16)) * 0x7feb352du;
  x = (x ^ (x >> 15)) * 0x846ca68bu;
  return x ^ (x >> 16);
}

// Sobol's second dimension (direction numbers m_k = 1: column k + 1 is
// column k xor column k shifted right by one); the first is the bit reversal.
AQ_DEV uint32_t sobol_second(uint32_t i) {
  uint32_t r = 0u;
  for (uint32_t col = 0x80000000u; i != 0u; i >>= 1, col ^= col >> 1) r ^= (i & 1u) ? col : 0u;
  return r;
}

AQ_DEV void stratified2(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t pattern, uint32_t sample, float* a,
                        float* b) {
  const uint32_t h = mix32(mix32(mix32(mix32(seed) ^ pixel) ^ frame) ^ pattern);
  const uint32_t i = owen_scramble(sample, h);
  *a = u01(owen_scramble(__builtin_bitreverse32(i), mix32(h ^ 0xa511e9b3u)));
  *b = u01(owen_scramble(sobol_second(i), mix32(h ^ 0x63d83595u)));
}

}  // namespace aqd
