// Morton keys of cache-query positions: 10 bits per axis of the position
// normalised to [0, 1]^3, interleaved x-lowest (30 bits). Shared by the
// shading kernel (which keys the queries it emits) and the sort module (which
// keys training rows).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aqd {

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t cell10(float x) {
  float c = fminf(fmaxf(x, 0.0f), 1.0f) * 1024.0f;
  return min(uint32_t(c), 1023u);
}

__device__ __forceinline__ uint32_t morton30(float x, float y, float z) {
  return spread3(cell10(x)) | (spread3(cell10(y)) << 1) | (spread3(cell10(z)) << 2);
}

}  // namespace aqd
