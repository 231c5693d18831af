// Image textures of base colours (room.json:6 `{"Image": ...}`): the hit's
// texcoords from its barycentrics, then a bilinear, repeating fetch of
// 8-bit sRGB texels decoded to linear through a 256-entry table. One fixed
// FP32 operation sequence (the CPU restatement runs the same), so textured
// films stay bit-exact.
#pragma once

#include "common.h"

namespace aqd {

struct DevTex {
  const uint32_t* texels;  // every image, RGBA8 (R in the low byte), row-major, row 0 = top
  const uint4* info;       // per image: (first texel, width, height, -)
  const float4* uv;        // 2 per triangle slot: (u0, v0, u1, v1), (u2, v2, -, -)
  const float* lut;        // sRGB code -> linear value
};

// Barycentrics (b1, b2) of point p in triangle (v0, v1, v2) from the normal
// equations of p - v0 = b1 e1 + b2 e2, clamped to [0, 1] (NaN -> 0).
AQ_DEV void bary_of(v3 v0, v3 v1, v3 v2, v3 p, float* b1, float* b2) {
  v3 e1 = sub(v1, v0), e2 = sub(v2, v0), w = sub(p, v0);
  float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
  float d20 = dot(w, e1), d21 = dot(w, e2);
  float den = d00 * d11 - d01 * d01;
  float x = 0.0f, y = 0.0f;
  if (den != 0.0f) {
    x = (d11 * d20 - d01 * d21) / den;
    y = (d00 * d21 - d01 * d20) / den;
  }
  *b1 = x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f;
  *b2 = y > 0.0f ? (y < 1.0f ? y : 1.0f) : 0.0f;
}

AQ_DEV float tex_channel(const float* lut, uint32_t a, uint32_t b, uint32_t c, uint32_t d, int k, float fx,
                         float fy) {
  const float la = lut[(a >> (8 * k)) & 0xFFu], lb = lut[(b >> (8 * k)) & 0xFFu];
  const float lc = lut[(c >> (8 * k)) & 0xFFu], ld = lut[(d >> (8 * k)) & 0xFFu];
  const float top = la * (1.0f - fx) + lb * fx;
  const float bot = lc * (1.0f - fx) + ld * fx;
  return top * (1.0f - fy) + bot * fy;
}

// Linear RGB of image `img` at the texcoords of triangle slot `slot` at
// barycentrics (b1, b2). Texel centres sit at half-integers; v = 0 is the
// bottom row (the .mesh texcoords follow the OBJ convention).
AQ_DEV v3 tex_base(const DevTex& T, uint32_t slot, uint32_t img, float b1, float b2) {
  const float4 a = T.uv[2 * size_t(slot)], b = T.uv[2 * size_t(slot) + 1];
  const float b0 = (1.0f - b1) - b2;
  float u = (a.x * b0 + a.z * b1) + b.x * b2;
  float v = (a.y * b0 + a.w * b1) + b.y * b2;
  u = u - floorf(u);
  v = v - floorf(v);
  const uint4 in = T.info[img];
  const int w = int(in.y), h = int(in.z);
  const float x = u * float(w) - 0.5f;
  const float y = (1.0f - v) * float(h) - 0.5f;
  const float x0f = floorf(x), y0f = floorf(y);
  const float fx = x - x0f, fy = y - y0f;
  const int x0 = int(x0f), y0 = int(y0f);  // in [-1, w - 1] and [-1, h - 1]
  const int xa = x0 < 0 ? w - 1 : x0, xb = x0 + 1 >= w ? 0 : x0 + 1;
  const int ya = y0 < 0 ? h - 1 : y0, yb = y0 + 1 >= h ? 0 : y0 + 1;
  const uint32_t* t = T.texels + in.x;
  const uint32_t p00 = t[size_t(ya) * w + xa], p10 = t[size_t(ya) * w + xb];
  const uint32_t p01 = t[size_t(yb) * w + xa], p11 = t[size_t(yb) * w + xb];
  return v3{tex_channel(T.lut, p00, p10, p01, p11, 0, fx, fy), tex_channel(T.lut, p00, p10, p01, p11, 1, fx, fy),
            tex_channel(T.lut, p00, p10, p01, p11, 2, fx, fy)};
}

}  // namespace aqd
