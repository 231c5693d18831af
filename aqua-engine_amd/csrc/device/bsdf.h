// Metallic-roughness BSDF: Lambert diffuse + GGX microfacet specular.
//
//   alpha  = max(roughness^2, 1e-3)
//   F0     = mix(0.08 * specular, base_color, metallic)          (per channel)
//   F(c)   = F0 + (f90 - F0) (1 - c)^5,  f90 = min(1, 50 * mean(F0))
//            (Schlick 1994; the f90 clamp makes specular = 0 a pure
//             Lambertian surface, which is what every shipped BSDF uses:
//             scenes/cbox.json:17-19)
//   D      = alpha^2 / (pi ((n.h)^2 (alpha^2 - 1) + 1)^2)         (GGX, Walter 2007)
//   G2     = 1 / (1 + Lambda(wo) + Lambda(wi))                    (Smith, height-correlated)
//   f      = (1 - metallic)(1 - Fd(wo.h)) base / pi + F(wo.h) D G2 / (4 |n.wo| |n.wi|)
//            Fd = the dielectric Schlick term (F0 = 0.08 * specular)
// Sampling picks the specular lobe with probability ws / (ws + wd) from the
// Fresnel-weighted albedos at n.wo, GGX by its visible normals (Heitz 2018,
// "Sampling the GGX Distribution of Visible Normals"), diffuse by the cosine
// hemisphere; the pdf is the lobe mixture.
#pragma once

#include "common.h"

namespace aqd {

struct Mat {
  v3 base;
  float metallic;
  v3 emission;
  float roughness;
  float specular;
  float light_pdf;  // pi * lum(Le) / total light power (area measure)
  uint32_t tex;     // 0, or image index + 1 of a textured base colour
};

AQ_DEV Mat load_mat(const float4* mats, uint32_t m) {
  float4 a = mats[3 * m], b = mats[3 * m + 1], c = mats[3 * m + 2];
  Mat r;
  r.base = xyz(a);
  r.metallic = a.w;
  r.emission = xyz(b);
  r.roughness = b.w;
  r.specular = c.x;
  r.light_pdf = c.y;
  r.tex = __float_as_uint(c.z);
  return r;
}

struct Bsdf {
  v3 dcol;    // (1 - metallic) * base / pi
  v3 F0;
  float f90;
  float F0d, f90d;
  float a2, alpha;
  float dm;   // 1 - metallic
  float base_mean;
};

AQ_DEV Bsdf make_bsdf(const Mat& m) {
  Bsdf b;
  float alpha = m.roughness * m.roughness;
  if (alpha < 1e-3f) alpha = 1e-3f;
  b.alpha = alpha;
  b.a2 = alpha * alpha;
  b.dm = 1.0f - m.metallic;
  float s08 = 0.08f * m.specular;
  b.F0 = v3{s08 * b.dm + m.base.x * m.metallic, s08 * b.dm + m.base.y * m.metallic,
            s08 * b.dm + m.base.z * m.metallic};
  float f90 = 50.0f * (((b.F0.x + b.F0.y) + b.F0.z) * (1.0f / 3.0f));
  b.f90 = f90 > 1.0f ? 1.0f : f90;
  b.F0d = s08;
  float f90d = 50.0f * s08;
  b.f90d = f90d > 1.0f ? 1.0f : f90d;
  b.dcol = v3{(m.base.x * b.dm) * kInvPi, (m.base.y * b.dm) * kInvPi, (m.base.z * b.dm) * kInvPi};
  b.base_mean = ((m.base.x + m.base.y) + m.base.z) * (1.0f / 3.0f);
  return b;
}

AQ_DEV float schlick5(float c) {
  float m = 1.0f - c;
  float m2 = m * m;
  return (m2 * m2) * m;
}

AQ_DEV float ggx_d(const Bsdf& b, float ch) {
  float t = ((ch * ch) * (b.a2 - 1.0f)) + 1.0f;
  return b.a2 / ((kPi * t) * t);
}

AQ_DEV float smith_lambda(const Bsdf& b, float c) {
  float c2 = c * c;
  float tan2 = (1.0f - c2) / c2;
  return (__builtin_sqrtf(1.0f + b.a2 * tan2) - 1.0f) * 0.5f;
}

AQ_DEV float spec_prob(const Bsdf& b, float co) {
  float s5 = schlick5(co);
  float fx = b.F0.x + (b.f90 - b.F0.x) * s5;
  float fy = b.F0.y + (b.f90 - b.F0.y) * s5;
  float fz = b.F0.z + (b.f90 - b.F0.z) * s5;
  float ws = ((fx + fy) + fz) * (1.0f / 3.0f);
  float fd = b.F0d + (b.f90d - b.F0d) * s5;
  float wd = (b.dm * b.base_mean) * (1.0f - fd);
  float sum = ws + wd;
  return sum > 0.0f ? ws / sum : 0.0f;
}

// f(wo, wi) and the sampling pdf of wi; n is the shading normal on wo's side.
AQ_DEV v3 bsdf_eval(const Bsdf& b, v3 n, v3 wo, v3 wi, float* pdf, float* cos_i) {
  float co = dot(n, wo), ci = dot(n, wi);
  *cos_i = ci;
  if (co <= 0.0f || ci <= 0.0f) {
    *pdf = 0.0f;
    return v3{0.0f, 0.0f, 0.0f};
  }
  v3 h = normalize(add(wo, wi));
  float ch = dot(n, h);
  float doh = dot(wo, h);
  if (doh < 0.0f) doh = 0.0f;
  float s5 = schlick5(doh);
  v3 F = v3{b.F0.x + (b.f90 - b.F0.x) * s5, b.F0.y + (b.f90 - b.F0.y) * s5,
            b.F0.z + (b.f90 - b.F0.z) * s5};
  float Fd = b.F0d + (b.f90d - b.F0d) * s5;
  float D = ggx_d(b, ch);
  float lo = smith_lambda(b, co), li = smith_lambda(b, ci);
  float G2 = 1.0f / ((1.0f + lo) + li);
  float spec = (D * G2) / ((4.0f * co) * ci);
  float kd = 1.0f - Fd;
  v3 f = v3{b.dcol.x * kd + F.x * spec, b.dcol.y * kd + F.y * spec, b.dcol.z * kd + F.z * spec};
  float ps = spec_prob(b, co);
  float G1o = 1.0f / (1.0f + lo);
  float pdf_s = (G1o * D) / (4.0f * co);
  float pdf_d = ci * kInvPi;
  *pdf = ps * pdf_s + (1.0f - ps) * pdf_d;
  return f;
}

// Heitz 2018, listing 1, isotropic alpha; wo in the local frame (z = n).
AQ_DEV v3 sample_vndf(float alpha, v3 wo, float u1, float u2) {
  v3 vh = normalize(v3{alpha * wo.x, alpha * wo.y, wo.z});
  float lensq = vh.x * vh.x + vh.y * vh.y;
  v3 T1;
  if (lensq > 0.0f) {
    float inv = 1.0f / __builtin_sqrtf(lensq);
    T1 = v3{-vh.y * inv, vh.x * inv, 0.0f};
  } else {
    T1 = v3{1.0f, 0.0f, 0.0f};
  }
  v3 T2 = cross(vh, T1);
  float r = __builtin_sqrtf(u1);
  float s, c;
  sincos_2pi(u2, &s, &c);
  float t1 = r * c;
  float t2 = r * s;
  float sw = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - sw) * __builtin_sqrtf(1.0f - t1 * t1) + sw * t2;
  float w2 = (1.0f - t1 * t1) - t2 * t2;
  float w = __builtin_sqrtf(w2 > 0.0f ? w2 : 0.0f);
  v3 nh = v3{(t1 * T1.x + t2 * T2.x) + w * vh.x, (t1 * T1.y + t2 * T2.y) + w * vh.y,
             (t1 * T1.z + t2 * T2.z) + w * vh.z};
  return normalize(v3{alpha * nh.x, alpha * nh.y, nh.z > 0.0f ? nh.z : 0.0f});
}

// Samples wi; returns false when no direction results. On success *f and
// *pdf are bsdf_eval's values for the sampled direction.
AQ_DEV bool bsdf_sample(const Bsdf& b, v3 n, v3 wo, float ul, float u1, float u2, v3* wi, v3* f,
                        float* pdf, float* cos_i) {
  float co = dot(n, wo);
  if (co <= 0.0f) return false;
  v3 t, bt;
  onb(n, &t, &bt);
  float ps = spec_prob(b, co);
  v3 w;
  if (ul < ps) {
    v3 wol = v3{dot(wo, t), dot(wo, bt), co};
    v3 hl = sample_vndf(b.alpha, wol, u1, u2);
    v3 h = from_local(t, bt, n, hl);
    float d = dot(wo, h);
    w = sub(scl(h, 2.0f * d), wo);
  } else {
    float r = __builtin_sqrtf(u1);
    float s, c;
    sincos_2pi(u2, &s, &c);
    float z2 = 1.0f - u1;
    v3 l = v3{r * c, r * s, __builtin_sqrtf(z2 > 0.0f ? z2 : 0.0f)};
    w = from_local(t, bt, n, l);
  }
  *wi = w;
  *f = bsdf_eval(b, n, wo, w, pdf, cos_i);
  return *pdf > 0.0f && *cos_i > 0.0f;
}

}  // namespace aqd
