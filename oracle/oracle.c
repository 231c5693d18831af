/*
 * oracle.c -- CPU restatement of the path-tracing hot path (test
 * infrastructure; see oracle.h for what pins each part).
 *
 * Reference anchors: the inputs are the reference's formats
 * (scenes/cbox.json:2-627 materials/camera/lights/shapes, scenes/<name>.mesh
 * geometry, scenes/integrator.json:3,5 spp and max_depth); the reference
 * crate has no renderer source (src/lib.rs is empty), so every algorithm
 * below cites the published method it restates.
 *
 * Compiled with -ffp-contract=off and without fast-math: every FP32
 * expression rounds exactly as written, fmaf() is the fused operation.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* vector helpers                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  float x, y, z;
} V;

static inline V v(float x, float y, float z) {
  V r = {x, y, z};
  return r;
}
static inline V vadd(V a, V b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V vsub(V a, V b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V vmul(V a, V b) { return v(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V vscl(V a, float s) { return v(a.x * s, a.y * s, a.z * s); }
static inline V vneg(V a) { return v(-a.x, -a.y, -a.z); }
static inline float vdot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline V vcross(V a, V b) {
  return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline V vnorm(V a) {
  float inv = 1.0f / sqrtf(vdot(a, a));
  return vscl(a, inv);
}
static inline float mn(float a, float b) { return a < b ? a : b; }
static inline float mx(float a, float b) { return a > b ? a : b; }
static inline float comp(V a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static inline uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static inline float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static const float PI_F = 3.14159265358979323846f;
static const float INV_PI_F = 0.31830988618379067154f;
static const float FLT_MAX_F = 3.402823466e+38f;

/* ------------------------------------------------------------------------ */
/* PCG32: O'Neill 2014, "PCG: A Family of Simple Fast Space-Efficient         */
/* Statistically Good Algorithms for Random Number Generation", XSH-RR 64/32. */
/* ------------------------------------------------------------------------ */

#define PCG_MULT 6364136223846793005ULL

typedef struct {
  uint64_t state, inc;
} Pcg;

static inline uint32_t pcg_next(Pcg* r) {
  uint64_t old = r->state;
  r->state = old * PCG_MULT + r->inc;
  uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

static Pcg pcg_seed(uint64_t initstate, uint64_t initseq) {
  Pcg r;
  r.state = 0u;
  r.inc = (initseq << 1u) | 1u;
  pcg_next(&r);
  r.state += initstate;
  pcg_next(&r);
  return r;
}

/* Arbitrary-stride LCG jump: F. Brown, "Random Number Generation with
 * Arbitrary Strides", Trans. Am. Nucl. Soc. 1994. */
static void pcg_advance(Pcg* r, uint64_t delta) {
  uint64_t cur_mult = PCG_MULT, cur_plus = r->inc;
  uint64_t acc_mult = 1u, acc_plus = 0u;
  while (delta > 0) {
    if (delta & 1u) {
      acc_mult *= cur_mult;
      acc_plus = acc_plus * cur_mult + cur_plus;
    }
    cur_plus = (cur_mult + 1u) * cur_plus;
    cur_mult *= cur_mult;
    delta /= 2u;
  }
  r->state = acc_mult * r->state + acc_plus;
}

static inline float unit_float(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

void orc_pcg32_seq(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t* out) {
  Pcg r = pcg_seed(initstate, initseq);
  pcg_advance(&r, advance);
  for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(&r);
}

/* ------------------------------------------------------------------------ */
/* Stratified sampler (integrator "sampler": "stratified"): the camera ray's  */
/* two dimensions and the first hit's six come from shuffled, Owen-scrambled  */
/* 2-D Sobol points over the sample index (Burley 2020, "Practical Hash-based */
/* Owen Scrambling", JCGT 9(4): nested uniform scramble by the Laine-Karras   */
/* hash on reversed bits, the index shuffled by the same scramble); the rest  */
/* stay on the PCG32 stream. Every power-of-two prefix of a pixel's samples   */
/* is then a (0, m, 2)-net in each pattern.                                   */
/* ------------------------------------------------------------------------ */

static inline uint32_t rev32(uint32_t x) {
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
}

/* Laine & Karras 2011 hash as given by Burley 2020 (listing 2). */
static inline uint32_t lk_hash(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}

/* Nested uniform (Owen) scramble of the bits of x, most significant first. */
static inline uint32_t owen32(uint32_t x, uint32_t seed) { return rev32(lk_hash(rev32(x), seed)); }

/* 32-bit integer mix ("lowbias32", C. Wellons, hash-prospector, public domain). */
static inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* Sobol's second dimension (Sobol 1967; every direction number m_k = 1): the
 * k-th column is the previous one xor itself shifted right by one. The first
 * dimension is the bit reversal of the index (van der Corput). */
static inline uint32_t sobol_dim1(uint32_t i) {
  uint32_t r = 0u, v = 0x80000000u;
  for (; i; i >>= 1, v ^= v >> 1)
    if (i & 1u) r ^= v;
  return r;
}

/* Point `index` of pattern `pair` of pixel `pixel` in frame `frame`. */
static void strat2(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t pair, uint32_t index, float* a,
                   float* b) {
  const uint32_t h = mix32(mix32(mix32(mix32(seed) ^ pixel) ^ frame) ^ pair);
  const uint32_t i = owen32(index, h);
  *a = unit_float(owen32(rev32(i), mix32(h ^ 0xa511e9b3u)));
  *b = unit_float(owen32(sobol_dim1(i), mix32(h ^ 0x63d83595u)));
}

void orc_strat2_seq(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t pair, uint32_t first, uint32_t n,
                    float* out) {
  for (uint32_t k = 0; k < n; ++k) strat2(seed, frame, pixel, pair, first + k, &out[2 * k], &out[2 * k + 1]);
}

/* ------------------------------------------------------------------------ */
/* sin / cos of 2 pi u: exact quadrant split of 4u, fold to [0, 1/2] with the */
/* exact complement, Taylor series of sin(pi/2 g) and cos(pi/2 g) in Horner   */
/* form (coefficients (-1)^k (pi/2)^n / n!).                                  */
/* ------------------------------------------------------------------------ */

static void sincos_2pi(float u, float* s, float* c) {
  float x = u * 4.0f;
  int q = (int)x;
  float f = x - (float)q;
  int comp_ = f > 0.5f;
  float g = comp_ ? 1.0f - f : f;
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
  float s0 = comp_ ? cg : sg;
  float c0 = comp_ ? sg : cg;
  switch (q & 3) {
    case 0: *s = s0; *c = c0; break;
    case 1: *s = c0; *c = -s0; break;
    case 2: *s = -s0; *c = -c0; break;
    default: *s = -c0; *c = s0; break;
  }
}

void orc_sincos_2pi(uint32_t n, const float* u, float* s, float* c) {
  for (uint32_t i = 0; i < n; ++i) sincos_2pi(u[i], &s[i], &c[i]);
}

/* ------------------------------------------------------------------------ */
/* scene + the oracle's own BVH (binned SAH, 12 bins, leaves <= 2; node =    */
/* box + (first child | first triangle) + count; boxes tested when popped)   */
/* ------------------------------------------------------------------------ */

typedef struct {
  float lo[3], hi[3];
  int32_t index;  /* inner: left child (right = index + 1); leaf: first ref */
  int32_t count;  /* 0 = inner */
} ONode;

typedef struct {
  float base[3], metallic, emission[3], roughness, specular, light_pdf;
  uint32_t tex; /* 0, or image index + 1 of a textured base colour */
} OMat;

struct orc_scene {
  uint32_t n_tris, n_mats, n_point;
  float* verts;      /* 9 per triangle */
  uint32_t* mat_ids;
  OMat* mats;
  float* points;     /* 6 per point light */
  ONode* nodes;
  uint32_t n_nodes, max_depth;
  uint32_t* refs;    /* leaf triangle references */
  float delta;
  /* light table */
  uint32_t n_area, n_lights;
  uint32_t* light_tri;
  float* light_area;
  float* light_pmf;
  float* light_cdf;
  /* image textures */
  float* uvs;        /* 6 per triangle (load order) */
  uint32_t n_images;
  uint32_t* img;     /* 3 per image: first texel, width, height */
  uint8_t* texels;   /* RGBA8, row 0 = top */
  float lut[256];    /* sRGB code -> linear */
};

typedef struct {
  float lo[3], hi[3];
} Box;

static void box_reset(Box* b) {
  for (int k = 0; k < 3; ++k) {
    b->lo[k] = INFINITY;
    b->hi[k] = -INFINITY;
  }
}
static void box_grow_pt(Box* b, const float* p) {
  for (int k = 0; k < 3; ++k) {
    if (p[k] < b->lo[k]) b->lo[k] = p[k];
    if (p[k] > b->hi[k]) b->hi[k] = p[k];
  }
}
static void box_grow(Box* b, const Box* o) {
  box_grow_pt(b, o->lo);
  box_grow_pt(b, o->hi);
}
static double box_area(const Box* b) {
  if (!(b->hi[0] >= b->lo[0])) return 0.0;
  double dx = (double)b->hi[0] - b->lo[0], dy = (double)b->hi[1] - b->lo[1], dz = (double)b->hi[2] - b->lo[2];
  return dx * dy + dy * dz + dz * dx;
}

typedef struct {
  orc_scene* s;
  Box* tbox;
  float* cen;
  uint32_t cap;
} Build;

static uint32_t build_rec(Build* B, uint32_t node, uint32_t first, uint32_t count, uint32_t depth) {
  orc_scene* s = B->s;
  Box b, cb;
  box_reset(&b);
  box_reset(&cb);
  for (uint32_t i = first; i < first + count; ++i) {
    box_grow(&b, &B->tbox[s->refs[i]]);
    box_grow_pt(&cb, &B->cen[3 * s->refs[i]]);
  }
  ONode* nd = &s->nodes[node];
  memcpy(nd->lo, b.lo, sizeof b.lo);
  memcpy(nd->hi, b.hi, sizeof b.hi);
  if (depth > s->max_depth) s->max_depth = depth;
  if (count <= 2) {
    nd->index = (int32_t)first;
    nd->count = (int32_t)count;
    return node;
  }
  enum { NB = 12 };
  double best = INFINITY;
  int best_axis = -1, best_bin = -1;
  for (int a = 0; a < 3; ++a) {
    float ext = cb.hi[a] - cb.lo[a];
    if (!(ext > 0.0f)) continue;
    Box bins[NB];
    uint32_t cnt[NB] = {0};
    for (int k = 0; k < NB; ++k) box_reset(&bins[k]);
    for (uint32_t i = first; i < first + count; ++i) {
      uint32_t t = s->refs[i];
      int k = (int)((B->cen[3 * t + a] - cb.lo[a]) / ext * NB);
      if (k >= NB) k = NB - 1;
      if (k < 0) k = 0;
      box_grow(&bins[k], &B->tbox[t]);
      cnt[k]++;
    }
    for (int split = 1; split < NB; ++split) {
      Box l, r;
      box_reset(&l);
      box_reset(&r);
      uint32_t nl = 0, nr = 0;
      for (int k = 0; k < split; ++k) {
        if (cnt[k]) box_grow(&l, &bins[k]);
        nl += cnt[k];
      }
      for (int k = split; k < NB; ++k) {
        if (cnt[k]) box_grow(&r, &bins[k]);
        nr += cnt[k];
      }
      if (!nl || !nr) continue;
      double c = box_area(&l) * nl + box_area(&r) * nr;
      if (c < best) {
        best = c;
        best_axis = a;
        best_bin = split;
      }
    }
  }
  uint32_t mid;
  if (best_axis < 0) {
    mid = first + count / 2;
  } else {
    float ext = cb.hi[best_axis] - cb.lo[best_axis];
    uint32_t i = first, j = first + count;
    while (i < j) {
      uint32_t t = s->refs[i];
      int k = (int)((B->cen[3 * t + best_axis] - cb.lo[best_axis]) / ext * NB);
      if (k >= NB) k = NB - 1;
      if (k < best_bin) {
        ++i;
      } else {
        --j;
        uint32_t tmp = s->refs[i];
        s->refs[i] = s->refs[j];
        s->refs[j] = tmp;
      }
    }
    mid = i;
    if (mid == first || mid == first + count) mid = first + count / 2;
  }
  uint32_t left = s->n_nodes;
  s->n_nodes += 2;
  nd = &s->nodes[node];
  nd->index = (int32_t)left;
  nd->count = 0;
  build_rec(B, left, first, mid - first, depth + 1);
  build_rec(B, left + 1, mid, first + count - mid, depth + 1);
  return node;
}

static double lum_d(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

/* Lights are picked in proportion to emitted power: pi * area * lum(Le) for
 * a one-sided Lambertian emitter triangle, 4 pi * lum(I) for a point light
 * (area lights first, in triangle order, then point lights). */
static int build_lights(orc_scene* s) {
  uint32_t cap = s->n_tris + s->n_point + 1;
  double* power = (double*)malloc(sizeof(double) * cap);
  s->light_tri = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  s->light_area = (float*)malloc(sizeof(float) * cap);
  s->light_pmf = (float*)malloc(sizeof(float) * cap);
  s->light_cdf = (float*)malloc(sizeof(float) * cap);
  if (!power || !s->light_tri || !s->light_area || !s->light_pmf || !s->light_cdf) {
    free(power);
    return -1;
  }
  uint32_t n = 0;
  for (uint32_t t = 0; t < s->n_tris; ++t) {
    const OMat* m = &s->mats[s->mat_ids[t]];
    if (!(m->emission[0] > 0.0f || m->emission[1] > 0.0f || m->emission[2] > 0.0f)) continue;
    double lum = lum_d(m->emission[0], m->emission[1], m->emission[2]);
    const float* p = &s->verts[9 * (size_t)t];
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)p[3 + k] - (double)p[k];
      e2[k] = (double)p[6 + k] - (double)p[k];
    }
    double c0 = e1[1] * e2[2] - e1[2] * e2[1];
    double c1 = e1[2] * e2[0] - e1[0] * e2[2];
    double c2 = e1[0] * e2[1] - e1[1] * e2[0];
    double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (!(area > 0.0) || !(lum > 0.0)) continue;
    s->light_tri[n] = t;
    s->light_area[n] = (float)area;
    power[n] = M_PI * area * lum;
    ++n;
  }
  s->n_area = n;
  for (uint32_t j = 0; j < s->n_point; ++j) {
    const float* pl = &s->points[6 * (size_t)j];
    double lum = lum_d(pl[3], pl[4], pl[5]);
    power[n++] = 4.0 * M_PI * (lum > 0.0 ? lum : 0.0);
  }
  double total = 0.0;
  for (uint32_t i = 0; i < n; ++i) total += power[i];
  s->n_lights = total > 0.0 ? n : 0;
  double acc = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    s->light_pmf[i] = total > 0.0 ? (float)(power[i] / total) : 0.0f;
    acc += power[i];
    s->light_cdf[i] = total > 0.0 ? (i + 1 == n ? 1.0f : (float)(acc / total)) : 0.0f;
  }
  for (uint32_t m = 0; m < s->n_mats; ++m) {
    OMat* mt = &s->mats[m];
    double lum = lum_d(mt->emission[0], mt->emission[1], mt->emission[2]);
    mt->light_pdf = (total > 0.0 && lum > 0.0) ? (float)(M_PI * lum / total) : 0.0f;
  }
  free(power);
  return 0;
}

int orc_scene_create(uint32_t n_tris, const float* verts, const uint32_t* mat_ids, uint32_t n_mats,
                     const float* mats, uint32_t n_point, const float* points, orc_scene** out) {
  orc_scene* s = (orc_scene*)calloc(1, sizeof *s);
  if (!s) return -1;
  s->n_tris = n_tris;
  s->n_mats = n_mats;
  s->n_point = n_point;
  s->verts = (float*)malloc(sizeof(float) * 9 * (size_t)n_tris + 4);
  s->mat_ids = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n_tris + 4);
  s->mats = (OMat*)calloc(n_mats + 1, sizeof(OMat));
  s->points = (float*)malloc(sizeof(float) * 6 * (size_t)n_point + 4);
  s->refs = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n_tris + 4);
  s->nodes = (ONode*)calloc(2 * (size_t)n_tris + 2, sizeof(ONode));
  if (!s->verts || !s->mat_ids || !s->mats || !s->points || !s->refs || !s->nodes) {
    orc_scene_destroy(s);
    return -1;
  }
  memcpy(s->verts, verts, sizeof(float) * 9 * (size_t)n_tris);
  memcpy(s->mat_ids, mat_ids, sizeof(uint32_t) * n_tris);
  if (n_point) memcpy(s->points, points, sizeof(float) * 6 * (size_t)n_point);
  for (uint32_t m = 0; m < n_mats; ++m) {
    const float* a = &mats[12 * (size_t)m];
    OMat* o = &s->mats[m];
    for (int k = 0; k < 3; ++k) {
      o->base[k] = a[k];
      o->emission[k] = a[4 + k];
    }
    o->metallic = a[3];
    o->roughness = a[7];
    o->specular = a[8];
    memcpy(&o->tex, &a[9], 4);
  }
  /* Culling slack: 2^-16 of the largest coordinate magnitude. */
  float mxc = 0.0f;
  for (size_t i = 0; i < 9 * (size_t)n_tris; ++i) {
    float a = fabsf(verts[i]);
    if (a > mxc) mxc = a;
  }
  s->delta = ldexpf(mxc > 0.0f ? mxc : 1.0f, -16);
  /* BVH */
  Build B;
  B.s = s;
  B.tbox = (Box*)malloc(sizeof(Box) * ((size_t)n_tris + 1));
  B.cen = (float*)malloc(sizeof(float) * 3 * ((size_t)n_tris + 1));
  if (!B.tbox || !B.cen) {
    free(B.tbox);
    free(B.cen);
    orc_scene_destroy(s);
    return -1;
  }
  for (uint32_t t = 0; t < n_tris; ++t) {
    box_reset(&B.tbox[t]);
    for (int k = 0; k < 3; ++k) box_grow_pt(&B.tbox[t], &verts[9 * (size_t)t + 3 * k]);
    for (int k = 0; k < 3; ++k) B.cen[3 * t + k] = 0.5f * (B.tbox[t].lo[k] + B.tbox[t].hi[k]);
    s->refs[t] = t;
  }
  s->n_nodes = 1;
  if (n_tris > 0) {
    build_rec(&B, 0, 0, n_tris, 0);
  } else {
    s->nodes[0].count = 0;
    s->nodes[0].index = -1;
  }
  free(B.tbox);
  free(B.cen);
  if (build_lights(s) != 0) {
    orc_scene_destroy(s);
    return -1;
  }
  *out = s;
  return 0;
}

void orc_scene_destroy(orc_scene* s) {
  if (!s) return;
  free(s->verts);
  free(s->mat_ids);
  free(s->mats);
  free(s->points);
  free(s->refs);
  free(s->nodes);
  free(s->light_tri);
  free(s->light_area);
  free(s->light_pmf);
  free(s->light_cdf);
  free(s->uvs);
  free(s->img);
  free(s->texels);
  free(s);
}

int orc_scene_set_textures(orc_scene* s, const float* uvs, uint32_t n_images, const uint32_t* dims,
                           const uint8_t* texels) {
  size_t total = 0;
  for (uint32_t k = 0; k < n_images; ++k) total += (size_t)dims[2 * k] * dims[2 * k + 1];
  s->uvs = (float*)malloc(sizeof(float) * 6 * (size_t)s->n_tris + 4);
  s->img = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)n_images + 4);
  s->texels = (uint8_t*)malloc(4 * total + 4);
  if (!s->uvs || !s->img || !s->texels) return -1;
  memcpy(s->uvs, uvs, sizeof(float) * 6 * (size_t)s->n_tris);
  memcpy(s->texels, texels, 4 * total);
  size_t at = 0;
  for (uint32_t k = 0; k < n_images; ++k) {
    s->img[3 * k] = (uint32_t)at;
    s->img[3 * k + 1] = dims[2 * k];
    s->img[3 * k + 2] = dims[2 * k + 1];
    at += (size_t)dims[2 * k] * dims[2 * k + 1];
  }
  s->n_images = n_images;
  /* IEC 61966-2-1 decode of each 8-bit code, in double, rounded once */
  for (int i = 0; i < 256; ++i) {
    double c = i / 255.0;
    s->lut[i] = (float)(c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4));
  }
  return 0;
}

void orc_scene_bvh_info(const orc_scene* s, uint32_t* n_nodes, uint32_t* max_depth) {
  if (n_nodes) *n_nodes = s->n_nodes;
  if (max_depth) *max_depth = s->max_depth;
}

/* ------------------------------------------------------------------------ */
/* ray / triangle and ray / box                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
  V o, d, inv, oi;
  int kx, ky, kz;
  float sx, sy, sz;
} Ray;

static inline float safe_inv(float d) {
  float dd = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
  return 1.0f / dd;
}

/* Woop, Benthin & Wald 2013, "Watertight Ray/Triangle Intersection",
 * JCGT 2(1): per-ray shear to the dominant axis, then 2D edge functions.
 * The paper's double-precision re-test of exactly-zero edge functions is
 * not used: zero counts as either sign, which keeps the test watertight. */
static Ray ray_make(V o, V d) {
  Ray r;
  r.o = o;
  r.d = d;
  r.inv = v(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  r.oi = vmul(o, r.inv);
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  int kx = (kz + 1) % 3;
  int ky = (kx + 1) % 3;
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

static int tri_isect(const Ray* r, const float* p, float* t_out) {
  V A = vsub(v(p[0], p[1], p[2]), r->o);
  V B = vsub(v(p[3], p[4], p[5]), r->o);
  V C = vsub(v(p[6], p[7], p[8]), r->o);
  float Akz = comp(A, r->kz), Bkz = comp(B, r->kz), Ckz = comp(C, r->kz);
  float Ax = comp(A, r->kx) - r->sx * Akz;
  float Ay = comp(A, r->ky) - r->sy * Akz;
  float Bx = comp(B, r->kx) - r->sx * Bkz;
  float By = comp(B, r->ky) - r->sy * Bkz;
  float Cx = comp(C, r->kx) - r->sx * Ckz;
  float Cy = comp(C, r->ky) - r->sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float Vv = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return 0;
  float det = (U + Vv) + W;
  if (det == 0.0f) return 0;
  float Az = r->sz * Akz, Bz = r->sz * Bkz, Cz = r->sz * Ckz;
  float T = fmaf(U, Az, fmaf(Vv, Bz, W * Cz));
  *t_out = T / det;
  return 1;
}

/* Ize 2013, "Robust BVH Ray Traversal", JCGT 2(2): far slab distances
 * scaled by 1 + 2 gamma_3; plus an absolute slack `delta` so no box whose
 * triangle the watertight test would accept at t <= tbest is culled. */
static const float FAR_SCALE = 1.00000036f;

static inline int box_test(const Ray* r, const float* lo, const float* hi, float tbest, float delta) {
  float t0x = fmaf(lo[0], r->inv.x, -r->oi.x), t1x = fmaf(hi[0], r->inv.x, -r->oi.x);
  float t0y = fmaf(lo[1], r->inv.y, -r->oi.y), t1y = fmaf(hi[1], r->inv.y, -r->oi.y);
  float t0z = fmaf(lo[2], r->inv.z, -r->oi.z), t1z = fmaf(hi[2], r->inv.z, -r->oi.z);
  float tn = mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mx(mn(t0z, t1z), 0.0f));
  float tf = mn(mn(mx(t0x, t1x), mx(t0y, t1y)), mn(mx(t0z, t1z), tbest));
  return tn <= tf * FAR_SCALE + delta;
}

/* Entry distance only, for ordering children (never for culling). */
static inline float box_entry(const Ray* r, const float* lo, const float* hi) {
  float t0x = fmaf(lo[0], r->inv.x, -r->oi.x), t1x = fmaf(hi[0], r->inv.x, -r->oi.x);
  float t0y = fmaf(lo[1], r->inv.y, -r->oi.y), t1y = fmaf(hi[1], r->inv.y, -r->oi.y);
  float t0z = fmaf(lo[2], r->inv.z, -r->oi.z), t1z = fmaf(hi[2], r->inv.z, -r->oi.z);
  return mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mx(mn(t0z, t1z), 0.0f));
}

typedef struct {
  uint64_t nodes, tris;
} TCount;

/* Closest hit: least (t, triangle id) with 0 < t < tmax. */
static uint32_t closest(const orc_scene* s, V o, V d, float tmax, int brute, float* t_out, TCount* tc) {
  Ray r = ray_make(o, d);
  float tbest = tmax;
  uint32_t best = 0xFFFFFFFFu;
  if (brute) {
    for (uint32_t id = 0; id < s->n_tris; ++id) {
      float t;
      if (tri_isect(&r, &s->verts[9 * (size_t)id], &t) && t > 0.0f && t < tmax &&
          (t < tbest || (t == tbest && id < best))) {
        tbest = t;
        best = id;
      }
    }
    *t_out = tbest;
    return best;
  }
  /* Depth-first, nearer child first; every box is (re)tested with the
   * current best distance when it is popped. */
  uint32_t stack[128];
  int sp = 0;
  if (s->n_tris) stack[sp++] = 0;
  while (sp > 0) {
    const ONode* nd = &s->nodes[stack[--sp]];
    if (tc) tc->nodes++;
    if (!box_test(&r, nd->lo, nd->hi, tbest, s->delta)) continue;
    if (nd->count > 0) {
      for (int32_t i = 0; i < nd->count; ++i) {
        uint32_t id = s->refs[nd->index + i];
        float t;
        if (tc) tc->tris++;
        if (tri_isect(&r, &s->verts[9 * (size_t)id], &t) && t > 0.0f && t < tmax &&
            (t < tbest || (t == tbest && id < best))) {
          tbest = t;
          best = id;
        }
      }
    } else {
      const ONode* l = &s->nodes[nd->index];
      const ONode* rr = &s->nodes[nd->index + 1];
      float tl = box_entry(&r, l->lo, l->hi);
      float tr = box_entry(&r, rr->lo, rr->hi);
      if (tl <= tr) {
        stack[sp++] = (uint32_t)nd->index + 1;
        stack[sp++] = (uint32_t)nd->index;
      } else {
        stack[sp++] = (uint32_t)nd->index;
        stack[sp++] = (uint32_t)nd->index + 1;
      }
    }
  }
  *t_out = tbest;
  return best;
}

static int occluded(const orc_scene* s, V o, V d, float tmax, int brute) {
  Ray r = ray_make(o, d);
  if (brute) {
    for (uint32_t id = 0; id < s->n_tris; ++id) {
      float t;
      if (tri_isect(&r, &s->verts[9 * (size_t)id], &t) && t > 0.0f && t < tmax) return 1;
    }
    return 0;
  }
  uint32_t stack[128];
  int sp = 0;
  if (s->n_tris) stack[sp++] = 0;
  while (sp > 0) {
    const ONode* nd = &s->nodes[stack[--sp]];
    if (!box_test(&r, nd->lo, nd->hi, tmax, s->delta)) continue;
    if (nd->count > 0) {
      for (int32_t i = 0; i < nd->count; ++i) {
        float t;
        if (tri_isect(&r, &s->verts[9 * (size_t)s->refs[nd->index + i]], &t) && t > 0.0f && t < tmax)
          return 1;
      }
    } else {
      stack[sp++] = (uint32_t)nd->index + 1;
      stack[sp++] = (uint32_t)nd->index;
    }
  }
  return 0;
}

int orc_trace_closest(const orc_scene* s, uint32_t n, const float* rays, int brute, uint32_t* ids, float* t) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* x = &rays[8 * (size_t)i];
    float tt;
    ids[i] = closest(s, v(x[0], x[1], x[2]), v(x[4], x[5], x[6]), x[3], brute, &tt, NULL);
    if (t) t[i] = tt;
  }
  return 0;
}

int orc_trace_occluded(const orc_scene* s, uint32_t n, const float* rays, int brute, uint8_t* occ) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* x = &rays[8 * (size_t)i];
    occ[i] = (uint8_t)occluded(s, v(x[0], x[1], x[2]), v(x[4], x[5], x[6]), x[3], brute);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* frames, offsets                                                           */
/* ------------------------------------------------------------------------ */

/* Duff, Burgess, Christensen, Hery, Kensler, Liani, Villemin 2017,
 * "Building an Orthonormal Basis, Revisited", JCGT 6(1). */
static void onb(V n, V* t, V* b) {
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = (n.x * n.y) * a;
  *t = v(1.0f + ((sign * n.x) * n.x) * a, sign * bb, -sign * n.x);
  *b = v(bb, sign + (n.y * n.y) * a, -n.y);
}
static V from_local(V t, V b, V n, V l) {
  return v((t.x * l.x + b.x * l.y) + n.x * l.z, (t.y * l.x + b.y * l.y) + n.y * l.z,
           (t.z * l.x + b.z * l.y) + n.z * l.z);
}

/* Waechter & Binder 2019, "A Fast and Robust Method for Avoiding
 * Self-Intersection", Ray Tracing Gems ch. 6. */
static float off1(float p, float n) {
  const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
  int o = (int)(int_scale * n);
  uint32_t bits = f2u(p) + (uint32_t)(p < 0.0f ? -o : o);
  return fabsf(p) < origin ? p + float_scale * n : u2f(bits);
}
static V offset_ray(V p, V n) { return v(off1(p.x, n.x), off1(p.y, n.y), off1(p.z, n.z)); }

/* ------------------------------------------------------------------------ */
/* metallic-roughness BSDF: Lambert + GGX (Walter, Marschner, Li, Torrance   */
/* 2007), Smith height-correlated G2 (Heitz 2014), Schlick 1994 Fresnel with */
/* f90 = min(1, 50 mean(F0)), VNDF sampling (Heitz 2018, JCGT 7(4) list. 1). */
/* ------------------------------------------------------------------------ */

typedef struct {
  V dcol, F0;
  float f90, F0d, f90d, a2, alpha, dm, base_mean;
} Bsdf;

static Bsdf bsdf_make(const OMat* m) {
  Bsdf b;
  float alpha = m->roughness * m->roughness;
  if (alpha < 1e-3f) alpha = 1e-3f;
  b.alpha = alpha;
  b.a2 = alpha * alpha;
  b.dm = 1.0f - m->metallic;
  float s08 = 0.08f * m->specular;
  b.F0 = v(s08 * b.dm + m->base[0] * m->metallic, s08 * b.dm + m->base[1] * m->metallic,
           s08 * b.dm + m->base[2] * m->metallic);
  float f90 = 50.0f * (((b.F0.x + b.F0.y) + b.F0.z) * (1.0f / 3.0f));
  b.f90 = f90 > 1.0f ? 1.0f : f90;
  b.F0d = s08;
  float f90d = 50.0f * s08;
  b.f90d = f90d > 1.0f ? 1.0f : f90d;
  b.dcol = v((m->base[0] * b.dm) * INV_PI_F, (m->base[1] * b.dm) * INV_PI_F, (m->base[2] * b.dm) * INV_PI_F);
  b.base_mean = ((m->base[0] + m->base[1]) + m->base[2]) * (1.0f / 3.0f);
  return b;
}

static float schlick5(float c) {
  float m = 1.0f - c;
  float m2 = m * m;
  return (m2 * m2) * m;
}

static float ggx_d(const Bsdf* b, float ch) {
  float t = ((ch * ch) * (b->a2 - 1.0f)) + 1.0f;
  return b->a2 / ((PI_F * t) * t);
}

static float smith_lambda(const Bsdf* b, float c) {
  float c2 = c * c;
  float tan2 = (1.0f - c2) / c2;
  return (sqrtf(1.0f + b->a2 * tan2) - 1.0f) * 0.5f;
}

static float spec_prob(const Bsdf* b, float co) {
  float s5 = schlick5(co);
  float fx = b->F0.x + (b->f90 - b->F0.x) * s5;
  float fy = b->F0.y + (b->f90 - b->F0.y) * s5;
  float fz = b->F0.z + (b->f90 - b->F0.z) * s5;
  float ws = ((fx + fy) + fz) * (1.0f / 3.0f);
  float fd = b->F0d + (b->f90d - b->F0d) * s5;
  float wd = (b->dm * b->base_mean) * (1.0f - fd);
  float sum = ws + wd;
  return sum > 0.0f ? ws / sum : 0.0f;
}

static V bsdf_eval(const Bsdf* b, V n, V wo, V wi, float* pdf, float* cos_i) {
  float co = vdot(n, wo), ci = vdot(n, wi);
  *cos_i = ci;
  if (co <= 0.0f || ci <= 0.0f) {
    *pdf = 0.0f;
    return v(0.0f, 0.0f, 0.0f);
  }
  V h = vnorm(vadd(wo, wi));
  float ch = vdot(n, h);
  float doh = vdot(wo, h);
  if (doh < 0.0f) doh = 0.0f;
  float s5 = schlick5(doh);
  V F = v(b->F0.x + (b->f90 - b->F0.x) * s5, b->F0.y + (b->f90 - b->F0.y) * s5,
          b->F0.z + (b->f90 - b->F0.z) * s5);
  float Fd = b->F0d + (b->f90d - b->F0d) * s5;
  float D = ggx_d(b, ch);
  float lo = smith_lambda(b, co), li = smith_lambda(b, ci);
  float G2 = 1.0f / ((1.0f + lo) + li);
  float spec = (D * G2) / ((4.0f * co) * ci);
  float kd = 1.0f - Fd;
  V f = v(b->dcol.x * kd + F.x * spec, b->dcol.y * kd + F.y * spec, b->dcol.z * kd + F.z * spec);
  float ps = spec_prob(b, co);
  float G1o = 1.0f / (1.0f + lo);
  float pdf_s = (G1o * D) / (4.0f * co);
  float pdf_d = ci * INV_PI_F;
  *pdf = ps * pdf_s + (1.0f - ps) * pdf_d;
  return f;
}

static V sample_vndf(float alpha, V wo, float u1, float u2) {
  V vh = vnorm(v(alpha * wo.x, alpha * wo.y, wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V T1;
  if (lensq > 0.0f) {
    float inv = 1.0f / sqrtf(lensq);
    T1 = v(-vh.y * inv, vh.x * inv, 0.0f);
  } else {
    T1 = v(1.0f, 0.0f, 0.0f);
  }
  V T2 = vcross(vh, T1);
  float r = sqrtf(u1);
  float s, c;
  sincos_2pi(u2, &s, &c);
  float t1 = r * c;
  float t2 = r * s;
  float sw = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - sw) * sqrtf(1.0f - t1 * t1) + sw * t2;
  float w2 = (1.0f - t1 * t1) - t2 * t2;
  float w = sqrtf(w2 > 0.0f ? w2 : 0.0f);
  V nh = v((t1 * T1.x + t2 * T2.x) + w * vh.x, (t1 * T1.y + t2 * T2.y) + w * vh.y,
           (t1 * T1.z + t2 * T2.z) + w * vh.z);
  return vnorm(v(alpha * nh.x, alpha * nh.y, nh.z > 0.0f ? nh.z : 0.0f));
}

static int bsdf_sample(const Bsdf* b, V n, V wo, float ul, float u1, float u2, V* wi, V* f, float* pdf,
                       float* cos_i) {
  float co = vdot(n, wo);
  if (co <= 0.0f) return 0;
  V t, bt;
  onb(n, &t, &bt);
  float ps = spec_prob(b, co);
  V w;
  if (ul < ps) {
    V wol = v(vdot(wo, t), vdot(wo, bt), co);
    V hl = sample_vndf(b->alpha, wol, u1, u2);
    V h = from_local(t, bt, n, hl);
    float d = vdot(wo, h);
    w = vsub(vscl(h, 2.0f * d), wo);
  } else {
    float r = sqrtf(u1);
    float s, c;
    sincos_2pi(u2, &s, &c);
    float z2 = 1.0f - u1;
    V l = v(r * c, r * s, sqrtf(z2 > 0.0f ? z2 : 0.0f));
    w = from_local(t, bt, n, l);
  }
  *wi = w;
  *f = bsdf_eval(b, n, wo, w, pdf, cos_i);
  return *pdf > 0.0f && *cos_i > 0.0f;
}

static void mat_from12(const float* a, OMat* o) {
  for (int k = 0; k < 3; ++k) {
    o->base[k] = a[k];
    o->emission[k] = a[4 + k];
  }
  o->metallic = a[3];
  o->roughness = a[7];
  o->specular = a[8];
  o->light_pdf = 0.0f;
  o->tex = 0;
}

/* ---- image textures: texcoords from the barycentrics of the hit point
 * (normal equations of p - v0 = b1 e1 + b2 e2, clamped to [0, 1]), then a
 * bilinear, repeating fetch of sRGB texels decoded through the table; texel
 * centres at half-integers, v = 0 the bottom row. */
static void bary_of(V v0, V v1, V v2, V p, float* b1, float* b2) {
  V e1 = vsub(v1, v0), e2 = vsub(v2, v0), w = vsub(p, v0);
  float d00 = vdot(e1, e1), d01 = vdot(e1, e2), d11 = vdot(e2, e2);
  float d20 = vdot(w, e1), d21 = vdot(w, e2);
  float den = d00 * d11 - d01 * d01;
  float x = 0.0f, y = 0.0f;
  if (den != 0.0f) {
    x = (d11 * d20 - d01 * d21) / den;
    y = (d00 * d21 - d01 * d20) / den;
  }
  *b1 = x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f;
  *b2 = y > 0.0f ? (y < 1.0f ? y : 1.0f) : 0.0f;
}

static float lerp2(const float* lut, const uint8_t* a, const uint8_t* b, const uint8_t* c, const uint8_t* d, int k,
                   float fx, float fy) {
  float top = lut[a[k]] * (1.0f - fx) + lut[b[k]] * fx;
  float bot = lut[c[k]] * (1.0f - fx) + lut[d[k]] * fx;
  return top * (1.0f - fy) + bot * fy;
}

static void tex_base(const orc_scene* s, uint32_t tri, uint32_t img, float b1, float b2, float out[3]) {
  const float* uv = &s->uvs[6 * (size_t)tri];
  float b0 = (1.0f - b1) - b2;
  float u = (uv[0] * b0 + uv[2] * b1) + uv[4] * b2;
  float vv = (uv[1] * b0 + uv[3] * b1) + uv[5] * b2;
  u = u - floorf(u);
  vv = vv - floorf(vv);
  int w = (int)s->img[3 * img + 1], h = (int)s->img[3 * img + 2];
  float x = u * (float)w - 0.5f;
  float y = (1.0f - vv) * (float)h - 0.5f;
  float xf = floorf(x), yf = floorf(y);
  float fx = x - xf, fy = y - yf;
  int x0 = (int)xf, y0 = (int)yf;
  int xa = x0 < 0 ? w - 1 : x0, xb = x0 + 1 >= w ? 0 : x0 + 1;
  int ya = y0 < 0 ? h - 1 : y0, yb = y0 + 1 >= h ? 0 : y0 + 1;
  const uint8_t* t = s->texels + 4 * (size_t)s->img[3 * img];
  const uint8_t* p00 = t + 4 * ((size_t)ya * w + xa);
  const uint8_t* p10 = t + 4 * ((size_t)ya * w + xb);
  const uint8_t* p01 = t + 4 * ((size_t)yb * w + xa);
  const uint8_t* p11 = t + 4 * ((size_t)yb * w + xb);
  for (int k = 0; k < 3; ++k) out[k] = lerp2(s->lut, p00, p10, p01, p11, k, fx, fy);
}

void orc_bsdf_eval(const float* mat, const float* n, const float* wo, const float* wi, float* out5) {
  OMat m;
  mat_from12(mat, &m);
  Bsdf b = bsdf_make(&m);
  float pdf, ci;
  V f = bsdf_eval(&b, v(n[0], n[1], n[2]), v(wo[0], wo[1], wo[2]), v(wi[0], wi[1], wi[2]), &pdf, &ci);
  out5[0] = f.x;
  out5[1] = f.y;
  out5[2] = f.z;
  out5[3] = pdf;
  out5[4] = ci;
}

void orc_bsdf_sample(const float* mat, const float* n, const float* wo, float ul, float u1, float u2,
                     float* out10) {
  OMat m;
  mat_from12(mat, &m);
  Bsdf b = bsdf_make(&m);
  V wi = v(0, 0, 0), f = v(0, 0, 0);
  float pdf = 0.0f, ci = 0.0f;
  int ok = bsdf_sample(&b, v(n[0], n[1], n[2]), v(wo[0], wo[1], wo[2]), ul, u1, u2, &wi, &f, &pdf, &ci);
  out10[0] = wi.x;
  out10[1] = wi.y;
  out10[2] = wi.z;
  out10[3] = f.x;
  out10[4] = f.y;
  out10[5] = f.z;
  out10[6] = pdf;
  out10[7] = ci;
  out10[8] = (float)ok;
  out10[9] = 0.0f;
}

void orc_ggx(float roughness, float cos_h, float cos_v, float* out2) {
  OMat m;
  memset(&m, 0, sizeof m);
  m.roughness = roughness;
  Bsdf b = bsdf_make(&m);
  out2[0] = ggx_d(&b, cos_h);
  out2[1] = smith_lambda(&b, cos_v);
}

uint32_t orc_light_table(const orc_scene* s, float* pmf, float* cdf, float* area, float* mat_pdf) {
  uint32_t n = s->n_area + s->n_point;
  for (uint32_t i = 0; i < n; ++i) {
    if (pmf) pmf[i] = s->light_pmf[i];
    if (cdf) cdf[i] = s->light_cdf[i];
  }
  for (uint32_t i = 0; i < s->n_area; ++i)
    if (area) area[i] = s->light_area[i];
  for (uint32_t m = 0; m < s->n_mats; ++m)
    if (mat_pdf) mat_pdf[m] = s->mats[m].light_pdf;
  return s->n_lights;
}

/* ------------------------------------------------------------------------ */
/* integrator                                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
  float m[12];
  float tan_half, aspect;
  uint32_t width, height;
} Cam;

static void cam_from16(const float* c, Cam* o) {
  memcpy(o->m, c, sizeof o->m);
  o->tan_half = c[12];
  o->aspect = c[13];
  memcpy(&o->width, &c[14], 4);
  memcpy(&o->height, &c[15], 4);
}

/* Pinhole camera: pixel (px, py) + jitter -> NDC -> camera space
 * (x * tan(fov/2) * aspect, y * tan(fov/2), -1) -> world, normalised. */
static V camera_dir(const Cam* c, uint32_t px, uint32_t py, float u0, float u1) {
  float fx = ((float)px + u0) / (float)c->width;
  float fy = ((float)py + u1) / (float)c->height;
  float nx = 2.0f * fx - 1.0f;
  float ny = 1.0f - 2.0f * fy;
  float cx = (nx * c->tan_half) * c->aspect;
  float cy = ny * c->tan_half;
  V d = v((c->m[0] * cx + c->m[1] * cy) - c->m[2], (c->m[4] * cx + c->m[5] * cy) - c->m[6],
          (c->m[8] * cx + c->m[9] * cy) - c->m[10]);
  return vnorm(d);
}

void orc_camera_ray(const float* camera, uint32_t px, uint32_t py, float u0, float u1, float* out6) {
  Cam cam;
  cam_from16(camera, &cam);
  V d = camera_dir(&cam, px, py, u0, u1);
  out6[0] = cam.m[3];
  out6[1] = cam.m[7];
  out6[2] = cam.m[11];
  out6[3] = d.x;
  out6[4] = d.y;
  out6[5] = d.z;
}

static uint32_t pick_light(const float* cdf, uint32_t n, float u) {
  /* first index with u < cdf[i] */
  uint32_t i = 0;
  while (i + 1 < n && !(u < cdf[i])) ++i;
  return i;
}

typedef struct {
  uint64_t cam, ext, shadow;
  TCount tc;
} PathCount;

/* One path: pbrt-style unidirectional path tracing with next-event
 * estimation and the power heuristic (Veach 1997, ch. 9), PCG32 stream
 * `pixel` advanced to ((frame << 20) | sample) * dims, 8 draws per vertex
 * (with the stratified sampler the camera's two and the first hit's six
 * draws that are used are replaced by strat2 points; the stream still
 * advances past them).
 * Each vertex: emission (MIS-weighted against the light pdf unless it is the
 * camera ray's hit), then below max_depth: light pick / point / BSDF eval /
 * shadow ray, BSDF sample, Russian roulette from depth rr_depth on with
 * q = min(0.95, max throughput). */
static V trace_path(const orc_scene* s, const Cam* cam, const orc_cfg* cfg, uint32_t frame, uint32_t pixel,
                    uint32_t sample, PathCount* pc) {
  uint32_t dims = 8u * (cfg->max_depth + 2u);
  Pcg rng = pcg_seed(cfg->seed, pixel);
  pcg_advance(&rng, (((uint64_t)frame << 20) | (uint64_t)sample) * (uint64_t)dims);
  float u[8];
  for (int k = 0; k < 8; ++k) u[k] = unit_float(pcg_next(&rng));
  if (cfg->sampler) strat2(cfg->seed, frame, pixel, 0u, sample, &u[0], &u[1]);
  uint32_t px = pixel % cam->width, py = pixel / cam->width;
  V o = v(cam->m[3], cam->m[7], cam->m[11]);
  V d = camera_dir(cam, px, py, u[0], u[1]);
  V tp = v(1.0f, 1.0f, 1.0f);
  V L = v(0.0f, 0.0f, 0.0f);
  float pdf_prev = 0.0f;
  for (uint32_t depth = 0;; ++depth) {
    float t;
    if (depth == 0)
      pc->cam++;
    else
      pc->ext++;
    uint32_t id = closest(s, o, d, FLT_MAX_F, 0, &t, &pc->tc);
    if (id == 0xFFFFFFFFu) break;
    const float* p9 = &s->verts[9 * (size_t)id];
    V v0 = v(p9[0], p9[1], p9[2]), v1 = v(p9[3], p9[4], p9[5]), v2 = v(p9[6], p9[7], p9[8]);
    const OMat* m = &s->mats[s->mat_ids[id]];
    V p = v(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    OMat mtex;
    if (m->tex) {
      float b1, b2;
      mtex = *m;
      bary_of(v0, v1, v2, p, &b1, &b2);
      tex_base(s, id, m->tex - 1u, b1, b2, mtex.base);
      m = &mtex;
    }
    V wo = vneg(d);
    V ng = vnorm(vcross(vsub(v1, v0), vsub(v2, v0)));
    float cg = vdot(ng, wo);
    if ((m->emission[0] > 0.0f || m->emission[1] > 0.0f || m->emission[2] > 0.0f) && cg > 0.0f) {
      float w = 1.0f;
      if (depth > 0u) {
        float pl = ((m->light_pdf * t) * t) / cg;
        float b2 = pdf_prev * pdf_prev, l2 = pl * pl;
        w = b2 / (b2 + l2);
      }
      L = v(L.x + (tp.x * m->emission[0]) * w, L.y + (tp.y * m->emission[1]) * w,
            L.z + (tp.z * m->emission[2]) * w);
    }
    if (depth >= cfg->max_depth) break;
    for (int k = 0; k < 8; ++k) u[k] = unit_float(pcg_next(&rng));
    if (cfg->sampler && depth == 0u) {
      strat2(cfg->seed, frame, pixel, 1u, sample, &u[1], &u[2]); /* light point */
      strat2(cfg->seed, frame, pixel, 2u, sample, &u[4], &u[5]); /* BSDF direction */
      strat2(cfg->seed, frame, pixel, 3u, sample, &u[0], &u[3]); /* light pick, lobe */
    }
    V ns = cg < 0.0f ? vneg(ng) : ng;
    Bsdf bs = bsdf_make(m);
    if (s->n_lights > 0u) {
      uint32_t li = pick_light(s->light_cdf, s->n_lights, u[0]);
      float psel = s->light_pmf[li];
      int is_area = li < s->n_area;
      V wi, Le;
      float d2, dist, pdf_l = 0.0f;
      int ok;
      if (is_area) {
        uint32_t tid = s->light_tri[li];
        const float* q9 = &s->verts[9 * (size_t)tid];
        V lv0 = v(q9[0], q9[1], q9[2]), lv1 = v(q9[3], q9[4], q9[5]), lv2 = v(q9[6], q9[7], q9[8]);
        float area = s->light_area[li];
        const OMat* lm = &s->mats[s->mat_ids[tid]];
        Le = v(lm->emission[0], lm->emission[1], lm->emission[2]);
        /* uniform point on the triangle (Osada et al. 2002): b0 = 1 - sqrt(u1), b1 = u2 sqrt(u1) */
        float su = sqrtf(u[1]);
        float b0 = 1.0f - su;
        float b1 = u[2] * su;
        float b2 = (1.0f - b0) - b1;
        V qp = v((lv0.x * b0 + lv1.x * b1) + lv2.x * b2, (lv0.y * b0 + lv1.y * b1) + lv2.y * b2,
                 (lv0.z * b0 + lv1.z * b1) + lv2.z * b2);
        V nl = vnorm(vcross(vsub(lv1, lv0), vsub(lv2, lv0)));
        V dv = vsub(qp, p);
        d2 = vdot(dv, dv);
        dist = sqrtf(d2);
        wi = vscl(dv, 1.0f / dist);
        float cl = -vdot(nl, wi);
        ok = cl > 0.0f && d2 > 0.0f;
        if (ok) pdf_l = (psel * d2) / (area * cl);
      } else {
        const float* pl = &s->points[6 * (size_t)(li - s->n_area)];
        Le = v(pl[3], pl[4], pl[5]);
        V dv = vsub(v(pl[0], pl[1], pl[2]), p);
        d2 = vdot(dv, dv);
        dist = sqrtf(d2);
        wi = vscl(dv, 1.0f / dist);
        ok = d2 > 0.0f;
      }
      if (ok) {
        float pb, ci;
        V f = bsdf_eval(&bs, ns, wo, wi, &pb, &ci);
        if (ci > 0.0f) {
          float sfac;
          if (is_area) {
            float l2 = pdf_l * pdf_l, bb2 = pb * pb;
            float w = l2 / (l2 + bb2);
            sfac = (ci * w) / pdf_l;
          } else {
            sfac = ci / (d2 * psel);
          }
          V con = v(((tp.x * f.x) * Le.x) * sfac, ((tp.y * f.y) * Le.y) * sfac, ((tp.z * f.z) * Le.z) * sfac);
          if (con.x > 0.0f || con.y > 0.0f || con.z > 0.0f) {
            pc->shadow++;
            if (!occluded(s, offset_ray(p, ns), wi, dist * 0.9999f, 0))
              L = v(L.x + con.x, L.y + con.y, L.z + con.z);
          }
        }
      }
    }
    V wi, f;
    float pdf, ci;
    if (!bsdf_sample(&bs, ns, wo, u[3], u[4], u[5], &wi, &f, &pdf, &ci)) break;
    float sc = ci / pdf;
    V tpn = v(tp.x * (f.x * sc), tp.y * (f.y * sc), tp.z * (f.z * sc));
    int cont = tpn.x > 0.0f || tpn.y > 0.0f || tpn.z > 0.0f;
    if (cont && depth >= cfg->rr_depth) {
      float qq = mx(tpn.x, mx(tpn.y, tpn.z));
      if (qq > 0.95f) qq = 0.95f;
      if (u[6] >= qq)
        cont = 0;
      else
        tpn = v(tpn.x / qq, tpn.y / qq, tpn.z / qq);
    }
    if (!cont) break;
    tp = tpn;
    pdf_prev = pdf;
    o = offset_ray(p, ns);
    d = wi;
  }
  return L;
}

/* ---- threaded film over rows ---- */

typedef struct {
  const orc_scene* s;
  const Cam* cam;
  const orc_cfg* cfg;
  uint32_t frame, spp, y0, y1;
  float* rgb;
  double* rgb64;
  atomic_uint next_row;
  atomic_ullong cam_rays, ext_rays, shadow_rays, nonfinite, nodes, tris;
} Job;

static void* worker(void* arg) {
  Job* J = (Job*)arg;
  PathCount pc;
  memset(&pc, 0, sizeof pc);
  uint64_t bad = 0;
  float inv = 1.0f / (float)J->spp;
  for (;;) {
    uint32_t y = atomic_fetch_add(&J->next_row, 1u);
    if (y >= J->y1) break;
    for (uint32_t x = 0; x < J->cam->width; ++x) {
      uint32_t pixel = y * J->cam->width + x;
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      double dx = 0.0, dy = 0.0, dz = 0.0;
      for (uint32_t sidx = 0; sidx < J->spp; ++sidx) {
        V L = trace_path(J->s, J->cam, J->cfg, J->frame, pixel, sidx, &pc);
        if (!(isfinite(L.x) && isfinite(L.y) && isfinite(L.z))) {
          ++bad;
          continue;
        }
        ax = ax + L.x;
        ay = ay + L.y;
        az = az + L.z;
        dx += (double)L.x;
        dy += (double)L.y;
        dz += (double)L.z;
      }
      size_t o = 3 * ((size_t)(y - J->y0) * J->cam->width + x);
      if (J->rgb) {
        J->rgb[o] = ax * inv;
        J->rgb[o + 1] = ay * inv;
        J->rgb[o + 2] = az * inv;
      }
      if (J->rgb64) {
        J->rgb64[o] = dx / (double)J->spp;
        J->rgb64[o + 1] = dy / (double)J->spp;
        J->rgb64[o + 2] = dz / (double)J->spp;
      }
    }
  }
  atomic_fetch_add(&J->cam_rays, pc.cam);
  atomic_fetch_add(&J->ext_rays, pc.ext);
  atomic_fetch_add(&J->shadow_rays, pc.shadow);
  atomic_fetch_add(&J->nonfinite, bad);
  atomic_fetch_add(&J->nodes, pc.tc.nodes);
  atomic_fetch_add(&J->tris, pc.tc.tris);
  return NULL;
}

static int run_job(Job* J, uint32_t threads) {
  if (threads == 0) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    threads = n > 0 ? (uint32_t)n : 1u;
  }
  if (threads > 256) threads = 256;
  pthread_t th[256];
  uint32_t started = 0;
  for (uint32_t i = 0; i < threads; ++i) {
    if (pthread_create(&th[i], NULL, worker, J) != 0) break;
    ++started;
  }
  if (started == 0) worker(J);
  for (uint32_t i = 0; i < started; ++i) pthread_join(th[i], NULL);
  return 0;
}

int orc_render(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame, uint32_t spp,
               uint32_t y0, uint32_t y1, float* rgb, orc_stats* stats) {
  Cam cam;
  cam_from16(camera, &cam);
  if (spp == 0 || y1 > cam.height || y0 > y1) return -1;
  Job J;
  memset(&J, 0, sizeof J);
  J.s = s;
  J.cam = &cam;
  J.cfg = cfg;
  J.frame = frame;
  J.spp = spp;
  J.y0 = y0;
  J.y1 = y1;
  J.rgb = rgb;
  atomic_init(&J.next_row, y0);
  run_job(&J, cfg->threads);
  if (stats) {
    stats->camera_rays = J.cam_rays;
    stats->extension_rays = J.ext_rays;
    stats->shadow_rays = J.shadow_rays;
    stats->nonfinite = J.nonfinite;
    stats->nodes = J.nodes;
    stats->tris = J.tris;
  }
  return 0;
}

int orc_render64(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame, uint32_t spp,
                 double* rgb) {
  Cam cam;
  cam_from16(camera, &cam);
  if (spp == 0) return -1;
  Job J;
  memset(&J, 0, sizeof J);
  J.s = s;
  J.cam = &cam;
  J.cfg = cfg;
  J.frame = frame;
  J.spp = spp;
  J.y0 = 0;
  J.y1 = cam.height;
  J.rgb64 = rgb;
  atomic_init(&J.next_row, 0u);
  run_job(&J, cfg->threads);
  return 0;
}

typedef struct {
  const orc_scene* s;
  const Cam* cam;
  const orc_cfg* cfg;
  uint32_t frame, sample, y0, y1;
  uint32_t* ids;
  float* t;
  atomic_uint next_row;
} HitJob;

static void* hit_worker(void* arg) {
  HitJob* J = (HitJob*)arg;
  const Cam* cam = J->cam;
  uint32_t dims = 8u * (J->cfg->max_depth + 2u);
  for (;;) {
    uint32_t y = atomic_fetch_add(&J->next_row, 1u);
    if (y >= J->y1) break;
    for (uint32_t x = 0; x < cam->width; ++x) {
      uint32_t pixel = y * cam->width + x;
      Pcg rng = pcg_seed(J->cfg->seed, pixel);
      pcg_advance(&rng, (((uint64_t)J->frame << 20) | (uint64_t)J->sample) * (uint64_t)dims);
      float u0 = unit_float(pcg_next(&rng));
      float u1 = unit_float(pcg_next(&rng));
      if (J->cfg->sampler) strat2(J->cfg->seed, J->frame, pixel, 0u, J->sample, &u0, &u1);
      V o = v(cam->m[3], cam->m[7], cam->m[11]);
      V d = camera_dir(cam, x, y, u0, u1);
      float tt;
      size_t k = (size_t)(y - J->y0) * cam->width + x;
      J->ids[k] = closest(J->s, o, d, FLT_MAX_F, 0, &tt, NULL);
      if (J->t) J->t[k] = tt;
    }
  }
  return NULL;
}

int orc_primary_hits_rows(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame,
                          uint32_t sample, uint32_t y0, uint32_t y1, uint32_t* ids, float* t) {
  Cam cam;
  cam_from16(camera, &cam);
  if (y1 > cam.height || y0 > y1) return -1;
  HitJob J;
  memset(&J, 0, sizeof J);
  J.s = s;
  J.cam = &cam;
  J.cfg = cfg;
  J.frame = frame;
  J.sample = sample;
  J.y0 = y0;
  J.y1 = y1;
  J.ids = ids;
  J.t = t;
  atomic_init(&J.next_row, y0);
  uint32_t threads = cfg->threads;
  if (threads == 0) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    threads = n > 0 ? (uint32_t)n : 1u;
  }
  if (threads > 256) threads = 256;
  pthread_t th[256];
  uint32_t started = 0;
  for (uint32_t i = 0; i < threads; ++i) {
    if (pthread_create(&th[i], NULL, hit_worker, &J) != 0) break;
    ++started;
  }
  if (started == 0) hit_worker(&J);
  for (uint32_t i = 0; i < started; ++i) pthread_join(th[i], NULL);
  return 0;
}

int orc_primary_hits(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame,
                     uint32_t sample, uint32_t* ids, float* t) {
  Cam cam;
  cam_from16(camera, &cam);
  return orc_primary_hits_rows(s, camera, cfg, frame, sample, 0, cam.height, ids, t);
}
