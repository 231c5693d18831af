/*
 * oracle.h -- CPU restatement of the path-tracing hot path. TEST
 * INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU-baseline leg of
 * the benchmark load it as the checker. The product library never links it.
 *
 * The reference (shiinamiyuki/aqua-engine, crate `arukas`) ships no renderer
 * code (src/lib.rs is empty), so this restates, from the published methods
 * each function names, the algorithm the GPU kernels implement on the
 * reference's inputs: the scene/integrator JSON and .mesh data of
 * scenes/cbox.json, scenes/integrator.json and scenes/<name>.mesh. It is its own
 * code -- its own BVH (a different tree from the GPU's), its own traversal
 * (box tested when popped), a one-path-at-a-time integrator -- with the same
 * FP32 operation sequence for every quantity that reaches a pixel, so that
 * GPU and CPU agree bit for bit.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   pcg32 / advance   published pcg32-demo vector; advance == n x next
 *   strat2            (0, m, 2)-net property of every power-of-two prefix;
 *                     Sobol dimension values from the direction numbers
 *   sincos_2pi        numpy float64 sin/cos, <= 2 ulp
 *   camera            closed forms (centre ray = -Z axis, corners at fov/2)
 *   closest / any hit brute force over all triangles; float64 Moller-Trumbore
 *                     on random rays; watertightness on closed meshes
 *   BSDF              Lambert rho/pi; GGX D normalisation; reciprocity;
 *                     pdf integrates to 1; sampling histogram vs pdf (chi^2)
 *   integrator        white furnace (closed form sum rho^k); Lambert's
 *                     polygon irradiance for direct light under a quad
 *   light table       area / power by hand for the box room
 */
#ifndef AQUA_ORACLE_H
#define AQUA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Layouts match include/aqua/aqua.h: material = 12 floats
 * (base rgb, metallic, emission rgb, roughness, specular, pad x3),
 * camera = 16 floats (3x4 matrix, tan_half, aspect, width bits, height bits),
 * point light = 6 floats (pos, intensity). */
int orc_scene_create(uint32_t n_tris, const float* verts, const uint32_t* mat_ids, uint32_t n_mats,
                     const float* mats, uint32_t n_point, const float* points, orc_scene** out);
void orc_scene_destroy(orc_scene* s);
/* Image textures: uvs 6 floats per triangle (u, v per corner, load order),
 * dims (width, height) per image, texels the images' 8-bit sRGB RGBA
 * concatenated (row 0 = top). Materials name image k as base_color_tex
 * (float slot 9, as uint32 bits) = k + 1. */
int orc_scene_set_textures(orc_scene* s, const float* uvs, uint32_t n_images, const uint32_t* dims,
                           const uint8_t* texels);
/* Oracle BVH statistics: nodes, max depth. */
void orc_scene_bvh_info(const orc_scene* s, uint32_t* n_nodes, uint32_t* max_depth);

typedef struct orc_cfg {
  uint32_t max_depth;
  uint32_t rr_depth;
  uint32_t seed;
  uint32_t threads;   /* 0 = all online CPUs */
  uint32_t sampler;   /* 0 = PCG32 only, 1 = stratified camera / first-hit dims */
} orc_cfg;

typedef struct orc_stats {
  uint64_t camera_rays, extension_rays, shadow_rays, nonfinite;
  uint64_t nodes, tris;  /* closest-hit traversal work on the oracle's own BVH */
} orc_stats;

/* Points first .. first + n - 1 of the stratified sampler's 2-D pattern
 * `pair` of one pixel and frame: out[2 k], out[2 k + 1]. */
void orc_strat2_seq(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t pair, uint32_t first, uint32_t n,
                    float* out);

/* Film over pixels [y0, y1) x [0, width): rgb receives the mean radiance
 * (float sum over samples 0..spp-1 in order, times 1/spp). */
int orc_render(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame, uint32_t spp,
               uint32_t y0, uint32_t y1, float* rgb, orc_stats* stats);
/* Same, accumulated in double. */
int orc_render64(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame, uint32_t spp,
                 double* rgb);

int orc_primary_hits(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame,
                     uint32_t sample, uint32_t* ids, float* t);
/* Same over pixel rows [y0, y1) only (ids / t hold (y1 - y0) * width entries);
 * rows are shared among cfg->threads threads (0 = all online CPUs). */
int orc_primary_hits_rows(const orc_scene* s, const float* camera, const orc_cfg* cfg, uint32_t frame,
                          uint32_t sample, uint32_t y0, uint32_t y1, uint32_t* ids, float* t);
/* rays: 8 floats each (o, tmax, d, 0). brute != 0 tests every triangle. */
int orc_trace_closest(const orc_scene* s, uint32_t n, const float* rays, int brute, uint32_t* ids, float* t);
int orc_trace_occluded(const orc_scene* s, uint32_t n, const float* rays, int brute, uint8_t* occ);

/* Building blocks, exposed for their own tests. */
void orc_pcg32_seq(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t* out);
void orc_sincos_2pi(uint32_t n, const float* u, float* s, float* c);
/* mat: 12 floats; n, wo, wi unit vectors (3 floats each). out: f rgb, pdf, cos_i. */
void orc_bsdf_eval(const float* mat, const float* n, const float* wo, const float* wi, float* out5);
/* out: wi xyz, f rgb, pdf, cos_i, ok. */
void orc_bsdf_sample(const float* mat, const float* n, const float* wo, float ul, float u1, float u2,
                     float* out10);
/* GGX normal distribution D(cos_h) and Smith Lambda(cos_v) for alpha =
 * max(roughness^2, 1e-3). */
void orc_ggx(float roughness, float cos_h, float cos_v, float* out2);
/* Camera ray (origin xyz, unit direction xyz) of pixel (px, py) with jitter. */
void orc_camera_ray(const float* camera, uint32_t px, uint32_t py, float u0, float u1, float* out6);
/* Light table as the integrator uses it: pmf / cdf per light, area per
 * emissive triangle, light pdf per material. Returns the light count. */
uint32_t orc_light_table(const orc_scene* s, float* pmf, float* cdf, float* area, float* mat_pdf);

#ifdef __cplusplus
}
#endif

#endif
