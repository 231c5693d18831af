/*
 * aqua.h -- C ABI of the MI355X (gfx950) path tracer.
 *
 * The reference crate (`arukas`, Cargo.toml:2-5) ships no code: src/lib.rs is
 * empty, so it has no FFI of its own. What it does define is data: the scene
 * JSON (scenes/cbox.json:2-627, scenes/room.json:2-3899), the integrator JSON
 * (scenes/integrator.json:2-8) and the BSON `.mesh` geometry files
 * (scenes/CornellBox-Original_floor_0.mesh). Every entry point below takes or
 * produces exactly those formats, so a Rust host of the reference binds this
 * header with `extern "C"` and `#[repr(C)]` mirrors of the structs.
 *
 * Conventions
 *   - Every call returns an `int` status (AQUA_OK == 0). On failure a
 *     thread-local message is available from aqua_last_error(). No C++
 *     exception and no abort crosses this boundary.
 *   - Structs hold fixed-width scalars only (no bool, no pointers to C++
 *     types); sizes are fixed and listed beside each struct.
 *   - Handles are opaque, created by *_load / *_create, released by *_destroy.
 *   - The renderer computes on a GPU only. There is no CPU fallback: a
 *     negative device index is rejected with AQUA_ERR_INVALID_ARG, and a host
 *     without a usable gfx950 device gets AQUA_ERR_NO_DEVICE.
 */
#ifndef AQUA_AQUA_H
#define AQUA_AQUA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AQUA_ABI_VERSION 1

enum aqua_status {
  AQUA_OK = 0,
  AQUA_ERR_IO = 1,            /* file missing / unreadable / unwritable          */
  AQUA_ERR_PARSE = 2,         /* malformed JSON or BSON, unknown enum tag         */
  AQUA_ERR_MISSING_MESH = 3,  /* a shape names a .mesh file that is absent        */
  AQUA_ERR_NO_DEVICE = 4,     /* no GPU, or the device index is out of range      */
  AQUA_ERR_OOM = 5,           /* host or device allocation failed                 */
  AQUA_ERR_HIP = 6,           /* a HIP runtime call failed (message has details)  */
  AQUA_ERR_INVALID_ARG = 7,   /* null handle, bad size, out-of-range index        */
  AQUA_ERR_NONFINITE = 8,     /* non-finite input geometry or parameters          */
  AQUA_ERR_UNSUPPORTED = 9    /* a feature of the format this build does not run  */
};

typedef struct aqua_scene aqua_scene;       /* immutable after load; shareable */
typedef struct aqua_renderer aqua_renderer; /* one per device and thread       */

/* ---- scene ------------------------------------------------------------- */

/* 32 bytes. */
typedef struct aqua_load_opts {
  uint32_t skip_missing;      /* 1: shapes whose .mesh is absent are skipped and counted
                                 (room.json:3021 names the withheld mesh recorded in
                                 .MISSING_LARGE_BLOBS:1); 0: AQUA_ERR_MISSING_MESH        */
  uint32_t image_as_constant; /* 1: `{"Image": path}` colours load as constant 0.5 grey
                                 (room.json:6 and 15 others); 0: the image (JPEG, path
                                 relative to the JSON) is decoded and sampled through the
                                 meshes' texcoords                                         */
  uint32_t emission_from_mtl; /* 1: a named BSDF whose emission is 0 takes the `Ke` of the
                                 same-named material of a .mtl file next to the JSON (the
                                 room's lamp quads: room.json:251-258 name MiddleLight,
                                 living_room.mtl:193-194 gives it Ke 5 5 5); 0: as written */
  uint32_t reserved[5];
} aqua_load_opts;

/* 48 bytes. The subset of the 17 Principled parameters (cbox.json:5-63) that
 * the metallic-roughness BSDF uses; the loader checks that the rest hold the
 * values the shipped scenes use everywhere (subsurface, sheen, clearcoat,
 * transmission 0) and rejects others with AQUA_ERR_UNSUPPORTED. */
typedef struct aqua_material {
  float base_color[3]; /* linear RGB (Srgb inputs decoded with IEC 61966-2-1)   */
  float metallic;
  float emission[3];   /* linear RGB radiance; > 0 makes the mesh an area light  */
  float roughness;     /* GGX alpha = max(roughness^2, 1e-3)                     */
  float specular;      /* dielectric F0 = 0.08 * specular                        */
  uint32_t base_color_tex; /* 0: base_color; k > 0: image k - 1 of the scene, sampled
                              bilinearly (repeat) at the hit's texcoords, sRGB texels
                              decoded to linear; base_color then holds its mean     */
  float pad[2];
} aqua_material;

/* 64 bytes. Pinhole camera. Camera space looks down -Z, +Y up (the shipped
 * cameras sit at +Z looking at the scene: cbox.json:518-541). */
typedef struct aqua_camera {
  float cam_to_world[12]; /* row-major 3x4; columns: right, up, back (+Z), origin */
  float tan_half_fov;     /* tan(vertical field of view / 2)                      */
  float aspect;           /* width / height                                       */
  uint32_t width, height;
} aqua_camera;

/* 24 bytes. `Point` light (cbox.json:544-561): radiant intensity, W/sr. */
typedef struct aqua_point_light {
  float pos[3];
  float intensity[3];
} aqua_point_light;

/* 64 bytes. */
typedef struct aqua_scene_info {
  uint32_t n_triangles;
  uint32_t n_materials;
  uint32_t n_shapes;            /* shapes loaded                          */
  uint32_t n_skipped_shapes;    /* shapes skipped for a missing .mesh     */
  uint32_t n_emissive_triangles;
  uint32_t n_point_lights;
  uint32_t n_image_textures;    /* Image colours (decoded, or replaced by a
                                   constant with image_as_constant)        */
  uint32_t n_vertices;          /* summed over the loaded meshes          */
  float bounds_lo[3];
  float bounds_hi[3];
  uint32_t n_images;            /* decoded images the scene holds         */
  uint32_t has_texcoords;       /* 1: per-triangle texcoords are present  */
} aqua_scene_info;

/* Load a scene JSON (schema of scenes/cbox.json) and every .mesh it names
 * (paths relative to the JSON's directory; '\\' separators accepted). */
int aqua_scene_load(const char* scene_json_path, const aqua_load_opts* opts, aqua_scene** out);

/* Build a scene straight from flat arrays: `verts` holds 9 floats per
 * triangle (v0, v1, v2), `mat_ids` one material index per triangle. */
int aqua_scene_from_arrays(uint32_t n_triangles, const float* verts, const uint32_t* mat_ids,
                           uint32_t n_materials, const aqua_material* materials,
                           const aqua_camera* camera, uint32_t n_point_lights,
                           const aqua_point_light* point_lights, aqua_scene** out);

int aqua_scene_get_info(const aqua_scene* scene, aqua_scene_info* out);
/* Flattened geometry in load order: verts[9*n], mat_ids[n] (either may be NULL). */
int aqua_scene_get_triangles(const aqua_scene* scene, float* verts, uint32_t* mat_ids);
int aqua_scene_get_materials(const aqua_scene* scene, aqua_material* out);
int aqua_scene_get_camera(const aqua_scene* scene, aqua_camera* out);
int aqua_scene_get_point_lights(const aqua_scene* scene, aqua_point_light* out);
/* Texcoords in load order, 6 floats per triangle (u, v of v0, v1, v2; the
 * .mesh `texcoords`, v up); AQUA_ERR_INVALID_ARG when the scene has none. */
int aqua_scene_get_texcoords(const aqua_scene* scene, float* uv);
/* Image k (0-based) of the scene: size, and 8-bit sRGB RGBA texels (row 0 =
 * top) when rgba is not NULL. */
int aqua_scene_get_image(const aqua_scene* scene, uint32_t k, uint32_t* width, uint32_t* height, uint8_t* rgba);
/* For scenes built from arrays, before a renderer is created from them:
 * attach texcoords (6 floats per triangle) / add an image (8-bit sRGB RGBA,
 * row 0 = top) that materials name as base_color_tex = *index + 1. */
int aqua_scene_set_texcoords(aqua_scene* scene, const float* uv);
int aqua_scene_add_image(aqua_scene* scene, uint32_t width, uint32_t height, const uint8_t* rgba,
                         uint32_t* index);
void aqua_scene_destroy(aqua_scene* scene);

/* ---- images ------------------------------------------------------------- */

/* Decodes a JPEG file (the `{"Image": path}` texture format of
 * scenes/room.json:6; baseline or progressive, 8-bit, grey or YCbCr/RGB) to
 * 8-bit RGBA, row 0 = top. With rgba == NULL only the size is returned. */
int aqua_image_load(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgba);
/* Writes 8-bit RGB (channels 3) or RGBA (channels 4) pixels, row 0 = top,
 * as a PNG file (zlib-deflated, no filter). */
int aqua_image_write_png(const char* path, uint32_t width, uint32_t height, uint32_t channels,
                         const uint8_t* pixels);
/* Writes float RGB pixels, row 0 = top, as a little-endian PFM file (the
 * portable float map: HDR radiance without loss). */
int aqua_image_write_pfm(const char* path, uint32_t width, uint32_t height, const float* rgb);

/* ---- procedural scenes (written in the reference schema: JSON + .mesh) -- */

enum aqua_procgen_kind {
  AQUA_PROCGEN_BOX_ROOM = 1,    /* 5 walls (10 tris), boxes 0.6/1.2 at +-17 deg (24 tris),
                                   ceiling quad light L=(17,12,4) (CornellBox-Original.mtl:88),
                                   512x512, fov 40                                             */
  AQUA_PROCGEN_INTERIOR = 2     /* room + n_meshes seeded subdivided boxes / icospheres,
                                   GGX metallic-roughness, 3 quad lights                       */
};

/* 32 bytes. */
typedef struct aqua_procgen_params {
  uint32_t kind;         /* aqua_procgen_kind                              */
  uint32_t seed;         /* PCG32 seed of every random choice              */
  uint32_t n_meshes;     /* INTERIOR: object meshes (e.g. 150, 400)        */
  uint32_t target_tris;  /* INTERIOR: approximate triangle budget          */
  uint32_t width, height;/* camera resolution written into the JSON        */
  uint32_t reserved[2];
} aqua_procgen_params;

/* Writes <out_dir>/scene.json plus one .mesh per shape; the JSON path is
 * copied into scene_path (capacity `cap`). Same params => same bytes. */
int aqua_procgen_write(const aqua_procgen_params* params, const char* out_dir, char* scene_path,
                       size_t cap);

/* ---- integrator (scenes/integrator.json:2-8) --------------------------- */

enum aqua_integrator_type { AQUA_INTEGRATOR_PT = 0, AQUA_INTEGRATOR_NRC = 1 };
enum aqua_sampler { AQUA_SAMPLER_INDEPENDENT = 0, AQUA_SAMPLER_STRATIFIED = 1 };

/* 64 bytes. */
typedef struct aqua_integrator_cfg {
  uint32_t type;            /* "type": "pt" | "nrc"            (integrator.json:2) */
  uint32_t spp;             /* "spp"                           (integrator.json:3) */
  uint32_t max_depth;       /* "max_depth": bounces            (integrator.json:5) */
  uint32_t batch_size;      /* "batch_size"                    (integrator.json:4) */
  uint32_t training_iters;  /* "training_iters"                (integrator.json:6):
                               the online cache's Adam steps in all; after them
                               frames trace no training paths and only query
                               (0: no limit)                                       */
  float learning_rate;      /* "learning_rate"                 (integrator.json:7) */
  uint32_t visualize_cache; /* "visualize_cache"               (integrator.json:8):
                               with "nrc", 1 ends every non-training path in a
                               cache query at its first hit, so the image is the
                               emission there plus the cache's prediction; the
                               cache trains as configured. No effect on "pt".   */
  uint32_t seed;            /* "seed" (extension, default 0): PCG32 initstate     */
  uint32_t rr_depth;        /* "rr_depth" (extension, default 3)                  */
  /* "nrc" only (extensions): */
  uint32_t nrc_precision;   /* "nrc_precision": "f16" | "bf16" | "fp8" (queries; training
                               runs in BF16 for fp8)                                 */
  uint32_t nrc_train_paths; /* "nrc_train_paths": suffix paths traced for training
                               per frame (default 65536)                             */
  uint32_t nrc_query_vertex;/* "nrc_query_vertex": path vertex (1 = first hit) at
                               which other paths end in a cache query (default 2)    */
  uint32_t sampler;         /* "sampler" (extension): "independent" (0, default): every
                               draw from the pixel's PCG32 stream; "stratified" (1):
                               the camera ray's 2 and the first hit's 6 draws from
                               shuffled Owen-scrambled 2-D Sobol points over the
                               sample index (Burley 2020), seeded by seed, pixel and
                               frame; later vertices keep the PCG32 stream          */
  uint32_t reserved[3];
} aqua_integrator_cfg;

void aqua_integrator_default(aqua_integrator_cfg* out);
int aqua_integrator_load(const char* integrator_json_path, aqua_integrator_cfg* out);

/* ---- renderer ----------------------------------------------------------- */

/* 128 bytes. Counters and HIP-event timings of one aqua_render_frame call. */
typedef struct aqua_frame_stats {
  uint64_t camera_rays;
  uint64_t extension_rays;     /* closest-hit rays after the camera rays   */
  uint64_t shadow_rays;
  uint64_t nonfinite_samples;  /* samples whose radiance was zeroed        */
  double ms_frame;             /* events around the whole frame            */
  double ms_intersect;         /* summed closest-hit kernel time           */
  double ms_shade;
  double ms_shadow;
  uint32_t n_intersect_launches;
  uint32_t n_passes;
  uint64_t nodes_visited;      /* only when counting is on, else 0         */
  uint64_t tris_tested;        /* closest-hit rays only                    */
  uint64_t shadow_nodes_visited;
  uint64_t shadow_tris_tested;
  uint32_t nrc_queries;        /* paths ended by a cache query          */
  uint32_t nrc_train_records;  /* training samples produced this frame  */
  float nrc_loss;              /* last training step's loss             */
  float ms_nrc;                /* cache inference + targets + training  */
  uint64_t spill_node_visits;  /* node visits made while the ray's traversal
                                  stack had entries spilled to global memory
                                  (both traversal kernels; counting on only) */
} aqua_frame_stats;

/* 32 bytes. */
typedef struct aqua_renderer_opts {
  uint32_t samples_per_pass;   /* 0 = choose (about 128 M paths in flight)    */
  uint32_t count_traversal;    /* 1 = count nodes / triangles (slower)        */
  uint32_t timing;             /* 1 = per-kernel HIP-event timing             */
  uint32_t sort_materials;     /* 1 = order each bounce's hits by material
                                  before shading (same film)                 */
  uint32_t train_ahead;        /* "nrc": 1 = train a frame's batch at the start
                                  of the next frame, beside its first bounces
                                  (its queries wait for the update, so they see
                                  the same weights as with 0); the batch waiting
                                  is part of the cache's checkpoint           */
  uint32_t reserved[3];
} aqua_renderer_opts;

int aqua_renderer_create(const aqua_scene* scene, const aqua_integrator_cfg* cfg, int device,
                         aqua_renderer** out);
int aqua_renderer_set_opts(aqua_renderer* r, const aqua_renderer_opts* opts);

/* Render `spp` samples per pixel of frame `frame` into the renderer's FP32
 * film (cleared first). `camera` NULL = the scene's camera. Sample s of
 * pixel p draws from PCG32 stream p, advanced to ((frame << 20) | s) * D,
 * D = 8 * (max_depth + 2), so results do not depend on the pass split. */
int aqua_render_frame(aqua_renderer* r, const aqua_camera* camera, uint32_t frame, uint32_t spp,
                      aqua_frame_stats* stats);
/* Mean radiance of the last frame, RGB float per pixel, row 0 = top. */
int aqua_film_read(aqua_renderer* r, float* rgb, uint32_t* spp);
/* The last frame for display, computed on the GPU: each channel of the mean
 * radiance times 2^exposure, clamped to [0, 1], encoded with the sRGB curve
 * (IEC 61966-2-1) and rounded to 8 bits; RGBA per pixel (alpha 255), row 0 =
 * top (host array of width * height * 4). aqua_image_write_png stores it. */
int aqua_film_tonemap(aqua_renderer* r, float exposure, uint8_t* rgba);

/* FP64-accumulated render: adds samples [spp_begin, spp_begin + spp_count)
 * of `frame` into an FP64 film, so a long oracle render can be resumed;
 * spp_begin == 0 clears it first. */
int aqua_render_oracle(aqua_renderer* r, const aqua_camera* camera, uint32_t frame,
                       uint32_t spp_begin, uint32_t spp_count);
/* Mean of the FP64 film (double RGB per pixel) and the samples it holds. */
int aqua_oracle_read(aqua_renderer* r, double* rgb, uint32_t* spp);
/* relMSE = mean_{p,c} (x - y)^2 / (y^2 + 1e-2) and relRMSE =
 * sqrt(sum (x - y)^2 / sum y^2) of the FP32 film (x) against the FP64 film
 * (y), reduced in FP64 on the GPU. */
int aqua_film_error(aqua_renderer* r, double* relmse, double* relrmse);

/* First-hit triangle id (load order; 0xFFFFFFFF = miss) and distance for
 * every pixel, sample `sample` of frame `frame`. */
int aqua_primary_hits(aqua_renderer* r, const aqua_camera* camera, uint32_t frame,
                      uint32_t sample, uint32_t* prim_ids, float* t_hit);

/* Closest hit of arbitrary rays given on the host, 8 floats each:
 * ox oy oz tmax dx dy dz 0 (directions need not be unit length). */
int aqua_trace_closest(aqua_renderer* r, uint32_t n, const float* rays, uint32_t* prim_ids,
                       float* t_hit);
/* aqua_trace_closest without any allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * caller passes device arrays on the renderer's device: d_rays (8 floats per
 * ray as above, 16-byte aligned; finite, non-zero directions -- not checked
 * here), d_prim_ids and d_t_hit (n each); results are those of
 * aqua_trace_closest. At most 2^30 rays per call (larger n:
 * AQUA_ERR_INVALID_ARG). The kernels' counters, work cursors and traversal
 * spill stacks live in a fixed 33.6 MB device array that every GPU the
 * library runs on holds until the process ends; all calls on one GPU share
 * it, so calls that overlap on one GPU (other threads, other renderers) race.
 * Queued on the renderer's stream; returns once the results are written. */
int aqua_trace_closest_noalloc(aqua_renderer* r, uint32_t n, const float* d_rays, uint32_t* d_prim_ids,
                               float* d_t_hit);
/* Any hit in (0, tmax): occluded[i] = 1 / 0. */
int aqua_trace_occluded(aqua_renderer* r, uint32_t n, const float* rays, uint8_t* occluded);

void aqua_renderer_destroy(aqua_renderer* r);

/* ---- neural radiance cache ("type": "nrc", integrator.json:2) ------------
 * Multiresolution hash encoding (16 levels x 2 FP16 features, 2^log2_table
 * entries per level) of the position, degree-4 SH of the direction, plus
 * normal / roughness / albedo, into a 64-wide MLP of 4 hidden ReLU layers
 * and 3 linear outputs, run on MFMA (FP16 or BF16, FP8 e4m3 for inference)
 * with FP32 accumulation; trained with the relative L2 loss and Adam.
 *
 * A record is 16 floats: position xyz normalised to [0,1]^3, roughness,
 * direction xyz, 0, normal xyz, 0, albedo rgb, 0. */

typedef struct aqua_nrc aqua_nrc;

enum aqua_nrc_precision { AQUA_NRC_F16 = 0, AQUA_NRC_BF16 = 1, AQUA_NRC_FP8 = 2 };

/* 64 bytes. */
typedef struct aqua_nrc_cfg {
  uint32_t n_levels;        /* must be 16                                    */
  uint32_t log2_table;      /* entries per level = 2^log2_table (10..22)    */
  uint32_t base_res;        /* coarsest grid resolution (16)                */
  float level_scale;        /* resolution growth per level (1.5)            */
  float learning_rate;      /* Adam step size (integrator.json:7)           */
  float beta1, beta2, eps;  /* Adam (0.9, 0.99, 1e-15)                      */
  uint32_t train_precision; /* AQUA_NRC_F16 or AQUA_NRC_BF16                */
  uint32_t seed;            /* PCG32 seed of the initial parameters         */
  uint32_t max_batch;       /* largest training batch (rows)                */
  uint32_t reserved[5];
} aqua_nrc_cfg;

/* 64 bytes. HIP-event timings of the last training call. */
typedef struct aqua_nrc_stats {
  double ms_total;          /* all steps                                    */
  double ms_fwd_bwd;        /* encode + forward + backward, then the
                               hash-table gradient scatter                  */
  double ms_wgrad;          /* weight gradients (MFMA), on a second stream
                               beside the scatter                           */
  double ms_adam;           /* after the scatter: the table's Adam step and
                               the operand repack (the MLP's Adam step runs
                               beside, after the weight gradients)          */
  float loss;               /* relative L2 loss of the last step            */
  uint32_t steps;
  uint32_t rows;
  float ms_scatter;         /* the scatter alone (part of ms_fwd_bwd)       */
  uint32_t reserved[4];
} aqua_nrc_stats;

void aqua_nrc_default_cfg(aqua_nrc_cfg* out);
int aqua_nrc_create(int device, const aqua_nrc_cfg* cfg, aqua_nrc** out);
void aqua_nrc_destroy(aqua_nrc* nrc);
/* Parameter counts: hash table floats (levels x entries x 2) and MLP floats
 * (4 x 64 x 64 + 64 x 16, row-major W[in][out] per layer). */
int aqua_nrc_param_counts(const aqua_nrc* nrc, uint64_t* n_table, uint32_t* n_mlp);
int aqua_nrc_get_params(aqua_nrc* nrc, float* table, float* mlp);
/* Replaces the parameters and clears the Adam state. */
int aqua_nrc_set_params(aqua_nrc* nrc, const float* table, const float* mlp);
/* The FP8 query path's per-layer scales (5 each): weights are quantised as
 * e4m3(sat448(w * wscale[L])), layer inputs as e4m3(sat448(x * ascale[L])),
 * and each layer's FP32 accumulator is multiplied by 1 / (ascale[L] *
 * wscale[L]). wscale = 448 / amax |W_L| of the current weights; ascale =
 * 448 / amax |layer input| seen by the last training forward pass (1 before
 * any training). Diagnostics for checking the FP8 path. */
int aqua_nrc_fp8_scales(aqua_nrc* nrc, float* wscale, float* ascale);
/* The order the cache evaluates n records in (aqua_nrc_infer, and the rows of
 * a training step): order[j] = the record evaluated j-th, the records sorted
 * stably by the Morton code of their position (10 bits per axis of the
 * position clamped to [0, 1], x lowest) with its low 6 bits dropped, i.e. by
 * cells of 1/256 of the unit cube per axis, ties in record order. No
 * counterpart in the reference (whose renderer source is empty);
 * diagnostics for checking the device sort. */
int aqua_nrc_query_order(aqua_nrc* nrc, uint32_t n, const float* records, uint32_t* order);
/* restore == 0: saves the cache's whole training state on the device
 * (parameters, Adam moments, step count, FP8 activation ranges); restore ==
 * 1: puts the saved state back, so later queries and training steps run
 * exactly as they would have from the save. One checkpoint per cache. */
int aqua_nrc_checkpoint(aqua_nrc* nrc, int restore);
/* Cache query: rgb[3i..3i+2] for each record (host arrays). */
int aqua_nrc_infer(aqua_nrc* nrc, uint32_t n, const float* records, uint32_t precision, float* rgb);
/* `steps` Adam steps on one batch (host arrays; targets 3 floats per row). */
int aqua_nrc_train(aqua_nrc* nrc, uint32_t n, const float* records, const float* targets, uint32_t steps,
                   aqua_nrc_stats* stats);
/* Diagnostics: gradients of the relative L2 loss (mean over rows and the 3
 * channels) with respect to every parameter, from the same kernels a
 * training step runs, without updating anything. Also returns the loss. */
int aqua_nrc_gradients(aqua_nrc* nrc, uint32_t n, const float* records, const float* targets,
                       float* g_table, float* g_mlp, float* loss);
/* Same on device-resident buffers: records 4 float4 per row, targets one
 * float4 per row (rgb, -), both on the NRC's device. */
int aqua_nrc_train_device(aqua_nrc* nrc, uint32_t n, const void* d_records, const void* d_targets,
                          uint32_t steps, aqua_nrc_stats* stats);

/* The cache of an "nrc" renderer (NULL for "pt"); owned by the renderer. */
aqua_nrc* aqua_renderer_nrc(aqua_renderer* r);
/* 0 disables the cache (frames render as the plain path tracer, which is
 * also the fallback when training diverges); 1 re-enables it. */
int aqua_renderer_set_nrc_enabled(aqua_renderer* r, int enabled);
/* The precision the cache's queries run in from the next frame on:
 * AQUA_NRC_F16, AQUA_NRC_BF16 or AQUA_NRC_FP8, from the same weights (the
 * training precision, fixed when the renderer was created, does not change).
 * With aqua_renderer_nrc_checkpoint it renders one cache state in two
 * precisions. */
int aqua_renderer_set_nrc_precision(aqua_renderer* r, uint32_t precision);
/* aqua_nrc_checkpoint of the renderer's cache, plus the renderer's own cache
 * schedule (training steps taken, cache on / off), between frames. */
int aqua_renderer_nrc_checkpoint(aqua_renderer* r, int restore);
/* The last frame's training batch, device-resident on the renderer's device:
 * records (4 float4 per row) and targets (float4 per row). Valid until the
 * next frame. */
int aqua_renderer_nrc_batch(aqua_renderer* r, uint32_t* n, const void** d_records, const void** d_targets);

/* ---- misc ---------------------------------------------------------------- */

/* Camera of frame `frame` of an `n_frames` fly-through: a closed Catmull-Rom
 * spline through 8 seeded control points on a jittered orbit around the
 * centre of the scene bounds (25-40 % of the half extent out, 70-80 % of the
 * way up), looking at the bounds' centre; fov and resolution as given. */
int aqua_camera_flythrough(const aqua_scene* scene, uint32_t seed, uint32_t frame, uint32_t n_frames,
                           float fov_deg, uint32_t width, uint32_t height, aqua_camera* out);

const char* aqua_last_error(void);
int aqua_abi_version(void);
/* Number of HIP devices visible (0 on a host without a GPU). */
int aqua_device_count(int* count);

/* Diagnostics: the BVH the renderer builds for `scene`, as exact 4-wide
 * nodes of 32 floats (128 bytes) -- lo.x[4], hi.x[4], lo.y[4], hi.y[4],
 * lo.z[4], hi.z[4], then int32 index[4] and count[4]; count > 0 = leaf over
 * triangle slots [index, index + count), count == 0 = inner node `index`,
 * index < 0 = empty. (On the device each node is stored in 64 bytes with its
 * child boxes quantised outward to 8 bits per plane.) Call with nodes == NULL to get *n_nodes, then
 * again with room for that many; order[slot] = triangle id; max_depth =
 * depth of the deepest inner node. */
int aqua_scene_bvh(const aqua_scene* scene, uint32_t* n_nodes, float* nodes, uint32_t* order,
                   uint32_t* max_depth);
/* Diagnostics: the same tree as the renderer stores it on the device, 16
 * words (64 bytes) per node: words 0-2 the node origin (float), word 3 the
 * exponents e_x | e_y << 8 | e_z << 16 of the per-axis power-of-two scales as
 * signed bytes, words 4-9 the child planes lo.x, hi.x, lo.y, hi.y, lo.z, hi.z
 * as one byte per child (byte k = child k; plane = origin + q * 2^e, rounded
 * outward), words 10-13 the child references (an inner child: its node
 * index, below 2^31; a leaf child: 0x80000000 | the mask of its triangle
 * slots over the node's leaf run, bit i = slot word14 + i; an empty child:
 * 0x80000000, a leaf of no triangles), word 14 the first triangle slot of
 * the node's leaf children (which take one run of slots), word 15 zero. Same
 * calling pattern as above. */
int aqua_scene_bvh_device(const aqua_scene* scene, uint32_t* n_nodes, uint32_t* words);

/* Diagnostics: the device's FP32 arithmetic on n inputs, on the current
 * device: out[5i..5i+4] = a/b, sqrt(a), fma(a, b, c), sin(2 pi u), cos(2 pi u),
 * where sin/cos are the kernels' own sincos_2pi. */
int aqua_fp_probe(uint32_t n, const float* a, const float* b, const float* c, const float* u,
                  float* out);

#ifdef __cplusplus
}
#endif

#endif /* AQUA_AQUA_H */
