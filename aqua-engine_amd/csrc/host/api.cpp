// C ABI entry points that need no GPU: scenes, procedural scenes, the
// integrator config and error reporting.
#include <cmath>
#include <cstring>
#include <new>

#include "bvh.h"
#include "internal.h"

namespace aqua {
int load_scene(const char* path, const aqua_load_opts* opts, aqua_scene** out);
int scene_from_arrays(uint32_t n_tris, const float* verts, const uint32_t* mat_ids, uint32_t n_mats,
                      const aqua_material* mats, const aqua_camera* cam, uint32_t n_pl,
                      const aqua_point_light* pl, aqua_scene** out);
int procgen_write(const aqua_procgen_params* p, const char* out_dir, char* scene_path, size_t cap);
void integrator_default(aqua_integrator_cfg* c);
int integrator_load(const char* path, aqua_integrator_cfg* out);
const char* last_error_cstr();
}  // namespace aqua

// Every entry point catches allocation failure so no exception crosses the ABI.
#define AQUA_GUARD(body)                                                  \
  try {                                                                   \
    body                                                                  \
  } catch (const std::bad_alloc&) {                                       \
    return aqua::fail(AQUA_ERR_OOM, "host allocation failed");            \
  } catch (...) {                                                         \
    return aqua::fail(AQUA_ERR_INVALID_ARG, "unexpected internal error"); \
  }

extern "C" {

const char* aqua_last_error(void) { return aqua::last_error_cstr(); }
int aqua_abi_version(void) { return AQUA_ABI_VERSION; }

int aqua_scene_load(const char* path, const aqua_load_opts* opts, aqua_scene** out) {
  if (!path || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_scene_load: null argument");
  *out = nullptr;
  AQUA_GUARD(return aqua::load_scene(path, opts, out);)
}

int aqua_scene_from_arrays(uint32_t n_triangles, const float* verts, const uint32_t* mat_ids,
                           uint32_t n_materials, const aqua_material* materials, const aqua_camera* camera,
                           uint32_t n_point_lights, const aqua_point_light* point_lights, aqua_scene** out) {
  AQUA_GUARD(return aqua::scene_from_arrays(n_triangles, verts, mat_ids, n_materials, materials, camera,
                                            n_point_lights, point_lights, out);)
}

int aqua_scene_get_info(const aqua_scene* s, aqua_scene_info* out) {
  if (!s || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  *out = s->info;
  return AQUA_OK;
}

int aqua_scene_get_triangles(const aqua_scene* s, float* verts, uint32_t* mat_ids) {
  if (!s) return aqua::fail(AQUA_ERR_INVALID_ARG, "null scene");
  if (verts && !s->verts.empty()) memcpy(verts, s->verts.data(), s->verts.size() * sizeof(float));
  if (mat_ids && !s->mat_ids.empty()) memcpy(mat_ids, s->mat_ids.data(), s->mat_ids.size() * sizeof(uint32_t));
  return AQUA_OK;
}

int aqua_scene_get_materials(const aqua_scene* s, aqua_material* out) {
  if (!s || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  memcpy(out, s->materials.data(), s->materials.size() * sizeof(aqua_material));
  return AQUA_OK;
}

int aqua_scene_get_camera(const aqua_scene* s, aqua_camera* out) {
  if (!s || !out) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  *out = s->camera;
  return AQUA_OK;
}

int aqua_scene_get_point_lights(const aqua_scene* s, aqua_point_light* out) {
  if (!s || (!out && !s->point_lights.empty())) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (!s->point_lights.empty())
    memcpy(out, s->point_lights.data(), s->point_lights.size() * sizeof(aqua_point_light));
  return AQUA_OK;
}

int aqua_scene_get_texcoords(const aqua_scene* s, float* uv) {
  if (!s || !uv) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (s->uvs.empty()) return aqua::fail(AQUA_ERR_INVALID_ARG, "the scene has no texcoords");
  memcpy(uv, s->uvs.data(), s->uvs.size() * sizeof(float));
  return AQUA_OK;
}

int aqua_scene_get_image(const aqua_scene* s, uint32_t k, uint32_t* width, uint32_t* height, uint8_t* rgba) {
  if (!s || !width || !height) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (k >= s->images.size()) return aqua::fail(AQUA_ERR_INVALID_ARG, "image %u of %zu", k, s->images.size());
  const AquaImage& im = s->images[k];
  *width = im.w;
  *height = im.h;
  if (rgba) memcpy(rgba, im.rgba.data(), im.rgba.size());
  return AQUA_OK;
}

int aqua_scene_set_texcoords(aqua_scene* s, const float* uv) {
  if (!s || !uv) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  const size_t n = 6 * s->mat_ids.size();
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(uv[i])) return aqua::fail(AQUA_ERR_NONFINITE, "texcoord %zu is not finite", i);
  AQUA_GUARD(s->uvs.assign(uv, uv + n); aqua::finalize_scene_info(s); return AQUA_OK;)
}

int aqua_scene_add_image(aqua_scene* s, uint32_t width, uint32_t height, const uint8_t* rgba, uint32_t* index) {
  if (!s || !rgba || !index) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (width == 0 || height == 0 || width > 32768 || height > 32768)
    return aqua::fail(AQUA_ERR_INVALID_ARG, "image size %ux%u out of range", width, height);
  AQUA_GUARD({
    AquaImage im;
    im.w = width;
    im.h = height;
    im.rgba.assign(rgba, rgba + size_t(width) * height * 4);
    *index = uint32_t(s->images.size());
    s->images.push_back(std::move(im));
    aqua::finalize_scene_info(s);
    return AQUA_OK;
  })
}

void aqua_scene_destroy(aqua_scene* s) { delete s; }

int aqua_scene_bvh(const aqua_scene* s, uint32_t* n_nodes, float* nodes, uint32_t* order, uint32_t* max_depth) {
  if (!s || !n_nodes) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  AQUA_GUARD({
    aqua::Bvh b;
    aqua::BvhBuildOpts o;
    aqua::build_bvh(s->verts.data(), uint32_t(s->mat_ids.size()), o, &b);
    aqua::Bvh4 b4;
    aqua::collapse_bvh4(b, &b4);
    if (nodes) {
      if (*n_nodes < b4.nodes.size()) return aqua::fail(AQUA_ERR_INVALID_ARG, "node buffer too small");
      memcpy(nodes, b4.nodes.data(), b4.nodes.size() * sizeof(aqua::Bvh4Node));
      if (order) memcpy(order, b4.tri_order.data(), b4.tri_order.size() * sizeof(uint32_t));
    }
    *n_nodes = uint32_t(b4.nodes.size());
    if (max_depth) *max_depth = b4.max_inner_depth;
    return AQUA_OK;
  })
}

int aqua_scene_bvh_device(const aqua_scene* s, uint32_t* n_nodes, uint32_t* words) {
  if (!s || !n_nodes) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  AQUA_GUARD({
    aqua::Bvh b;
    aqua::BvhBuildOpts o;
    aqua::build_bvh(s->verts.data(), uint32_t(s->mat_ids.size()), o, &b);
    aqua::Bvh4 b4;
    aqua::collapse_bvh4(b, &b4);
    if (words) {
      if (*n_nodes < b4.nodes.size()) return aqua::fail(AQUA_ERR_INVALID_ARG, "node buffer too small");
      std::vector<uint32_t> q;
      if (!aqua::quantize_bvh4(b4, &q)) return aqua::fail(AQUA_ERR_UNSUPPORTED, "BVH4 leaf runs do not fit the node format");
      memcpy(words, q.data(), q.size() * sizeof(uint32_t));
    }
    *n_nodes = uint32_t(b4.nodes.size());
    return AQUA_OK;
  })
}

int aqua_procgen_write(const aqua_procgen_params* p, const char* out_dir, char* scene_path, size_t cap) {
  AQUA_GUARD(return aqua::procgen_write(p, out_dir, scene_path, cap);)
}

int aqua_camera_flythrough(const aqua_scene* s, uint32_t seed, uint32_t frame, uint32_t n_frames, float fov_deg,
                           uint32_t width, uint32_t height, aqua_camera* out) {
  if (!s || !out || n_frames == 0 || width == 0 || height == 0 || width > 16384 || height > 16384 ||
      !(fov_deg > 0.0f && fov_deg < 180.0f))
    return aqua::fail(AQUA_ERR_INVALID_ARG, "aqua_camera_flythrough: bad argument");
  // 8 control points (PCG32, stream 0x666c79) on a jittered orbit around the
  // centre of the bounds: angle 2 pi i / 8 +- 0.3 rad, radius 25-40 % of the
  // half extent in x and z, height 70-80 % up the bounds (above floor-standing
  // objects, below a ceiling light); closed uniform Catmull-Rom spline through
  // them; the camera looks at the centre of the bounds with +y up.
  const aqua_scene_info& in = s->info;
  double c[3], he[3];
  for (int k = 0; k < 3; ++k) {
    double a = in.bounds_lo[k], b = in.bounds_hi[k];
    c[k] = 0.5 * (a + b);
    he[k] = 0.5 * (b - a);
  }
  uint64_t state = 0, inc = (0x666c79ull << 1u) | 1u;
  auto next = [&]() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u), rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  };
  auto u01 = [&]() { return double(next() >> 8) / 16777216.0; };
  next();
  state += seed;
  next();
  double P[8][3];
  for (int i = 0; i < 8; ++i) {
    double ang = 2.0 * M_PI * i / 8.0 + 0.6 * (u01() - 0.5);
    double rad = 0.25 + 0.15 * u01();
    double hy = 0.7 + 0.1 * u01();
    P[i][0] = c[0] + rad * he[0] * std::cos(ang);
    P[i][1] = in.bounds_lo[1] + hy * 2.0 * he[1];
    P[i][2] = c[2] + rad * he[2] * std::sin(ang);
  }
  double u = double(frame % n_frames) / double(n_frames) * 8.0;
  int seg = int(u);
  double t = u - seg;
  const double* p0 = P[(seg + 7) % 8];
  const double* p1 = P[seg % 8];
  const double* p2 = P[(seg + 1) % 8];
  const double* p3 = P[(seg + 2) % 8];
  double pos[3];
  for (int k = 0; k < 3; ++k) {
    double t2 = t * t, t3 = t2 * t;
    pos[k] = 0.5 * ((2 * p1[k]) + (-p0[k] + p2[k]) * t + (2 * p0[k] - 5 * p1[k] + 4 * p2[k] - p3[k]) * t2 +
                    (-p0[k] + 3 * p1[k] - 3 * p2[k] + p3[k]) * t3);
  }
  double f[3] = {c[0] - pos[0], c[1] - pos[1], c[2] - pos[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  if (!(fl > 1e-9)) f[0] = 0, f[1] = 0, f[2] = -1, fl = 1;
  for (double& x : f) x /= fl;
  double up[3] = {0, 1, 0};
  // right = f x up, true up = right x f; camera looks down -Z: back = -f
  double rgt[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double rl = std::sqrt(rgt[0] * rgt[0] + rgt[1] * rgt[1] + rgt[2] * rgt[2]);
  if (!(rl > 1e-9)) rgt[0] = 1, rgt[1] = 0, rgt[2] = 0, rl = 1;
  for (double& x : rgt) x /= rl;
  double u2[3] = {rgt[1] * f[2] - rgt[2] * f[1], rgt[2] * f[0] - rgt[0] * f[2], rgt[0] * f[1] - rgt[1] * f[0]};
  memset(out, 0, sizeof *out);
  for (int k = 0; k < 3; ++k) {
    out->cam_to_world[k * 4 + 0] = float(rgt[k]);
    out->cam_to_world[k * 4 + 1] = float(u2[k]);
    out->cam_to_world[k * 4 + 2] = float(-f[k]);
    out->cam_to_world[k * 4 + 3] = float(pos[k]);
  }
  out->tan_half_fov = float(std::tan(0.5 * double(fov_deg) * M_PI / 180.0));
  out->width = width;
  out->height = height;
  out->aspect = float(width) / float(height);
  return AQUA_OK;
}

void aqua_integrator_default(aqua_integrator_cfg* out) {
  if (out) aqua::integrator_default(out);
}

int aqua_integrator_load(const char* path, aqua_integrator_cfg* out) {
  AQUA_GUARD(return aqua::integrator_load(path, out);)
}

}  // extern "C"
