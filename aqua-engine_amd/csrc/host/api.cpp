// C ABI entry points that need no GPU: scenes, procedural scenes, the
// integrator config and error reporting.
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
      if (order) memcpy(order, b.tri_order.data(), b.tri_order.size() * sizeof(uint32_t));
    }
    *n_nodes = uint32_t(b4.nodes.size());
    if (max_depth) *max_depth = b4.max_inner_depth;
    return AQUA_OK;
  })
}

int aqua_procgen_write(const aqua_procgen_params* p, const char* out_dir, char* scene_path, size_t cap) {
  AQUA_GUARD(return aqua::procgen_write(p, out_dir, scene_path, cap);)
}

void aqua_integrator_default(aqua_integrator_cfg* out) {
  if (out) aqua::integrator_default(out);
}

int aqua_integrator_load(const char* path, aqua_integrator_cfg* out) {
  AQUA_GUARD(return aqua::integrator_load(path, out);)
}

}  // extern "C"
