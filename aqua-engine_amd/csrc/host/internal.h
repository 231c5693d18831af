// Host-side internals shared by the loader, the procedural generators, the
// BVH builder and the C ABI. Nothing here crosses the ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "aqua/aqua.h"

// A decoded image texture: 8-bit sRGB-encoded RGBA, row 0 at the top.
struct AquaImage {
  uint32_t w = 0, h = 0;
  std::vector<uint8_t> rgba;
};

struct aqua_scene {
  std::vector<float> verts;        // 9 floats per triangle, load order
  std::vector<uint32_t> mat_ids;   // one per triangle
  std::vector<aqua_material> materials;
  std::vector<std::string> material_names;
  aqua_camera camera{};
  std::vector<aqua_point_light> point_lights;
  std::vector<float> uvs;          // 6 floats per triangle (u, v per corner), or empty
  std::vector<AquaImage> images;   // textures named by aqua_material::base_color_tex
  aqua_scene_info info{};
};

namespace aqua {

// JPEG (T.81 Huffman, baseline or progressive) -> 8-bit RGBA.
int decode_jpeg(const std::vector<uint8_t>& data, uint32_t* w, uint32_t* h, std::vector<uint8_t>* rgba);
// Reads and decodes an image file (JPEG).
int read_image_file(const std::string& path, AquaImage* out);

// Thread-local last error; returns `code` so call sites can `return fail(...)`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
const char* last_error_cstr();

// Fills info (counts, bounds) from the arrays of a scene.
void finalize_scene_info(aqua_scene* s);

// IEC 61966-2-1 sRGB decode of one channel, in double.
double srgb_to_linear(double c);

// Camera matrix from the reference's TRS transform (cbox.json:524-541):
// rotate = Euler degrees applied X, then Y, then Z (R = Rz Ry Rx); columns of
// the linear part are the camera's right, up and back axes.
void camera_from_trs(const double translate[3], const double rotate_deg[3], const double scale[3],
                     double fov_deg, uint32_t w, uint32_t h, aqua_camera* out);

// ---- BSON (.mesh) --------------------------------------------------------
struct MeshData {
  std::string name;
  std::vector<double> vertices;   // 3 per vertex
  std::vector<double> normals;    // 3 per vertex (may be empty)
  std::vector<double> texcoords;  // 2 per vertex (may be empty)
  std::vector<int64_t> indices;   // 3 per triangle
};
int read_mesh_bson(const std::string& path, MeshData* out);   // AQUA_OK or error
int write_mesh_bson(const std::string& path, const MeshData& m);

// ---- light table (area lights = emissive triangles, then point lights) ---
// Selection is proportional to emitted power: pi * area * lum(Le) for a
// one-sided Lambertian emitter, 4 pi * lum(I) for a point light.
struct LightTable {
  std::vector<uint32_t> tri;      // emissive triangle ids (load order)
  std::vector<float> area;        // per emissive triangle
  std::vector<float> pmf;         // per light, area lights first
  std::vector<float> cdf;         // inclusive prefix of pmf, last == 1
  std::vector<float> mat_pdf;     // per material: pi * lum(Le) / total power
  double total_power = 0.0;
};
void build_light_table(const aqua_scene* s, LightTable* out);
double luminance(double r, double g, double b);

}  // namespace aqua
