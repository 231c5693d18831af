// Scene JSON loader: the reference's scene schema (scenes/cbox.json:2-627,
// scenes/room.json:2-3899) flattened to FP32 triangles, a material table, a
// camera and point lights.
//
//   { "named_bsdfs": { <name>: { "Principled": { 17 params as <Tex>, "hint" } } },
//     "camera": { "Perspective": { "res", "fov", "lens_radius", "focal", "transform" } },
//     "lights": [ { "Point": { "pos", "emission" } } ],
//     "shapes": [ { "Mesh": [ "<file>.mesh", { "Named": "<bsdf>" } ] } ] }
//   <Tex> := {"Float": f} | {"Float3": [3]} | {"Srgb": [3]} | {"Image": path}
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "internal.h"
#include "json.h"

namespace aqua {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
void clear_error() { g_last_error.clear(); }
const char* last_error_cstr() { return g_last_error.c_str(); }

double srgb_to_linear(double c) {
  // IEC 61966-2-1: linear = c / 12.92 below the 0.04045 knee, else
  // ((c + 0.055) / 1.055)^2.4.
  if (c <= 0.04045) return c / 12.92;
  return std::pow((c + 0.055) / 1.055, 2.4);
}

double luminance(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

void camera_from_trs(const double tr[3], const double rot_deg[3], const double sc[3],
                     double fov_deg, uint32_t w, uint32_t h, aqua_camera* out) {
  const double d2r = M_PI / 180.0;
  double cx = std::cos(rot_deg[0] * d2r), sx = std::sin(rot_deg[0] * d2r);
  double cy = std::cos(rot_deg[1] * d2r), sy = std::sin(rot_deg[1] * d2r);
  double cz = std::cos(rot_deg[2] * d2r), sz = std::sin(rot_deg[2] * d2r);
  double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
  double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
  double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
  double T[3][3], R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T[i][j] = 0;
      for (int k = 0; k < 3; ++k) T[i][j] += Ry[i][k] * Rx[k][j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R[i][j] = 0;
      for (int k = 0; k < 3; ++k) R[i][j] += Rz[i][k] * T[k][j];
    }
  memset(out, 0, sizeof *out);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) out->cam_to_world[i * 4 + j] = float(R[i][j] * sc[j]);
    out->cam_to_world[i * 4 + 3] = float(tr[i]);
  }
  out->tan_half_fov = float(std::tan(0.5 * fov_deg * d2r));
  out->width = w;
  out->height = h;
  out->aspect = float(w) / float(h);
}

void finalize_scene_info(aqua_scene* s) {
  aqua_scene_info& in = s->info;
  in.n_triangles = uint32_t(s->mat_ids.size());
  in.n_materials = uint32_t(s->materials.size());
  in.n_point_lights = uint32_t(s->point_lights.size());
  uint32_t ne = 0;
  for (uint32_t m : s->mat_ids) {
    const aqua_material& mt = s->materials[m];
    if (mt.emission[0] > 0.f || mt.emission[1] > 0.f || mt.emission[2] > 0.f) ++ne;
  }
  in.n_emissive_triangles = ne;
  in.n_images = uint32_t(s->images.size());
  in.has_texcoords = s->uvs.empty() ? 0u : 1u;
  for (int k = 0; k < 3; ++k) {
    in.bounds_lo[k] = INFINITY;
    in.bounds_hi[k] = -INFINITY;
  }
  for (size_t i = 0; i < s->verts.size(); i += 3)
    for (int k = 0; k < 3; ++k) {
      float v = s->verts[i + size_t(k)];
      if (v < in.bounds_lo[k]) in.bounds_lo[k] = v;
      if (v > in.bounds_hi[k]) in.bounds_hi[k] = v;
    }
}

void build_light_table(const aqua_scene* s, LightTable* out) {
  *out = LightTable();
  std::vector<double> power;
  size_t n = s->mat_ids.size();
  for (size_t t = 0; t < n; ++t) {
    const aqua_material& m = s->materials[s->mat_ids[t]];
    double lum = luminance(m.emission[0], m.emission[1], m.emission[2]);
    if (!(m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f)) continue;
    const float* v = &s->verts[9 * t];
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = double(v[3 + k]) - double(v[k]);
      e2[k] = double(v[6 + k]) - double(v[k]);
    }
    double c0 = e1[1] * e2[2] - e1[2] * e2[1];
    double c1 = e1[2] * e2[0] - e1[0] * e2[2];
    double c2 = e1[0] * e2[1] - e1[1] * e2[0];
    double area = 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (!(area > 0.0) || !(lum > 0.0)) continue;  // degenerate emitters carry no power
    out->tri.push_back(uint32_t(t));
    out->area.push_back(float(area));
    power.push_back(M_PI * area * lum);
  }
  for (const aqua_point_light& p : s->point_lights) {
    double lum = luminance(p.intensity[0], p.intensity[1], p.intensity[2]);
    power.push_back(4.0 * M_PI * (lum > 0.0 ? lum : 0.0));
  }
  double total = 0.0;
  for (double p : power) total += p;
  out->total_power = total;
  out->mat_pdf.assign(s->materials.size(), 0.f);
  if (!(total > 0.0)) {
    out->pmf.assign(power.size(), 0.f);
    out->cdf.assign(power.size(), 0.f);
    return;
  }
  double acc = 0.0;
  for (size_t i = 0; i < power.size(); ++i) {
    out->pmf.push_back(float(power[i] / total));
    acc += power[i];
    out->cdf.push_back(i + 1 == power.size() ? 1.0f : float(acc / total));
  }
  for (size_t m = 0; m < s->materials.size(); ++m) {
    const aqua_material& mt = s->materials[m];
    double lum = luminance(mt.emission[0], mt.emission[1], mt.emission[2]);
    out->mat_pdf[m] = lum > 0.0 ? float(M_PI * lum / total) : 0.f;
  }
}

namespace {

bool read_text(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

std::string slashes(std::string p) {
  for (char& c : p)
    if (c == '\\') c = '/';
  return p;
}

struct Loader {
  const aqua_load_opts& opts;
  std::string base;                           // directory of the scene JSON
  std::vector<AquaImage>* images = nullptr;   // the scene's decoded images
  std::map<std::string, uint32_t> image_of;   // path -> index (shared images load once)
  uint32_t n_images = 0;

  // Loads (once per path) the image `rel` names; *index = its position.
  int image(const std::string& rel, uint32_t* index) {
    std::string file = slashes(rel);
    std::string full = (!file.empty() && file[0] == '/') ? file : base + "/" + file;
    auto it = image_of.find(full);
    if (it != image_of.end()) {
      *index = it->second;
      return AQUA_OK;
    }
    AquaImage img;
    if (int e = read_image_file(full, &img)) return e;
    *index = uint32_t(images->size());
    image_of[full] = *index;
    images->push_back(std::move(img));
    return AQUA_OK;
  }

  int num(const Json* j, const char* what, double* v) {
    if (!j || !j->is_num()) return fail(AQUA_ERR_PARSE, "scene: '%s' must be a number", what);
    *v = j->num;
    return AQUA_OK;
  }
  int vec3(const Json* j, const char* what, double v[3]) {
    if (!j || j->kind != Json::Array || j->arr.size() != 3)
      return fail(AQUA_ERR_PARSE, "scene: '%s' must be an array of 3 numbers", what);
    for (int k = 0; k < 3; ++k) {
      if (!j->arr[k].is_num()) return fail(AQUA_ERR_PARSE, "scene: '%s' must hold numbers", what);
      v[k] = j->arr[k].num;
    }
    return AQUA_OK;
  }
  // <Tex> -> linear RGB. `tex` (colours only) receives image index + 1 for an
  // Image texture, whose mean linear colour goes to `out`.
  int tex_rgb(const Json* j, const char* what, double out[3], uint32_t* tex = nullptr) {
    if (!j || j->kind != Json::Object || j->obj.size() != 1)
      return fail(AQUA_ERR_PARSE, "scene: texture '%s' must be a one-key object", what);
    const std::string& tag = j->obj[0].first;
    const Json& v = j->obj[0].second;
    if (tag == "Float") {
      double f = 0.0;
      if (int e = num(&v, what, &f)) return e;
      out[0] = out[1] = out[2] = f;
    } else if (tag == "Float3") {
      if (int e = vec3(&v, what, out)) return e;
    } else if (tag == "Srgb") {
      if (int e = vec3(&v, what, out)) return e;
      for (int k = 0; k < 3; ++k) out[k] = srgb_to_linear(out[k]);
    } else if (tag == "Image") {
      if (v.kind != Json::String) return fail(AQUA_ERR_PARSE, "scene: Image '%s' must name a file", what);
      ++n_images;
      if (opts.image_as_constant) {
        out[0] = out[1] = out[2] = 0.5;
        return AQUA_OK;
      }
      if (!tex)
        return fail(AQUA_ERR_UNSUPPORTED, "scene: '%s' is an Image; only colours may be image textures", what);
      uint32_t k = 0;
      if (int e = image(v.str, &k)) return e;
      *tex = k + 1;
      const AquaImage& im = (*images)[k];
      double lut[256], sum[3] = {0, 0, 0};
      for (int i = 0; i < 256; ++i) lut[i] = srgb_to_linear(i / 255.0);
      for (size_t p = 0; p < size_t(im.w) * im.h; ++p)
        for (int c = 0; c < 3; ++c) sum[c] += lut[im.rgba[4 * p + size_t(c)]];
      for (int c = 0; c < 3; ++c) out[c] = sum[c] / double(size_t(im.w) * im.h);
    } else {
      return fail(AQUA_ERR_PARSE, "scene: unknown texture tag '%s' in '%s'", tag.c_str(), what);
    }
    return AQUA_OK;
  }
  int tex_scalar(const Json* j, const char* what, double* out) {
    double rgb[3];
    if (int e = tex_rgb(j, what, rgb)) return e;
    if (!(rgb[0] == rgb[1] && rgb[1] == rgb[2]))
      return fail(AQUA_ERR_UNSUPPORTED, "scene: '%s' must be a scalar texture", what);
    *out = rgb[0];
    return AQUA_OK;
  }

  int bsdf(const std::string& name, const Json& j, aqua_material* m) {
    const Json* p = j.get("Principled");
    if (!p || p->kind != Json::Object || j.obj.size() != 1)
      return fail(AQUA_ERR_PARSE, "scene: bsdf '%s' must be {\"Principled\": {...}}", name.c_str());
    memset(m, 0, sizeof *m);
    double rgb[3];
    if (int e = tex_rgb(p->get("color"), "color", rgb, &m->base_color_tex)) return e;
    for (int k = 0; k < 3; ++k) m->base_color[k] = float(rgb[k]);
    double v = 0.0;
    if (int e = tex_scalar(p->get("metallic"), "metallic", &v)) return e;
    m->metallic = float(v);
    if (int e = tex_scalar(p->get("roughness"), "roughness", &v)) return e;
    m->roughness = float(v);
    if (int e = tex_scalar(p->get("specular"), "specular", &v)) return e;
    m->specular = float(v);
    if (p->get("emission")) {
      if (int e = tex_rgb(p->get("emission"), "emission", rgb)) return e;
      for (int k = 0; k < 3; ++k) m->emission[k] = float(rgb[k] > 0.0 ? rgb[k] : 0.0);
    }
    // Lobes the metallic-roughness BSDF does not model must be off (the
    // shipped scenes hold 0 for all of them: cbox.json:14,20-22,41-52,59).
    static const char* off[] = {"subsurface", "anisotropic", "sheen", "clearcoat", "transmission"};
    for (const char* k : off) {
      const Json* t = p->get(k);
      if (!t) continue;
      double x = 0.0;
      if (int e = tex_scalar(t, k, &x)) return e;
      if (x != 0.0)
        return fail(AQUA_ERR_UNSUPPORTED, "scene: bsdf '%s' sets '%s' = %g; only 0 is supported",
                    name.c_str(), k, x);
    }
    for (float c : m->base_color)
      if (!std::isfinite(c)) return fail(AQUA_ERR_NONFINITE, "scene: bsdf '%s' colour", name.c_str());
    if (!std::isfinite(m->metallic) || !std::isfinite(m->roughness) || !std::isfinite(m->specular))
      return fail(AQUA_ERR_NONFINITE, "scene: bsdf '%s' parameter", name.c_str());
    return AQUA_OK;
  }
};

std::string dir_of(const std::string& path) {
  size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

// `Ke` (emitted radiance, linear RGB) of every `newmtl` in the .mtl files of
// directory `dir` (Wavefront MTL: `newmtl <name>` opens a material, `Ke r g
// b` sets its emission; living_room.mtl:11-13 BackLight Ke 2 2 2).
int read_mtl_emission(const std::string& dir, std::map<std::string, std::array<double, 3>>* out) {
  std::error_code ec;
  std::vector<std::string> files;
  for (const auto& de : std::filesystem::directory_iterator(dir, ec))
    if (de.path().extension() == ".mtl") files.push_back(de.path().string());
  if (ec) return fail(AQUA_ERR_IO, "cannot list '%s' for .mtl files", dir.c_str());
  std::sort(files.begin(), files.end());
  for (const std::string& f : files) {
    std::string text;
    if (!read_text(f, &text)) return fail(AQUA_ERR_IO, "cannot read '%s'", f.c_str());
    std::istringstream in(text);
    std::string line, cur;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string key;
      ls >> key;
      if (key == "newmtl") {
        ls >> cur;
      } else if (key == "Ke" && !cur.empty()) {
        std::array<double, 3> k{0, 0, 0};
        ls >> k[0] >> k[1] >> k[2];
        if (!ls) return fail(AQUA_ERR_PARSE, "%s: bad Ke line for '%s'", f.c_str(), cur.c_str());
        (*out)[cur] = k;
      }
    }
  }
  return AQUA_OK;
}

}  // namespace

int load_scene(const char* path, const aqua_load_opts* opts_in, aqua_scene** out) {
  aqua_load_opts opts{};
  if (opts_in) opts = *opts_in;
  std::string text;
  if (!read_text(path, &text)) return fail(AQUA_ERR_IO, "cannot open scene '%s'", path);
  Json root;
  std::string perr;
  if (!json_parse(text, &root, &perr)) return fail(AQUA_ERR_PARSE, "%s: %s", path, perr.c_str());
  if (root.kind != Json::Object) return fail(AQUA_ERR_PARSE, "%s: top level must be an object", path);

  auto s = std::make_unique<aqua_scene>();
  std::string base = dir_of(path);
  Loader L{opts, base, &s->images, {}, 0};
  std::map<std::string, uint32_t> by_name;

  const Json* bsdfs = root.get("named_bsdfs");
  if (!bsdfs || bsdfs->kind != Json::Object)
    return fail(AQUA_ERR_PARSE, "%s: 'named_bsdfs' must be an object", path);
  for (const auto& kv : bsdfs->obj) {
    aqua_material m;
    if (int e = L.bsdf(kv.first, kv.second, &m)) return e;
    by_name[kv.first] = uint32_t(s->materials.size());
    s->materials.push_back(m);
    s->material_names.push_back(kv.first);
  }
  if (opts.emission_from_mtl) {
    std::map<std::string, std::array<double, 3>> ke;
    if (int e = read_mtl_emission(base, &ke)) return e;
    for (size_t m = 0; m < s->materials.size(); ++m) {
      aqua_material& mt = s->materials[m];
      auto it = ke.find(s->material_names[m]);
      if (it == ke.end() || mt.emission[0] > 0.f || mt.emission[1] > 0.f || mt.emission[2] > 0.f) continue;
      for (int k = 0; k < 3; ++k) mt.emission[k] = float(it->second[size_t(k)] > 0.0 ? it->second[size_t(k)] : 0.0);
    }
  }

  const Json* cam = root.get("camera");
  const Json* persp = cam ? cam->get("Perspective") : nullptr;
  if (!persp) return fail(AQUA_ERR_PARSE, "%s: 'camera' must be {\"Perspective\": {...}}", path);
  {
    const Json* res = persp->get("res");
    if (!res || res->kind != Json::Array || res->arr.size() != 2 || !res->arr[0].is_num() ||
        !res->arr[1].is_num())
      return fail(AQUA_ERR_PARSE, "%s: camera 'res' must be [w, h]", path);
    double w = res->arr[0].num, h = res->arr[1].num;
    if (!(w >= 1 && h >= 1 && w <= 16384 && h <= 16384 && w == std::floor(w) && h == std::floor(h)))
      return fail(AQUA_ERR_INVALID_ARG, "%s: camera resolution out of range", path);
    double fov = 0.0, lens = 0.0;
    if (int e = L.num(persp->get("fov"), "fov", &fov)) return e;
    if (!(fov > 0.0 && fov < 180.0)) return fail(AQUA_ERR_INVALID_ARG, "%s: fov out of (0,180)", path);
    if (persp->get("lens_radius")) {
      if (int e = L.num(persp->get("lens_radius"), "lens_radius", &lens)) return e;
    }
    if (lens != 0.0)
      return fail(AQUA_ERR_UNSUPPORTED, "%s: lens_radius %g: only pinhole cameras are supported", path,
                  lens);
    double tr[3] = {0, 0, 0}, rot[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
    if (const Json* xf = persp->get("transform")) {
      if (xf->get("translate"))
        if (int e = L.vec3(xf->get("translate"), "translate", tr)) return e;
      if (xf->get("rotate"))
        if (int e = L.vec3(xf->get("rotate"), "rotate", rot)) return e;
      if (xf->get("scale"))
        if (int e = L.vec3(xf->get("scale"), "scale", sc)) return e;
    }
    camera_from_trs(tr, rot, sc, fov, uint32_t(w), uint32_t(h), &s->camera);
  }

  if (const Json* lights = root.get("lights")) {
    if (lights->kind != Json::Array) return fail(AQUA_ERR_PARSE, "%s: 'lights' must be an array", path);
    for (const Json& l : lights->arr) {
      const Json* pt = l.get("Point");
      if (!pt) {
        std::string tag = (l.kind == Json::Object && !l.obj.empty()) ? l.obj[0].first : "?";
        return fail(AQUA_ERR_UNSUPPORTED, "%s: light type '%s' is not supported", path, tag.c_str());
      }
      aqua_point_light p{};
      double pos[3], em[3];
      if (int e = L.vec3(pt->get("pos"), "pos", pos)) return e;
      if (int e = L.tex_rgb(pt->get("emission"), "emission", em)) return e;
      for (int k = 0; k < 3; ++k) {
        p.pos[k] = float(pos[k]);
        p.intensity[k] = float(em[k] > 0.0 ? em[k] : 0.0);
      }
      s->point_lights.push_back(p);
    }
  }

  const Json* shapes = root.get("shapes");
  if (!shapes || shapes->kind != Json::Array)
    return fail(AQUA_ERR_PARSE, "%s: 'shapes' must be an array", path);
  const bool want_uv = !s->images.empty();  // texcoords only matter to image textures
  uint32_t n_loaded = 0, n_skipped = 0, n_vertices = 0;
  for (const Json& sh : shapes->arr) {
    const Json* mesh = sh.get("Mesh");
    if (!mesh || mesh->kind != Json::Array || mesh->arr.size() != 2 ||
        mesh->arr[0].kind != Json::String)
      return fail(AQUA_ERR_PARSE, "%s: shape must be {\"Mesh\": [file, {\"Named\": bsdf}]}", path);
    const Json* named = mesh->arr[1].get("Named");
    if (!named || named->kind != Json::String)
      return fail(AQUA_ERR_PARSE, "%s: shape material must be {\"Named\": name}", path);
    auto it = by_name.find(named->str);
    if (it == by_name.end())
      return fail(AQUA_ERR_PARSE, "%s: unknown bsdf '%s'", path, named->str.c_str());
    std::string file = slashes(mesh->arr[0].str);
    std::string full = (!file.empty() && file[0] == '/') ? file : base + "/" + file;
    FILE* probe = fopen(full.c_str(), "rb");
    if (!probe) {
      if (opts.skip_missing) {
        ++n_skipped;
        continue;
      }
      return fail(AQUA_ERR_MISSING_MESH, "%s: mesh '%s' is missing", path, full.c_str());
    }
    fclose(probe);
    MeshData md;
    if (int e = read_mesh_bson(full, &md)) return e;
    size_t nt = md.indices.size() / 3;
    const bool has_uv = md.texcoords.size() == 2 * (md.vertices.size() / 3);
    for (size_t t = 0; t < nt; ++t) {
      for (int k = 0; k < 3; ++k) {
        size_t vi = size_t(md.indices[3 * t + size_t(k)]);
        for (int c = 0; c < 3; ++c) s->verts.push_back(float(md.vertices[3 * vi + size_t(c)]));
        if (want_uv) {
          s->uvs.push_back(has_uv ? float(md.texcoords[2 * vi]) : 0.0f);
          s->uvs.push_back(has_uv ? float(md.texcoords[2 * vi + 1]) : 0.0f);
        }
      }
      s->mat_ids.push_back(it->second);
    }
    n_vertices += uint32_t(md.vertices.size() / 3);
    ++n_loaded;
  }
  if (s->mat_ids.size() >= 0x7FFFFFFFu)
    return fail(AQUA_ERR_INVALID_ARG, "%s: too many triangles", path);
  finalize_scene_info(s.get());
  s->info.n_shapes = n_loaded;
  s->info.n_skipped_shapes = n_skipped;
  s->info.n_image_textures = L.n_images;
  s->info.n_vertices = n_vertices;
  *out = s.release();
  return AQUA_OK;
}

int scene_from_arrays(uint32_t n_tris, const float* verts, const uint32_t* mat_ids, uint32_t n_mats,
                      const aqua_material* mats, const aqua_camera* cam, uint32_t n_pl,
                      const aqua_point_light* pl, aqua_scene** out) {
  if (!out || !cam || (n_tris && (!verts || !mat_ids)) || (n_mats && !mats) || (n_pl && !pl))
    return fail(AQUA_ERR_INVALID_ARG, "aqua_scene_from_arrays: null argument");
  if (n_mats == 0) return fail(AQUA_ERR_INVALID_ARG, "aqua_scene_from_arrays: no materials");
  if (n_tris >= 0x7FFFFFFFu) return fail(AQUA_ERR_INVALID_ARG, "aqua_scene_from_arrays: too many triangles");
  if (cam->width == 0 || cam->height == 0 || cam->width > 16384 || cam->height > 16384)
    return fail(AQUA_ERR_INVALID_ARG, "aqua_scene_from_arrays: camera resolution out of range");
  for (uint32_t i = 0; i < n_tris; ++i)
    if (mat_ids[i] >= n_mats)
      return fail(AQUA_ERR_INVALID_ARG, "aqua_scene_from_arrays: material id %u out of range", mat_ids[i]);
  for (size_t i = 0; i < size_t(n_tris) * 9; ++i)
    if (!std::isfinite(verts[i])) return fail(AQUA_ERR_NONFINITE, "aqua_scene_from_arrays: vertex");
  auto s = std::make_unique<aqua_scene>();
  s->verts.assign(verts, verts + size_t(n_tris) * 9);
  s->mat_ids.assign(mat_ids, mat_ids + n_tris);
  s->materials.assign(mats, mats + n_mats);
  for (uint32_t m = 0; m < n_mats; ++m) s->material_names.push_back("m" + std::to_string(m));
  s->camera = *cam;
  s->point_lights.assign(pl, pl + n_pl);
  finalize_scene_info(s.get());
  s->info.n_shapes = 1;
  s->info.n_vertices = 3 * n_tris;
  *out = s.release();
  return AQUA_OK;
}

}  // namespace aqua
