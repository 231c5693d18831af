// Procedural scenes, written in the reference's own schema (scene JSON as in
// scenes/cbox.json, one BSON .mesh per shape as in
// scenes/CornellBox-Original_floor_0.mesh) so every render goes through the
// loader.
//
//  BOX_ROOM  An axis-aligned idealisation of scenes/CornellBox-Original.obj:
//            walls on [-1,1]x[0,2]x[-1,1] (obj:14-55 spans x -1.02..1,
//            y 0..1.99, z -1.04..0.99), a short 0.6 box and a tall 1.2 box
//            rotated -17 / +17 degrees about y at the OBJ's footprints
//            (obj:61-99 and obj:109-147, faces at 16.6 and 17.5 degrees), the
//            ceiling light quad of obj:152-156 with Ke = (17, 12, 4)
//            (CornellBox-Original.mtl:88), Lambertian albedos from the MTL
//            (red mtl:17, green mtl:27) and 0.75 grey.
//  INTERIOR  A closed room with `n_meshes` seeded subdivided boxes and
//            icospheres (about `target_tris` triangles), GGX
//            metallic-roughness materials (roughness U[0.05,1], metallic
//            Bernoulli(0.3), albedo U[0.1,0.9]^3) and 3 ceiling quad lights.
#include <sys/stat.h>

#include <cmath>
#include <cstring>
#include <string>

#include "internal.h"
#include "json.h"

namespace aqua {
namespace {

// PCG32 (O'Neill 2014, PCG-XSH-RR 64/32) for the generators' random choices.
struct Pcg {
  uint64_t state = 0, inc = 1;
  Pcg(uint64_t initstate, uint64_t initseq) {
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  double uniform() { return double(next()) * (1.0 / 4294967296.0); }
  double range(double a, double b) { return a + (b - a) * uniform(); }
};

struct Mat {
  std::string name;
  double color[3];
  double metallic, roughness, specular;
  double emission[3];
};

struct Shape {
  std::string file;
  std::string mat;
};

struct Builder {
  std::string dir;
  std::vector<Mat> mats;
  std::vector<Shape> shapes;
  int err = AQUA_OK;

  int add_mesh(const std::string& stem, const std::string& mat, MeshData& m) {
    std::string file = stem + ".mesh";
    if (int e = write_mesh_bson(dir + "/" + file, m)) return e;
    shapes.push_back({file, mat});
    return AQUA_OK;
  }
};

void push_v(MeshData& m, const double p[3], const double n[3]) {
  for (int k = 0; k < 3; ++k) m.vertices.push_back(p[k]);
  for (int k = 0; k < 3; ++k) m.normals.push_back(n[k]);
}

// Round a double to the nearest FP32 value, so the written doubles are
// FP32-exact like every coordinate of the shipped meshes.
double f32(double v) { return double(float(v)); }

// Quad (p0,p1,p2,p3) fan-triangulated (0,1,2),(0,2,3) like the shipped
// meshes; normal from the winding.
void quad(MeshData& m, const double p[4][3]) {
  double e1[3], e2[3], n[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = p[1][k] - p[0][k];
    e2[k] = p[2][k] - p[0][k];
  }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (double& c : n) c = f32(c / l);
  int64_t base = int64_t(m.vertices.size() / 3);
  for (int i = 0; i < 4; ++i) {
    double q[3] = {f32(p[i][0]), f32(p[i][1]), f32(p[i][2])};
    push_v(m, q, n);
  }
  int64_t idx[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
  m.indices.insert(m.indices.end(), idx, idx + 6);
}

// Box with per-face n x n grids. Local cube [-1,1]^3 scaled by half-extents
// `he`, rotated by `rot` (3x3 row-major), translated to `c`. Windings and
// normals face outward, or inward when `inward` (a closed room).
void box(MeshData& m, const double c[3], const double he[3], const double rot[9], int n,
         bool inward = false) {
  // face: fixed axis a with sign s; u, v span the other two axes so that
  // cross(u, v) points along s * axis a.
  for (int a = 0; a < 3; ++a)
    for (int s = -1; s <= 1; s += 2) {
      int u = (a + 1) % 3, v = (a + 2) % 3;
      if ((s < 0) != inward) std::swap(u, v);
      double nl[3] = {0, 0, 0};
      nl[a] = inward ? -s : s;
      double nw[3];
      for (int r = 0; r < 3; ++r)
        nw[r] = f32(rot[r * 3 + 0] * nl[0] + rot[r * 3 + 1] * nl[1] + rot[r * 3 + 2] * nl[2]);
      int64_t base = int64_t(m.vertices.size() / 3);
      for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
          double l[3];
          l[a] = s;
          l[u] = -1.0 + 2.0 * i / n;
          l[v] = -1.0 + 2.0 * j / n;
          double q[3];
          for (int r = 0; r < 3; ++r) {
            double x = rot[r * 3 + 0] * l[0] * he[0] + rot[r * 3 + 1] * l[1] * he[1] +
                       rot[r * 3 + 2] * l[2] * he[2];
            q[r] = f32(c[r] + x);
          }
          push_v(m, q, nw);
        }
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          int64_t i0 = base + j * (n + 1) + i;
          int64_t i1 = i0 + 1, i2 = i0 + (n + 1) + 1, i3 = i0 + (n + 1);
          int64_t idx[6] = {i0, i1, i2, i0, i2, i3};
          m.indices.insert(m.indices.end(), idx, idx + 6);
        }
    }
}

// Icosphere: icosahedron subdivided `level` times, 20 * 4^level triangles,
// vertices projected to the sphere (smooth normals = radial directions).
void icosphere(MeshData& m, const double c[3], double radius, int level) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> P = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  std::vector<int> F = {0, 11, 5, 0, 5,  1,  0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                        4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3,  9,  4, 3,  4,  2, 3, 2, 6, 3,
                        6, 8,  3, 8, 9,  4, 9, 5, 2, 4, 11, 6,  2, 10, 8,  6, 7, 9, 8, 1};
  auto norm = [&](size_t i) {
    double l = std::sqrt(P[3 * i] * P[3 * i] + P[3 * i + 1] * P[3 * i + 1] + P[3 * i + 2] * P[3 * i + 2]);
    for (int k = 0; k < 3; ++k) P[3 * i + k] /= l;
  };
  for (size_t i = 0; i < P.size() / 3; ++i) norm(i);
  for (int lv = 0; lv < level; ++lv) {
    std::vector<int> G;
    std::vector<std::pair<uint64_t, int>> cache;  // sorted edge -> midpoint index
    std::vector<std::vector<std::pair<int, int>>> adj(P.size() / 3);
    auto mid = [&](int a, int b) {
      int lo = a < b ? a : b, hi = a < b ? b : a;
      for (auto& e : adj[size_t(lo)])
        if (e.first == hi) return e.second;
      int id = int(P.size() / 3);
      for (int k = 0; k < 3; ++k) P.push_back(0.5 * (P[3 * size_t(a) + k] + P[3 * size_t(b) + k]));
      norm(size_t(id));
      adj[size_t(lo)].push_back({hi, id});
      adj.emplace_back();
      return id;
    };
    for (size_t f = 0; f < F.size(); f += 3) {
      int a = F[f], b = F[f + 1], cc = F[f + 2];
      int ab = mid(a, b), bc = mid(b, cc), ca = mid(cc, a);
      int tri[12] = {a, ab, ca, b, bc, ab, cc, ca, bc, ab, bc, ca};
      G.insert(G.end(), tri, tri + 12);
    }
    F.swap(G);
  }
  int64_t base = int64_t(m.vertices.size() / 3);
  for (size_t i = 0; i < P.size() / 3; ++i) {
    double n[3] = {f32(P[3 * i]), f32(P[3 * i + 1]), f32(P[3 * i + 2])};
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = f32(c[k] + radius * P[3 * i + size_t(k)]);
    push_v(m, q, n);
  }
  for (int idx : F) m.indices.push_back(base + idx);
}

void rot_y(double deg, double r[9]) {
  double a = deg * M_PI / 180.0, c = std::cos(a), s = std::sin(a);
  double R[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
  memcpy(r, R, sizeof R);
}

// Random rotation from a random unit quaternion (Shoemake 1992).
void rot_random(Pcg& g, double r[9]) {
  double u1 = g.uniform(), u2 = g.uniform(), u3 = g.uniform();
  double a = std::sqrt(1 - u1), b = std::sqrt(u1);
  double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
  double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
  double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                 2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                 2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
  memcpy(r, R, sizeof R);
}

std::string tex_float(double v) { return "{\"Float\": " + json_number(v) + "}"; }
std::string tex_float3(const double v[3]) {
  return "{\"Float3\": [" + json_number(v[0]) + ", " + json_number(v[1]) + ", " + json_number(v[2]) + "]}";
}

std::string principled_json(const Mat& m) {
  // All 17 Principled parameters in the order of cbox.json:5-63, plus "hint".
  const double ssr[3] = {1.0, double(0.2f), double(0.1f)};
  std::string s = "{\"Principled\": {";
  s += "\"color\": " + tex_float3(m.color);
  s += ", \"subsurface\": " + tex_float(0.0);
  s += ", \"subsurface_radius\": " + tex_float3(ssr);
  s += ", \"subsurface_color\": " + tex_float(0.0);
  s += ", \"metallic\": " + tex_float(m.metallic);
  s += ", \"specular\": " + tex_float(m.specular);
  s += ", \"specular_tint\": " + tex_float(0.0);
  s += ", \"roughness\": " + tex_float(m.roughness);
  s += ", \"anisotropic\": " + tex_float(0.0);
  s += ", \"anisotropic_rotation\": " + tex_float(0.0);
  s += ", \"sheen\": " + tex_float(0.0);
  s += ", \"sheen_tint\": " + tex_float(0.5);
  s += ", \"clearcoat\": " + tex_float(0.0);
  s += ", \"clearcoat_roughness\": " + tex_float(double(0.03f));
  s += ", \"ior\": " + tex_float(double(1.45f));
  s += ", \"transmission\": " + tex_float(0.0);
  if (m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0)
    s += ", \"emission\": " + tex_float3(m.emission);
  else
    s += ", \"emission\": " + tex_float(0.0);
  s += ", \"hint\": \"ltc\"}}";
  return s;
}

int write_scene_json(const Builder& b, const char* path, uint32_t w, uint32_t h, double fov,
                     const double tr[3], const double rot[3]) {
  std::string s = "{\n  \"named_bsdfs\": {\n";
  for (size_t i = 0; i < b.mats.size(); ++i) {
    s += "    " + json_escape(b.mats[i].name) + ": " + principled_json(b.mats[i]);
    s += (i + 1 < b.mats.size()) ? ",\n" : "\n";
  }
  s += "  },\n  \"camera\": {\"Perspective\": {\"res\": [" + std::to_string(w) + ", " + std::to_string(h) +
       "], \"fov\": " + json_number(fov) + ", \"lens_radius\": 0.0, \"focal\": 1.0, \"transform\": {" +
       "\"translate\": [" + json_number(tr[0]) + ", " + json_number(tr[1]) + ", " + json_number(tr[2]) +
       "], \"rotate\": [" + json_number(rot[0]) + ", " + json_number(rot[1]) + ", " + json_number(rot[2]) +
       "], \"scale\": [1.0, 1.0, 1.0]}}},\n  \"lights\": [],\n  \"shapes\": [\n";
  for (size_t i = 0; i < b.shapes.size(); ++i) {
    s += "    {\"Mesh\": [" + json_escape(b.shapes[i].file) + ", {\"Named\": " + json_escape(b.shapes[i].mat) +
         "}]}";
    s += (i + 1 < b.shapes.size()) ? ",\n" : "\n";
  }
  s += "  ]\n}\n";
  FILE* f = fopen(path, "wb");
  if (!f) return fail(AQUA_ERR_IO, "cannot write '%s'", path);
  size_t n = fwrite(s.data(), 1, s.size(), f);
  int cl = fclose(f);
  if (n != s.size() || cl != 0) return fail(AQUA_ERR_IO, "short write to '%s'", path);
  return AQUA_OK;
}

int gen_box_room(Builder& b, const aqua_procgen_params& p, std::string* json_path) {
  b.mats = {
      {"white", {0.75, 0.75, 0.75}, 0, 1, 0, {0, 0, 0}},
      {"red", {0.63, 0.065, 0.05}, 0, 1, 0, {0, 0, 0}},
      {"green", {0.14, 0.45, 0.091}, 0, 1, 0, {0, 0, 0}},
      {"light", {0.78, 0.78, 0.78}, 0, 1, 0, {17, 12, 4}},
  };
  for (Mat& m : b.mats)
    for (int k = 0; k < 3; ++k) m.color[k] = f32(m.color[k]);
  // Walls: inward-facing windings.
  const double W[5][4][3] = {
      {{-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, -1}},      // floor    (+y)
      {{-1, 2, 1}, {-1, 2, -1}, {1, 2, -1}, {1, 2, 1}},      // ceiling  (-y)
      {{-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}},    // back     (+z)
      {{1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}},        // right    (-x)
      {{-1, 0, 1}, {-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}},    // left     (+x)
  };
  const char* names[5] = {"floor", "ceiling", "backWall", "rightWall", "leftWall"};
  const char* wm[5] = {"white", "white", "white", "green", "red"};
  for (int i = 0; i < 5; ++i) {
    MeshData m;
    m.name = names[i];
    quad(m, W[i]);
    if (int e = b.add_mesh(std::string("box_room_") + names[i], wm[i], m)) return e;
  }
  double R[9];
  {
    MeshData m;
    m.name = "shortBox";
    double c[3] = {0.3275, 0.3, 0.3725}, he[3] = {0.3, 0.3, 0.3};
    rot_y(-17.0, R);
    box(m, c, he, R, 1);
    if (int e = b.add_mesh("box_room_shortBox", "white", m)) return e;
  }
  {
    MeshData m;
    m.name = "tallBox";
    double c[3] = {-0.335, 0.6, -0.29}, he[3] = {0.3, 0.6, 0.3};
    rot_y(17.0, R);
    box(m, c, he, R, 1);
    if (int e = b.add_mesh("box_room_tallBox", "white", m)) return e;
  }
  {
    MeshData m;
    m.name = "light";
    const double L[4][3] = {{-0.24, 1.98, 0.16}, {-0.24, 1.98, -0.22}, {0.23, 1.98, -0.22}, {0.23, 1.98, 0.16}};
    quad(m, L);  // winding gives a -y (downward) normal
    if (int e = b.add_mesh("box_room_light", "light", m)) return e;
  }
  *json_path = b.dir + "/scene.json";
  uint32_t w = p.width ? p.width : 512, h = p.height ? p.height : 512;
  const double tr[3] = {0.0, 1.0, 3.8}, rot[3] = {0, 0, 0};
  return write_scene_json(b, json_path->c_str(), w, h, 40.0, tr, rot);
}

int gen_interior(Builder& b, const aqua_procgen_params& p, std::string* json_path) {
  Pcg g(p.seed, 0x1f2e3d4cULL);
  const double R = 10.0, H = 6.0;
  uint32_t n = p.n_meshes ? p.n_meshes : 150;
  double budget = double(p.target_tris ? p.target_tris : 1000000u) / n;
  b.mats.push_back({"room", {0.7, 0.7, 0.7}, 0, 0.8, 0.5, {0, 0, 0}});
  b.mats.push_back({"lamp", {0.8, 0.8, 0.8}, 0, 1, 0, {12, 11, 9.5}});
  for (Mat& m : b.mats)
    for (int k = 0; k < 3; ++k) m.color[k] = f32(m.color[k]);
  {
    MeshData m;
    m.name = "room";
    // Closed room: a box whose faces point inward.
    double c[3] = {0, H / 2, 0}, he[3] = {R, H / 2, R}, I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    box(m, c, he, I, 1, true);
    if (int e = b.add_mesh("interior_room", "room", m)) return e;
  }
  const double lights[3][2] = {{-5.0, -3.0}, {0.5, 2.0}, {5.0, -4.5}};
  for (int i = 0; i < 3; ++i) {
    MeshData m;
    m.name = "lamp" + std::to_string(i);
    double x = lights[i][0], z = lights[i][1], y = H - 0.01, s = 1.0;
    const double Q[4][3] = {{x - s, y, z + s}, {x - s, y, z - s}, {x + s, y, z - s}, {x + s, y, z + s}};
    quad(m, Q);
    if (int e = b.add_mesh("interior_lamp" + std::to_string(i), "lamp", m)) return e;
  }
  for (uint32_t i = 0; i < n; ++i) {
    Mat mt;
    mt.name = "obj" + std::to_string(i);
    for (int k = 0; k < 3; ++k) mt.color[k] = f32(g.range(0.1, 0.9));
    mt.roughness = f32(g.range(0.05, 1.0));
    mt.metallic = g.uniform() < 0.3 ? 1.0 : 0.0;
    mt.specular = 0.5;
    mt.emission[0] = mt.emission[1] = mt.emission[2] = 0;
    b.mats.push_back(mt);
    double size = g.range(0.3, 1.4);
    double c[3] = {g.range(-R + 1.5, R - 1.5), g.range(size, H - 1.5 - size * 0.5), g.range(-R + 1.5, R - 1.5)};
    double jitter = g.range(0.6, 1.4);
    MeshData m;
    m.name = mt.name;
    if (g.uniform() < 0.5) {
      int sub = int(std::lround(std::sqrt(budget * jitter / 12.0)));
      if (sub < 1) sub = 1;
      double he[3] = {size * g.range(0.5, 1.0), size * g.range(0.5, 1.0), size * g.range(0.5, 1.0)};
      double Rm[9];
      rot_random(g, Rm);
      box(m, c, he, Rm, sub);
    } else {
      double lv = std::log(budget * jitter / 20.0) / std::log(4.0);
      int level = int(std::floor(lv));
      if (g.uniform() < lv - level) ++level;
      if (level < 0) level = 0;
      if (level > 8) level = 8;
      icosphere(m, c, size * 0.7, level);
    }
    if (int e = b.add_mesh("interior_obj" + std::to_string(i), mt.name, m)) return e;
  }
  *json_path = b.dir + "/scene.json";
  uint32_t w = p.width ? p.width : 1920, h = p.height ? p.height : 1080;
  const double tr[3] = {0.0, 3.0, 9.5}, rot[3] = {-8.0, 0, 0};
  return write_scene_json(b, json_path->c_str(), w, h, 60.0, tr, rot);
}

}  // namespace

int procgen_write(const aqua_procgen_params* p, const char* out_dir, char* scene_path, size_t cap) {
  if (!p || !out_dir || !scene_path || cap == 0)
    return fail(AQUA_ERR_INVALID_ARG, "aqua_procgen_write: null argument");
  struct stat st;
  if (stat(out_dir, &st) != 0) {
    if (mkdir(out_dir, 0755) != 0) return fail(AQUA_ERR_IO, "cannot create directory '%s'", out_dir);
  } else if (!S_ISDIR(st.st_mode)) {
    return fail(AQUA_ERR_IO, "'%s' is not a directory", out_dir);
  }
  Builder b;
  b.dir = out_dir;
  std::string path;
  int e;
  if (p->kind == AQUA_PROCGEN_BOX_ROOM)
    e = gen_box_room(b, *p, &path);
  else if (p->kind == AQUA_PROCGEN_INTERIOR)
    e = gen_interior(b, *p, &path);
  else
    return fail(AQUA_ERR_INVALID_ARG, "aqua_procgen_write: unknown kind %u", p->kind);
  if (e) return e;
  if (path.size() + 1 > cap) return fail(AQUA_ERR_INVALID_ARG, "aqua_procgen_write: path buffer too small");
  memcpy(scene_path, path.c_str(), path.size() + 1);
  return AQUA_OK;
}

}  // namespace aqua
