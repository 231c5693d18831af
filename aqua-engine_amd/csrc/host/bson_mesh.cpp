// BSON reader/writer for the reference's `.mesh` files.
//
// Layout (scenes/CornellBox-Original_floor_0.mesh@0-489): one BSON document
//   { name: string,
//     vertices:  [[x, y, z], ...]  doubles,
//     normals:   [[x, y, z], ...]  doubles, one per vertex,
//     texcoords: [[u, v], ...]     doubles (empty in the Cornell box meshes),
//     indices:   [[i, j, k], ...]  int64 }
// Arrays are BSON documents whose keys are "0", "1", ... (bsonspec.org 1.1).
// The reader accepts int32/int64/double for every numeric leaf and rejects
// anything else with AQUA_ERR_PARSE; every read is bounds-checked.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "internal.h"

namespace aqua {
namespace {

struct Reader {
  const uint8_t* p;
  size_t n;
  std::string err;

  bool bad(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool i32(size_t o, int32_t* v) {
    if (o + 4 > n) return bad("truncated int32");
    memcpy(v, p + o, 4);
    return true;
  }
  bool i64(size_t o, int64_t* v) {
    if (o + 8 > n) return bad("truncated int64");
    memcpy(v, p + o, 8);
    return true;
  }
  bool f64(size_t o, double* v) {
    if (o + 8 > n) return bad("truncated double");
    memcpy(v, p + o, 8);
    return true;
  }
  // Reads a cstring key at o; returns offset after the NUL.
  bool key(size_t o, std::string* k, size_t* next) {
    size_t e = o;
    while (e < n && p[e] != 0) ++e;
    if (e >= n) return bad("unterminated key");
    k->assign(reinterpret_cast<const char*>(p + o), e - o);
    *next = e + 1;
    return true;
  }
  // Bounds of the document starting at o: [o+4, end-1) holds elements.
  bool doc_bounds(size_t o, size_t limit, size_t* end) {
    int32_t len;
    if (!i32(o, &len)) return false;
    if (len < 5 || o + size_t(len) > limit) return bad("document length out of range");
    if (p[o + len - 1] != 0) return bad("document not NUL-terminated");
    *end = o + size_t(len);
    return true;
  }
  // Size in bytes of the value of type t at offset o (within limit).
  bool value_size(uint8_t t, size_t o, size_t limit, size_t* sz) {
    switch (t) {
      case 0x01: *sz = 8; break;
      case 0x10: *sz = 4; break;
      case 0x12: case 0x09: case 0x11: *sz = 8; break;
      case 0x08: *sz = 1; break;
      case 0x0A: *sz = 0; break;
      case 0x02: {
        int32_t l;
        if (!i32(o, &l)) return false;
        if (l < 1) return bad("bad string length");
        *sz = 4 + size_t(l);
        break;
      }
      case 0x03: case 0x04: {
        size_t e;
        if (!doc_bounds(o, limit, &e)) return false;
        *sz = e - o;
        break;
      }
      case 0x05: {
        int32_t l;
        if (!i32(o, &l)) return false;
        if (l < 0) return bad("bad binary length");
        *sz = 5 + size_t(l);
        break;
      }
      default: {
        char b[64];
        snprintf(b, sizeof b, "unsupported BSON element type 0x%02x", t);
        return bad(b);
      }
    }
    if (o + *sz > limit) return bad("element overruns document");
    return true;
  }
  bool number(uint8_t t, size_t o, double* v) {
    if (t == 0x01) return f64(o, v);
    if (t == 0x10) {
      int32_t x;
      if (!i32(o, &x)) return false;
      *v = x;
      return true;
    }
    if (t == 0x12) {
      int64_t x;
      if (!i64(o, &x)) return false;
      *v = double(x);
      return true;
    }
    return bad("numeric leaf expected");
  }
  bool integer(uint8_t t, size_t o, int64_t* v) {
    if (t == 0x12) return i64(o, v);
    if (t == 0x10) {
      int32_t x;
      if (!i32(o, &x)) return false;
      *v = x;
      return true;
    }
    if (t == 0x01) {
      double d;
      if (!f64(o, &d)) return false;
      if (d != std::floor(d) || std::fabs(d) > 9.0e15) return bad("non-integer index");
      *v = int64_t(d);
      return true;
    }
    return bad("integer index expected");
  }
  // Array of fixed-arity numeric arrays -> flat vector.
  template <class T, class F>
  bool tuples(size_t o, size_t limit, int arity, std::vector<T>* out, const char* what, F leaf) {
    size_t end;
    if (!doc_bounds(o, limit, &end)) return false;
    size_t q = o + 4;
    size_t idx = 0;
    while (q < end - 1) {
      uint8_t t = p[q];
      std::string k;
      size_t vo;
      if (!key(q + 1, &k, &vo)) return false;
      if (k != std::to_string(idx)) return bad(std::string(what) + ": array key out of order");
      if (t != 0x04) return bad(std::string(what) + ": element is not an array");
      size_t iend;
      if (!doc_bounds(vo, end, &iend)) return false;
      size_t r = vo + 4;
      int got = 0;
      while (r < iend - 1) {
        uint8_t lt = p[r];
        std::string lk;
        size_t lvo;
        if (!key(r + 1, &lk, &lvo)) return false;
        if (got >= arity) return bad(std::string(what) + ": tuple too long");
        T v;
        if (!leaf(lt, lvo, &v)) return false;
        out->push_back(v);
        size_t sz;
        if (!value_size(lt, lvo, iend, &sz)) return false;
        r = lvo + sz;
        ++got;
      }
      if (got != arity) return bad(std::string(what) + ": tuple too short");
      q = iend;
      ++idx;
    }
    return true;
  }
};

bool read_file(const std::string& path, std::string* data) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *data = ss.str();
  return bool(f) || f.eof();
}

// ---- writer ----
struct Writer {
  std::string b;
  void u8(uint8_t v) { b.push_back(char(v)); }
  void i32(int32_t v) { b.append(reinterpret_cast<const char*>(&v), 4); }
  void i64(int64_t v) { b.append(reinterpret_cast<const char*>(&v), 8); }
  void f64(double v) { b.append(reinterpret_cast<const char*>(&v), 8); }
  void cstr(const std::string& s) {
    b += s;
    b.push_back('\0');
  }
  size_t begin() {
    size_t at = b.size();
    i32(0);
    return at;
  }
  void end(size_t at) {
    b.push_back('\0');
    int32_t len = int32_t(b.size() - at);
    memcpy(&b[at], &len, 4);
  }
  void tuples_f64(const char* name, const std::vector<double>& v, int arity) {
    u8(0x04);
    cstr(name);
    size_t a = begin();
    size_t n = arity ? v.size() / size_t(arity) : 0;
    for (size_t i = 0; i < n; ++i) {
      u8(0x04);
      cstr(std::to_string(i));
      size_t e = begin();
      for (int k = 0; k < arity; ++k) {
        u8(0x01);
        cstr(std::to_string(k));
        f64(v[i * size_t(arity) + size_t(k)]);
      }
      end(e);
    }
    end(a);
  }
  void tuples_i64(const char* name, const std::vector<int64_t>& v, int arity) {
    u8(0x04);
    cstr(name);
    size_t a = begin();
    size_t n = v.size() / size_t(arity);
    for (size_t i = 0; i < n; ++i) {
      u8(0x04);
      cstr(std::to_string(i));
      size_t e = begin();
      for (int k = 0; k < arity; ++k) {
        u8(0x12);
        cstr(std::to_string(k));
        i64(v[i * size_t(arity) + size_t(k)]);
      }
      end(e);
    }
    end(a);
  }
};

}  // namespace

int read_mesh_bson(const std::string& path, MeshData* out) {
  std::string data;
  if (!read_file(path, &data)) return fail(AQUA_ERR_MISSING_MESH, "cannot open mesh '%s'", path.c_str());
  Reader r{reinterpret_cast<const uint8_t*>(data.data()), data.size(), {}};
  *out = MeshData();
  size_t end;
  bool ok = r.doc_bounds(0, data.size(), &end);
  if (ok && end != data.size()) ok = r.bad("trailing bytes after document");
  bool have_v = false, have_i = false;
  size_t q = 4;
  auto num = [&](uint8_t t, size_t o, double* v) { return r.number(t, o, v); };
  auto idx = [&](uint8_t t, size_t o, int64_t* v) { return r.integer(t, o, v); };
  while (ok && q < end - 1) {
    uint8_t t = data[q];
    std::string k;
    size_t vo;
    if (!r.key(q + 1, &k, &vo)) {
      ok = false;
      break;
    }
    size_t sz;
    if (!r.value_size(t, vo, end, &sz)) {
      ok = false;
      break;
    }
    if (k == "name" && t == 0x02) {
      int32_t l;
      memcpy(&l, data.data() + vo, 4);
      out->name.assign(data.data() + vo + 4, size_t(l) - 1);
    } else if (k == "vertices" && t == 0x04) {
      ok = r.tuples(vo, end, 3, &out->vertices, "vertices", num);
      have_v = true;
    } else if (k == "normals" && t == 0x04) {
      ok = r.tuples(vo, end, 3, &out->normals, "normals", num);
    } else if (k == "texcoords" && t == 0x04) {
      ok = r.tuples(vo, end, 2, &out->texcoords, "texcoords", num);
    } else if (k == "indices" && t == 0x04) {
      ok = r.tuples(vo, end, 3, &out->indices, "indices", idx);
      have_i = true;
    }
    q = vo + sz;
  }
  if (ok && !(have_v && have_i)) ok = r.bad("missing 'vertices' or 'indices'");
  if (ok) {
    size_t nv = out->vertices.size() / 3;
    for (int64_t i : out->indices)
      if (i < 0 || uint64_t(i) >= nv) {
        ok = r.bad("index out of range");
        break;
      }
    if (ok && !out->normals.empty() && out->normals.size() != out->vertices.size())
      ok = r.bad("normals count != vertices count");
    for (double v : out->vertices)
      if (!std::isfinite(v)) {
        return fail(AQUA_ERR_NONFINITE, "mesh '%s': non-finite vertex coordinate", path.c_str());
      }
  }
  if (!ok) return fail(AQUA_ERR_PARSE, "mesh '%s': %s", path.c_str(), r.err.c_str());
  return AQUA_OK;
}

int write_mesh_bson(const std::string& path, const MeshData& m) {
  Writer w;
  size_t top = w.begin();
  w.u8(0x02);
  w.cstr("name");
  w.i32(int32_t(m.name.size() + 1));
  w.cstr(m.name);
  w.tuples_f64("vertices", m.vertices, 3);
  w.tuples_f64("normals", m.normals, 3);
  w.tuples_f64("texcoords", m.texcoords, 2);
  w.tuples_i64("indices", m.indices, 3);
  w.end(top);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return fail(AQUA_ERR_IO, "cannot write '%s'", path.c_str());
  size_t n = fwrite(w.b.data(), 1, w.b.size(), f);
  int cl = fclose(f);
  if (n != w.b.size() || cl != 0) return fail(AQUA_ERR_IO, "short write to '%s'", path.c_str());
  return AQUA_OK;
}

}  // namespace aqua
