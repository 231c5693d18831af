#include "json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace aqua {

const Json* Json::get(const char* key) const {
  if (kind != Object) return nullptr;
  for (const auto& kv : obj)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;
  std::string err;
  int depth = 0;

  explicit Parser(const std::string& t) : s(t) {}

  bool error(const char* what) {
    if (err.empty()) {
      char buf[160];
      snprintf(buf, sizeof buf, "JSON: %s at byte %zu", what, i);
      err = buf;
    }
    return false;
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
  }
  bool lit(const char* w) {
    size_t n = strlen(w);
    if (s.compare(i, n, w) != 0) return false;
    i += n;
    return true;
  }
  static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
  }
  static void put_utf8(std::string& o, uint32_t cp) {
    if (cp < 0x80) {
      o += char(cp);
    } else if (cp < 0x800) {
      o += char(0xC0 | (cp >> 6));
      o += char(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += char(0xE0 | (cp >> 12));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    } else {
      o += char(0xF0 | (cp >> 18));
      o += char(0x80 | ((cp >> 12) & 0x3F));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(uint32_t* v) {
    if (i + 4 > s.size()) return error("truncated \\u escape");
    uint32_t r = 0;
    for (int k = 0; k < 4; ++k) {
      int h = hexval(s[i + k]);
      if (h < 0) return error("bad \\u escape");
      r = r * 16 + uint32_t(h);
    }
    i += 4;
    *v = r;
    return true;
  }
  bool string(std::string* out) {
    if (i >= s.size() || s[i] != '"') return error("expected string");
    ++i;
    out->clear();
    while (true) {
      if (i >= s.size()) return error("unterminated string");
      char c = s[i++];
      if (c == '"') return true;
      if ((unsigned char)c < 0x20) return error("control character in string");
      if (c != '\\') {
        *out += c;
        continue;
      }
      if (i >= s.size()) return error("unterminated escape");
      char e = s[i++];
      switch (e) {
        case '"': *out += '"'; break;
        case '\\': *out += '\\'; break;
        case '/': *out += '/'; break;
        case 'b': *out += '\b'; break;
        case 'f': *out += '\f'; break;
        case 'n': *out += '\n'; break;
        case 'r': *out += '\r'; break;
        case 't': *out += '\t'; break;
        case 'u': {
          uint32_t cp;
          if (!hex4(&cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00) {
            uint32_t lo;
            if (!(lit("\\u") && hex4(&lo)) || lo < 0xDC00 || lo >= 0xE000)
              return error("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(*out, cp);
          break;
        }
        default: return error("bad escape");
      }
    }
  }
  bool number(double* out) {
    size_t st = i;
    if (i < s.size() && s[i] == '-') ++i;
    if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return error("bad number");
    if (s[i] == '0') {
      ++i;
    } else {
      while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
    }
    if (i < s.size() && s[i] == '.') {
      ++i;
      if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return error("bad fraction");
      while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
    }
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
      ++i;
      if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
      if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) return error("bad exponent");
      while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
    }
    std::string tok = s.substr(st, i - st);
    // strtod is locale-dependent only for the decimal point; the C locale is
    // the process default and is never changed by this library.
    *out = strtod(tok.c_str(), nullptr);
    return true;
  }
  bool value(Json* v) {
    if (++depth > 256) return error("nesting too deep");
    ws();
    if (i >= s.size()) return error("unexpected end");
    char c = s[i];
    bool ok = true;
    if (c == '{') {
      ++i;
      v->kind = Json::Object;
      ws();
      if (i < s.size() && s[i] == '}') {
        ++i;
      } else {
        while (true) {
          ws();
          std::string key;
          if (!string(&key)) return false;
          ws();
          if (i >= s.size() || s[i] != ':') return error("expected ':'");
          ++i;
          v->obj.emplace_back(std::move(key), Json());
          if (!value(&v->obj.back().second)) return false;
          ws();
          if (i < s.size() && s[i] == ',') {
            ++i;
            continue;
          }
          if (i < s.size() && s[i] == '}') {
            ++i;
            break;
          }
          return error("expected ',' or '}'");
        }
      }
    } else if (c == '[') {
      ++i;
      v->kind = Json::Array;
      ws();
      if (i < s.size() && s[i] == ']') {
        ++i;
      } else {
        while (true) {
          v->arr.emplace_back();
          if (!value(&v->arr.back())) return false;
          ws();
          if (i < s.size() && s[i] == ',') {
            ++i;
            continue;
          }
          if (i < s.size() && s[i] == ']') {
            ++i;
            break;
          }
          return error("expected ',' or ']'");
        }
      }
    } else if (c == '"') {
      v->kind = Json::String;
      ok = string(&v->str);
    } else if (c == 't') {
      if (!lit("true")) return error("bad literal");
      v->kind = Json::Bool;
      v->b = true;
    } else if (c == 'f') {
      if (!lit("false")) return error("bad literal");
      v->kind = Json::Bool;
      v->b = false;
    } else if (c == 'n') {
      if (!lit("null")) return error("bad literal");
      v->kind = Json::Null;
    } else {
      v->kind = Json::Number;
      ok = number(&v->num);
    }
    --depth;
    return ok;
  }
};

}  // namespace

bool json_parse(const std::string& text, Json* out, std::string* err) {
  Parser p(text);
  *out = Json();
  if (!p.value(out)) {
    if (err) *err = p.err;
    return false;
  }
  p.ws();
  if (p.i != text.size()) {
    p.error("trailing characters");
    if (err) *err = p.err;
    return false;
  }
  return true;
}

std::string json_number(double v) {
  if (!std::isfinite(v)) return "0";
  char buf[40];
  for (int prec = 1; prec <= 17; ++prec) {
    snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (strtod(buf, nullptr) == v) break;
  }
  std::string r = buf;
  // JSON needs a digit before an exponent and no bare "inf"; %g already
  // produces e.g. "1e-05", which is valid JSON.
  if (r.find_first_of(".eE") == std::string::npos) r += ".0";
  return r;
}

std::string json_escape(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\t': o += "\\t"; break;
      case '\r': o += "\\r"; break;
      default:
        if ((unsigned char)c < 0x20) {
          char b[8];
          snprintf(b, sizeof b, "\\u%04x", (unsigned char)c);
          o += b;
        } else {
          o += c;
        }
    }
  }
  o += "\"";
  return o;
}

}  // namespace aqua
