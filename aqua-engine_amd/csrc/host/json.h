// Minimal JSON DOM + recursive-descent parser/writer (RFC 8259 subset that the
// reference's scene and integrator files use: objects, arrays, strings with
// escapes, numbers, true/false/null). serde_json plays this role in the
// reference's dependency set (Cargo.toml:16).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace aqua {

struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;  // insertion order kept

  const Json* get(const char* key) const;  // nullptr if absent or not an object
  bool is_num() const { return kind == Number; }
};

// Parses `text`; on failure returns false and writes a message with the byte
// offset into *err.
bool json_parse(const std::string& text, Json* out, std::string* err);

// Writer helpers used by the procedural generators.
std::string json_number(double v);  // shortest round-trip repr ("%.17g" trimmed)
std::string json_escape(const std::string& s);

}  // namespace aqua
