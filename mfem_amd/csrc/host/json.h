// Minimal JSON reader for scene files (reference scenes/*.json, consumed by
// src/simulation_state.cpp:72-366). Supports objects, arrays, strings,
// numbers, true/false/null; duplicate keys keep the last value (as
// nlohmann::json does when parsing).
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mfem {

class Json {
 public:
  enum Type { NUL, BOOL, NUM, STR, ARR, OBJ };
  Type type = NUL;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  static Json parse(const std::string& text);
  static Json parse_file(const std::string& path);

  bool is_object() const { return type == OBJ; }
  bool is_array() const { return type == ARR; }
  bool has(const std::string& k) const { return type == OBJ && obj.count(k) > 0; }
  const Json& operator[](const std::string& k) const;
  const Json& operator[](size_t i) const { return arr.at(i); }
  size_t size() const { return type == ARR ? arr.size() : (type == OBJ ? obj.size() : 0); }

  double as_number() const;
  int as_int() const { return (int)as_number(); }
  bool as_bool() const;
  const std::string& as_string() const;
  std::vector<double> as_numbers() const;
  std::vector<std::string> as_strings() const;
};

}  // namespace mfem
