// lh_json.h -- minimal JSON reader for the scene configs.
//
// The reference parses its config with nlohmann::json (SceneParameter.h:5,
// 20-151): numbers are read as double and converted with get<float>/get<int>;
// a missing key throws json::out_of_range.  This reader keeps those semantics:
// at() throws lh::json_error for a missing key or a type mismatch.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lh {

struct json_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum Kind { Null, Bool, Number, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  static Json parse(const std::string& text) {
    size_t i = 0;
    Json v = parse_value(text, i);
    skip_ws(text, i);
    if (i != text.size()) throw json_error("json: trailing characters at offset " + std::to_string(i));
    return v;
  }

  bool has(const std::string& k) const { return kind == Object && obj.count(k) > 0; }
  const Json& at(const std::string& k) const {
    if (kind != Object) throw json_error("json: not an object when reading key '" + k + "'");
    auto it = obj.find(k);
    if (it == obj.end()) throw json_error("json: missing key '" + k + "'");
    return it->second;
  }
  const Json& at(size_t i) const {
    if (kind != Array || i >= arr.size()) throw json_error("json: array index out of range");
    return arr[i];
  }
  size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
  double as_double() const {
    if (kind != Number) throw json_error("json: expected a number");
    return num;
  }
  float as_float() const { return (float)as_double(); }
  int as_int() const {
    if (kind != Number) throw json_error("json: expected a number");
    return (int)num;
  }
  bool as_bool() const {
    if (kind != Bool) throw json_error("json: expected a boolean");
    return b;
  }
  const std::string& as_string() const {
    if (kind != String) throw json_error("json: expected a string");
    return str;
  }
  // optional accessors with defaults (build extensions only)
  double get(const std::string& k, double dflt) const { return has(k) ? at(k).as_double() : dflt; }
  std::string get(const std::string& k, const char* dflt) const { return has(k) ? at(k).as_string() : dflt; }

 private:
  static void skip_ws(const std::string& s, size_t& i) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
  }
  static Json parse_value(const std::string& s, size_t& i) {
    skip_ws(s, i);
    if (i >= s.size()) throw json_error("json: unexpected end of input");
    char c = s[i];
    Json v;
    if (c == '{') {
      v.kind = Object;
      ++i;
      skip_ws(s, i);
      if (i < s.size() && s[i] == '}') { ++i; return v; }
      while (true) {
        skip_ws(s, i);
        Json key = parse_value(s, i);
        if (key.kind != String) throw json_error("json: object key must be a string");
        skip_ws(s, i);
        if (i >= s.size() || s[i] != ':') throw json_error("json: expected ':'");
        ++i;
        v.obj[key.str] = parse_value(s, i);
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        if (i < s.size() && s[i] == '}') { ++i; break; }
        throw json_error("json: expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = Array;
      ++i;
      skip_ws(s, i);
      if (i < s.size() && s[i] == ']') { ++i; return v; }
      while (true) {
        v.arr.push_back(parse_value(s, i));
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        if (i < s.size() && s[i] == ']') { ++i; break; }
        throw json_error("json: expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = String;
      ++i;
      while (i < s.size() && s[i] != '"') {
        if (s[i] == '\\' && i + 1 < s.size()) {
          char e = s[i + 1];
          v.str.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
          i += 2;
        } else {
          v.str.push_back(s[i++]);
        }
      }
      if (i >= s.size()) throw json_error("json: unterminated string");
      ++i;
    } else if (s.compare(i, 4, "true") == 0) {
      v.kind = Bool; v.b = true; i += 4;
    } else if (s.compare(i, 5, "false") == 0) {
      v.kind = Bool; v.b = false; i += 5;
    } else if (s.compare(i, 4, "null") == 0) {
      v.kind = Null; i += 4;
    } else {
      const char* start = s.c_str() + i;
      char* end = nullptr;
      double d = std::strtod(start, &end);
      if (end == start) throw json_error("json: bad value at offset " + std::to_string(i));
      v.kind = Number;
      v.num = d;
      i += (size_t)(end - start);
    }
    return v;
  }
};

}  // namespace lh
