// Minimal JSON reader for the fluid / simulation property files.
// Numbers are read with strtod (as picojson does) and narrowed by the caller.
// Parsing stops after the first value, so a trailing ';' (as in the shipped
// property files) is ignored.
#ifndef CLSPH_API_JSON_H_
#define CLSPH_API_JSON_H_

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace clsph_json {

struct value {
  enum kind_t { null_t, bool_t, number_t, string_t, array_t, object_t } kind = null_t;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<value> arr;
  std::map<std::string, value> obj;

  const value& at(const std::string& key) const {
    if (kind != object_t) throw std::runtime_error("JSON: not an object when looking up '" + key + "'");
    auto it = obj.find(key);
    if (it == obj.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
    return it->second;
  }
  double number(const std::string& what) const {
    if (kind != number_t) throw std::runtime_error("JSON: '" + what + "' is not a number");
    return num;
  }
  bool boolean(const std::string& what) const {
    if (kind != bool_t) throw std::runtime_error("JSON: '" + what + "' is not a boolean");
    return b;
  }
};

value parse(const std::string& text);
value parse_file(const std::string& path);

}  // namespace clsph_json

#endif  // CLSPH_API_JSON_H_
