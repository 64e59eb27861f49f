#include "json.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace clsph_json {
namespace {

struct reader {
  const std::string& s;
  size_t i = 0;

  [[noreturn]] void error(const std::string& what) {
    throw std::runtime_error("JSON parse error at offset " + std::to_string(i) + ": " + what);
  }
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s.compare(i, n, w) == 0) {
      i += n;
      return true;
    }
    return false;
  }
  std::string str() {
    if (s[i] != '"') error("expected '\"'");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) error("bad escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 > s.size()) error("bad \\u escape");
            unsigned cp = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
            i += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) error("unterminated string");
    ++i;
    return out;
  }
  value val() {
    ws();
    if (i >= s.size()) error("unexpected end");
    value v;
    char c = s[i];
    if (c == '{') {
      ++i;
      v.kind = value::object_t;
      ws();
      if (s[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws();
        std::string k = str();
        ws();
        if (s[i] != ':') error("expected ':'");
        ++i;
        v.obj[k] = val();
        ws();
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == '}') {
          ++i;
          return v;
        }
        error("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++i;
      v.kind = value::array_t;
      ws();
      if (s[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.arr.push_back(val());
        ws();
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == ']') {
          ++i;
          return v;
        }
        error("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = value::string_t;
      v.str = str();
      return v;
    }
    if (lit("true")) {
      v.kind = value::bool_t;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = value::bool_t;
      return v;
    }
    if (lit("null")) return v;
    if (c == '-' || std::isdigit((unsigned char)c)) {
      const char* b = s.c_str() + i;
      char* e = nullptr;
      v.kind = value::number_t;
      v.num = std::strtod(b, &e);
      if (e == b) error("bad number");
      i += (size_t)(e - b);
      return v;
    }
    error(std::string("unexpected character '") + c + "'");
  }
};

}  // namespace

value parse(const std::string& text) {
  reader r{text};
  return r.val();
}

value parse_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace clsph_json
