#include "json.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mfem {

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& t) : s(t) {}

  [[noreturn]] void fail(const char* what) const {
    throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(i));
  }
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s.compare(i, n, w) == 0) { i += n; return true; }
    return false;
  }
  std::string string_() {
    if (s[i] != '"') fail("expected string");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) fail("bad escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 > s.size()) fail("bad \\u escape");
            unsigned cp = std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
            i += 4;
            if (cp < 0x80) out += (char)cp;
            else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
            else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return out;
  }
  Json value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    Json j;
    char c = s[i];
    if (c == '{') {
      ++i;
      j.type = Json::OBJ;
      ws();
      if (s[i] == '}') { ++i; return j; }
      while (true) {
        ws();
        std::string k = string_();
        ws();
        if (s[i] != ':') fail("expected ':'");
        ++i;
        j.obj[k] = value();
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == '}') { ++i; break; }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      ++i;
      j.type = Json::ARR;
      ws();
      if (s[i] == ']') { ++i; return j; }
      while (true) {
        j.arr.push_back(value());
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == ']') { ++i; break; }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      j.type = Json::STR;
      j.str = string_();
    } else if (lit("true")) {
      j.type = Json::BOOL; j.b = true;
    } else if (lit("false")) {
      j.type = Json::BOOL; j.b = false;
    } else if (lit("null")) {
      j.type = Json::NUL;
    } else {
      char* end = nullptr;
      double v = std::strtod(s.c_str() + i, &end);
      if (end == s.c_str() + i) fail("bad value");
      i = end - s.c_str();
      j.type = Json::NUM;
      j.num = v;
    }
    return j;
  }
};

}  // namespace

Json Json::parse(const std::string& text) {
  Parser p(text);
  Json j = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("trailing characters");
  return j;
}

Json Json::parse_file(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Unable to open file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

const Json& Json::operator[](const std::string& k) const {
  auto it = obj.find(k);
  if (type != OBJ || it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
  return it->second;
}

double Json::as_number() const {
  if (type == NUM) return num;
  if (type == BOOL) return b ? 1.0 : 0.0;
  throw std::runtime_error("json: not a number");
}

bool Json::as_bool() const {
  if (type == BOOL) return b;
  if (type == NUM) return num != 0;
  throw std::runtime_error("json: not a boolean");
}

const std::string& Json::as_string() const {
  if (type != STR) throw std::runtime_error("json: not a string");
  return str;
}

std::vector<double> Json::as_numbers() const {
  std::vector<double> v;
  if (type != ARR) throw std::runtime_error("json: not an array");
  for (const auto& e : arr) v.push_back(e.as_number());
  return v;
}

std::vector<std::string> Json::as_strings() const {
  std::vector<std::string> v;
  if (type != ARR) throw std::runtime_error("json: not an array");
  for (const auto& e : arr) v.push_back(e.as_string());
  return v;
}

}  // namespace mfem
