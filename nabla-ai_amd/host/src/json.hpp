// Minimal JSON reader for case files: objects, arrays, strings, numbers,
// booleans and null.  Lookups mirror the reference's accessors
// (core/include/nabla/io/json.hpp:45-60): number/string/integer/boolean with a
// fallback when the key is absent or of another type; integer() truncates.
#pragma once

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace nabla::json {

struct Value {
  enum class Kind { Null, Bool, Number, String, Array, Object } kind = Kind::Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  const Value* find(const std::string& k) const {
    if (kind != Kind::Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  double number(const std::string& k, double fb) const {
    const Value* v = find(k);
    return (v && v->kind == Kind::Number) ? v->num : fb;
  }
  int integer(const std::string& k, int fb) const {
    const Value* v = find(k);
    return (v && v->kind == Kind::Number) ? static_cast<int>(v->num) : fb;
  }
  bool boolean(const std::string& k, bool fb) const {
    const Value* v = find(k);
    return (v && v->kind == Kind::Bool) ? v->b : fb;
  }
  std::string string(const std::string& k, const std::string& fb) const {
    const Value* v = find(k);
    return (v && v->kind == Kind::String) ? v->str : fb;
  }
};

class Reader {
 public:
  explicit Reader(const std::string& text) : s_(text) {}
  Value document() {
    Value v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  [[noreturn]] void fail(const char* why) const {
    throw std::runtime_error(std::string("json: ") + why + " at offset " + std::to_string(i_));
  }
  void ws() {
    while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
  }
  bool eat(char c) {
    ws();
    if (i_ < s_.size() && s_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  Value value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    Value v;
    if (c == '{') {
      ++i_;
      v.kind = Value::Kind::Object;
      if (eat('}')) return v;
      do {
        ws();
        std::string k = text();
        if (!eat(':')) fail("expected ':'");
        v.obj.emplace_back(std::move(k), value());
      } while (eat(','));
      if (!eat('}')) fail("expected '}'");
    } else if (c == '[') {
      ++i_;
      v.kind = Value::Kind::Array;
      if (eat(']')) return v;
      do v.arr.push_back(value());
      while (eat(','));
      if (!eat(']')) fail("expected ']'");
    } else if (c == '"') {
      v.kind = Value::Kind::String;
      v.str = text();
    } else if (s_.compare(i_, 4, "true") == 0) {
      i_ += 4;
      v.kind = Value::Kind::Bool;
      v.b = true;
    } else if (s_.compare(i_, 5, "false") == 0) {
      i_ += 5;
      v.kind = Value::Kind::Bool;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
    } else {
      const char* b = s_.c_str() + i_;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("bad value");
      i_ += size_t(e - b);
      v.kind = Value::Kind::Number;
    }
    return v;
  }
  std::string text() {
    if (i_ >= s_.size() || s_[i_] != '"') fail("expected string");
    ++i_;
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\' && i_ < s_.size()) {
        const char e = s_[i_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {
            if (i_ + 4 > s_.size()) fail("bad escape");
            const unsigned cp = unsigned(std::stoul(s_.substr(i_, 4), nullptr, 16));
            i_ += 4;
            if (cp < 0x80) {
              out += char(cp);
            } else if (cp < 0x800) {
              out += char(0xC0 | (cp >> 6));
              out += char(0x80 | (cp & 0x3F));
            } else {
              out += char(0xE0 | (cp >> 12));
              out += char(0x80 | ((cp >> 6) & 0x3F));
              out += char(0x80 | (cp & 0x3F));
            }
            continue;
          }
          default: c = e;
        }
      }
      out += c;
    }
    if (i_ >= s_.size()) fail("unterminated string");
    ++i_;
    return out;
  }
  const std::string& s_;
  size_t i_ = 0;
};

inline Value parseFile(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("json: cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  return Reader(text).document();
}

}  // namespace nabla::json
