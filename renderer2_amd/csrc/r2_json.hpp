// Minimal JSON document model for the front end: parse (with // and /* */
// comments, as the reference's config files use them) and write.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace r2json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;  // insertion order kept

  bool is_null() const { return kind == Null; }
  bool is_number() const { return kind == Number; }
  bool is_string() const { return kind == String; }
  bool is_array() const { return kind == Array; }
  bool is_object() const { return kind == Object; }
  bool is_bool() const { return kind == Bool; }
  const Value* find(const std::string& k) const {
    if (kind != Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  Value& set(const std::string& k) {
    kind = Object;
    for (auto& kv : obj)
      if (kv.first == k) return kv.second;
    obj.emplace_back(k, Value());
    return obj.back().second;
  }
  static Value number(double v) {
    Value x;
    x.kind = Number;
    x.num = v;
    return x;
  }
  static Value string(const std::string& s) {
    Value x;
    x.kind = String;
    x.str = s;
    return x;
  }
  static Value array() {
    Value x;
    x.kind = Array;
    return x;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& t) : s_(t) {}
  Value parse() {
    Value v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const char* what) const {
    size_t line = 1;
    for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
    throw std::runtime_error(std::string("json: ") + what + " at line " + std::to_string(line));
  }
  void ws() {
    for (;;) {
      while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
      if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
        while (i_ < s_.size() && s_[i_] != '\n') ++i_;
        continue;
      }
      if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
        const size_t e = s_.find("*/", i_ + 2);
        if (e == std::string::npos) fail("unterminated comment");
        i_ = e + 2;
        continue;
      }
      break;
    }
  }
  Value value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') return Value::string(string());
    if (s_.compare(i_, 4, "true") == 0) {
      i_ += 4;
      Value v;
      v.kind = Value::Bool;
      v.b = true;
      return v;
    }
    if (s_.compare(i_, 5, "false") == 0) {
      i_ += 5;
      Value v;
      v.kind = Value::Bool;
      return v;
    }
    if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
      return Value();
    }
    return number();
  }
  Value number() {
    const char* b = s_.c_str() + i_;
    char* e = nullptr;
    const double v = strtod(b, &e);
    if (e == b) fail("bad value");
    i_ += (size_t)(e - b);
    return Value::number(v);
  }
  std::string string() {
    ++i_;  // opening quote
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\') {
        if (i_ >= s_.size()) fail("bad escape");
        const char e = s_[i_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {
            if (i_ + 4 > s_.size()) fail("bad \\u escape");
            const unsigned cp = (unsigned)strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
            i_ += 4;
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
            continue;
          }
          default: c = e;  // \" \\ \/
        }
      }
      out += c;
    }
    if (i_ >= s_.size()) fail("unterminated string");
    ++i_;
    return out;
  }
  Value array() {
    ++i_;
    Value v = Value::array();
    ws();
    if (i_ < s_.size() && s_[i_] == ']') {
      ++i_;
      return v;
    }
    for (;;) {
      v.arr.push_back(value());
      ws();
      if (i_ < s_.size() && s_[i_] == ',') {
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == ']') {  // trailing comma
          ++i_;
          return v;
        }
        continue;
      }
      if (i_ < s_.size() && s_[i_] == ']') {
        ++i_;
        return v;
      }
      fail("expected , or ]");
    }
  }
  Value object() {
    ++i_;
    Value v;
    v.kind = Value::Object;
    ws();
    if (i_ < s_.size() && s_[i_] == '}') {
      ++i_;
      return v;
    }
    for (;;) {
      ws();
      if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
      std::string k = string();
      ws();
      if (i_ >= s_.size() || s_[i_] != ':') fail("expected :");
      ++i_;
      v.obj.emplace_back(k, value());
      ws();
      if (i_ < s_.size() && s_[i_] == ',') {
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == '}') {
          ++i_;
          return v;
        }
        continue;
      }
      if (i_ < s_.size() && s_[i_] == '}') {
        ++i_;
        return v;
      }
      fail("expected , or }");
    }
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

inline void write(const Value& v, std::string& out, int indent = 0) {
  const std::string pad((size_t)indent, ' ');
  switch (v.kind) {
    case Value::Null: out += "null"; break;
    case Value::Bool: out += v.b ? "true" : "false"; break;
    case Value::Number: {
      char buf[64];
      snprintf(buf, sizeof(buf), "%.9g", v.num);
      out += buf;
      break;
    }
    case Value::String: out += "\"" + v.str + "\""; break;
    case Value::Array:
      out += "[";
      for (size_t k = 0; k < v.arr.size(); ++k) {
        if (k) out += ", ";
        write(v.arr[k], out, indent);
      }
      out += "]";
      break;
    case Value::Object:
      out += "{\n";
      for (size_t k = 0; k < v.obj.size(); ++k) {
        out += pad + "  \"" + v.obj[k].first + "\": ";
        write(v.obj[k].second, out, indent + 2);
        out += (k + 1 < v.obj.size()) ? ",\n" : "\n";
      }
      out += pad + "}";
      break;
  }
}

}  // namespace r2json
