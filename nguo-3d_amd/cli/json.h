// A small JSON reader and the output formatting the mguso process contract uses: a DOM of
// null / bool / number / string / array / object, parsed strictly (RFC 8259), with errors that
// name the byte offset. Numbers keep their text so integers can be checked exactly.
#pragma once

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mguso {

struct Json {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string text;  // the number's literal, or the string's value
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  // The member `key` of an object (the last one when repeated, as serde keeps), or null.
  const Json *get(const std::string &key) const {
    const Json *r = nullptr;
    for (auto &kv : obj)
      if (kv.first == key) r = &kv.second;
    return r;
  }
};

class JsonParser {
 public:
  explicit JsonParser(const std::string &s) : s_(s) {}
  Json parse() {
    Json v = value(0);
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string &what) const {
    throw std::runtime_error(what + " at byte " + std::to_string(i_));
  }
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
  }
  bool lit(const char *w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
    return false;
  }
  Json value(int depth) {
    if (depth > 512) fail("nesting too deep");
    ws();
    if (i_ >= s_.size()) fail("unexpected end of input");
    Json v;
    char c = s_[i_];
    if (c == '{') {
      v.kind = Json::Obj;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
      for (;;) {
        ws();
        if (i_ >= s_.size() || s_[i_] != '"') fail("expected a key");
        std::string k = string();
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
        ++i_;
        v.obj.emplace_back(std::move(k), value(depth + 1));
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Json::Arr;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
      for (;;) {
        v.arr.push_back(value(depth + 1));
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') { v.kind = Json::Str; v.text = string(); return v; }
    if (lit("true")) { v.kind = Json::Bool; v.b = true; return v; }
    if (lit("false")) { v.kind = Json::Bool; v.b = false; return v; }
    if (lit("null")) return v;
    if (c == '-' || (c >= '0' && c <= '9')) return number();
    fail("unexpected character");
  }
  Json number() {
    size_t b = i_;
    if (s_[i_] == '-') ++i_;
    auto digits = [&]() {
      size_t d = i_;
      while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
      if (i_ == d) fail("malformed number");
    };
    if (i_ < s_.size() && s_[i_] == '0') ++i_;
    else digits();
    if (i_ < s_.size() && s_[i_] == '.') { ++i_; digits(); }
    if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
      ++i_;
      if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
      digits();
    }
    Json v;
    v.kind = Json::Num;
    v.text = s_.substr(b, i_ - b);
    v.num = std::strtod(v.text.c_str(), nullptr);
    return v;
  }
  static void utf8(std::string &o, uint32_t cp) {
    if (cp < 0x80) o += (char)cp;
    else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) { o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
    else {
      o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (i_ + 4 > s_.size()) fail("truncated \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      char h = s_[i_++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
      else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string() {
    ++i_;  // opening quote
    std::string o;
    for (;;) {
      if (i_ >= s_.size()) fail("unterminated string");
      char c = s_[i_++];
      if (c == '"') return o;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') { o += c; continue; }
      if (i_ >= s_.size()) fail("unterminated escape");
      char e = s_[i_++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (!(i_ + 1 < s_.size() && s_[i_] == '\\' && s_[i_ + 1] == 'u')) fail("lone surrogate");
            i_ += 2;
            uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            fail("lone surrogate");
          }
          utf8(o, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
};

// ---------------------------------------------------------------- output formatting
// Shortest text that reads back to the same value (the form serde_json writes), with ".0"
// on integral values; non-finite values become null, as serde_json writes them.
inline std::string num_f64(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v);
  std::string s(buf, r.ptr);
  if (s.find_first_of(".en") == std::string::npos) s += ".0";
  return s;
}
inline std::string num_f32(float v) {
  if (!std::isfinite(v)) return "null";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v);
  std::string s(buf, r.ptr);
  if (s.find_first_of(".en") == std::string::npos) s += ".0";
  return s;
}
// Rust's `{:.Ne}`: mantissa with N decimals, 'e', exponent without sign padding ("1.500e-7").
inline std::string rust_exp(double v, int prec) {
  if (!std::isfinite(v)) return std::isnan(v) ? "NaN" : (v > 0 ? "inf" : "-inf");
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*e", prec, v);
  std::string s(buf);
  size_t e = s.find('e');
  std::string mant = s.substr(0, e), ex = s.substr(e + 1);
  bool neg = ex[0] == '-';
  size_t k = (ex[0] == '+' || ex[0] == '-') ? 1 : 0;
  while (k + 1 < ex.size() && ex[k] == '0') ++k;
  return mant + "e" + (neg ? "-" : "") + ex.substr(k);
}
// Rust's `{:.N}` for a float: fixed with N decimals.
inline std::string rust_fixed(double v, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*f", prec, v);
  return buf;
}
inline std::string quote(const std::string &s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case '\b': o += "\\b"; break;
      case '\f': o += "\\f"; break;
      default:
        if (c < 0x20) {
          char b[8];
          std::snprintf(b, sizeof b, "\\u%04x", c);
          o += b;
        } else {
          o += (char)c;
        }
    }
  }
  return o + "\"";
}

// A pretty-printed JSON object with two-space indentation and keys in sorted order, the
// layout serde_json::to_string_pretty gives a json! map. Values are pre-rendered JSON text.
struct PrettyObject {
  std::vector<std::pair<std::string, std::string>> kv;
  void put(const std::string &k, const std::string &json_text) { kv.emplace_back(k, json_text); }
  std::string render(int indent = 0) const {
    auto sorted = kv;
    std::stable_sort(sorted.begin(), sorted.end(), [](auto &a, auto &b) { return a.first < b.first; });
    if (sorted.empty()) return "{}";
    std::string pad(indent + 2, ' '), o = "{\n";
    for (size_t k = 0; k < sorted.size(); ++k) {
      o += pad + quote(sorted[k].first) + ": " + sorted[k].second;
      o += k + 1 < sorted.size() ? ",\n" : "\n";
    }
    return o + std::string(indent, ' ') + "}";
  }
};

// A pretty-printed array of pre-rendered items at `indent`.
inline std::string pretty_array(const std::vector<std::string> &items, int indent) {
  if (items.empty()) return "[]";
  std::string pad(indent + 2, ' '), o = "[\n";
  for (size_t k = 0; k < items.size(); ++k) {
    o += pad + items[k];
    o += k + 1 < items.size() ? ",\n" : "\n";
  }
  return o + std::string(indent, ' ') + "]";
}

}  // namespace mguso
