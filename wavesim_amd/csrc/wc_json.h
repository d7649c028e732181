// JSON values as the run header holds them, printed byte for byte as the reference's writer
// prints them (serde_json::to_string_pretty, crates/waveio/src/run.rs:199-212):
//   - objects given as values (`waves`, `source`, notes) keep their keys sorted by bytes, as
//     serde_json's default map (a BTreeMap) does; the header's own fields keep their order;
//   - two-space indentation, "key": value, empty arrays and objects as [] and {};
//   - integers as integers; floats in the shortest digits that read back to the same double,
//     laid out as the ryu crate's format64 lays them out (1.0, 0.001, 1e-7, 1.5e300, -0.0);
//     non-finite floats as null;
//   - strings escape only '"', '\\' and the control characters below 0x20.
#pragma once

#include <stdint.h>

#include <string.h>

#include <charconv>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace wc::json {

struct Value {
  enum Kind { Null, Bool, PosInt, NegInt, Float, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  uint64_t u = 0;
  int64_t i = 0;
  double f = 0.0;
  std::string s;
  std::vector<Value> a;
  std::map<std::string, Value> o;  // std::string orders by unsigned bytes, as Rust's String

  static Value null() { return Value{}; }
  static Value boolean(bool v) {
    Value x;
    x.kind = Bool;
    x.b = v;
    return x;
  }
  static Value uint(uint64_t v) {
    Value x;
    x.kind = PosInt;
    x.u = v;
    return x;
  }
  static Value number(double v) {  // serde_json::Value::from(f64): non-finite is null
    Value x;
    if (std::isfinite(v)) {
      x.kind = Float;
      x.f = v;
    }
    return x;
  }
  static Value string(std::string v) {
    Value x;
    x.kind = String;
    x.s = std::move(v);
    return x;
  }
  static Value object() {
    Value x;
    x.kind = Object;
    return x;
  }
};

// ---- numbers ----------------------------------------------------------------------------------

// The ryu crate's format64: the shortest round-trip digits d1..dn with v = 0.d1..dn * 10^kk,
// integers up to 16 digits as "123.0", decimals with the point inside up to 16 digits before
// it, small numbers down to 1e-5 as "0.000123", everything else as "1.23e-7" / "1e30".
inline std::string format_f64(double v) {
  if (!std::isfinite(v)) return "null";
  std::string out = std::signbit(v) ? "-" : "";
  if (v == 0.0) return out + "0.0";
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf, std::fabs(v), std::chars_format::scientific);
  const std::string sci(buf, r.ptr);  // d[.ddd]e(+|-)XX, shortest round-trip digits
  const size_t e = sci.find('e');
  std::string digits;
  for (size_t q = 0; q < e; ++q)
    if (sci[q] != '.') digits += sci[q];
  const int length = (int)digits.size();
  const int kk = std::stoi(sci.substr(e + 1)) + 1;
  const int k = kk - length;
  if (0 <= k && kk <= 16) {
    out += digits + std::string(k, '0') + ".0";
  } else if (0 < kk && kk <= 16) {
    out += digits.substr(0, kk) + "." + digits.substr(kk);
  } else if (-5 < kk && kk <= 0) {
    out += "0." + std::string(-kk, '0') + digits;
  } else if (length == 1) {
    out += digits + "e" + std::to_string(kk - 1);
  } else {
    out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
  }
  return out;
}

// ---- printing -----------------------------------------------------------------------------------

inline void put_string(std::string& out, const std::string& s) {
  static const char* hex = "0123456789abcdef";
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      default:
        if (c < 0x20) {
          out += "\\u00";
          out += hex[c >> 4];
          out += hex[c & 15];
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

inline void indent(std::string& out, int level) { out.append(2 * (size_t)level, ' '); }

inline void put(std::string& out, const Value& v, int level) {
  switch (v.kind) {
    case Value::Null: out += "null"; return;
    case Value::Bool: out += v.b ? "true" : "false"; return;
    case Value::PosInt: out += std::to_string(v.u); return;
    case Value::NegInt: out += std::to_string(v.i); return;
    case Value::Float: out += format_f64(v.f); return;
    case Value::String: put_string(out, v.s); return;
    case Value::Array: {
      if (v.a.empty()) {
        out += "[]";
        return;
      }
      out += "[\n";
      for (size_t q = 0; q < v.a.size(); ++q) {
        indent(out, level + 1);
        put(out, v.a[q], level + 1);
        out += q + 1 < v.a.size() ? ",\n" : "\n";
      }
      indent(out, level);
      out += ']';
      return;
    }
    case Value::Object: {
      if (v.o.empty()) {
        out += "{}";
        return;
      }
      out += "{\n";
      size_t q = 0;
      for (const auto& kv : v.o) {
        indent(out, level + 1);
        put_string(out, kv.first);
        out += ": ";
        put(out, kv.second, level + 1);
        out += ++q < v.o.size() ? ",\n" : "\n";
      }
      indent(out, level);
      out += '}';
      return;
    }
  }
}

// A struct printed in its fields' order (serde's derive), with serde_json's pretty layout.
struct Fields {
  std::string out;
  int level;
  bool first = true;
  explicit Fields(int level_) : level(level_) { out = "{\n"; }
  void key(const char* k) {
    if (!first) out += ",\n";
    first = false;
    indent(out, level + 1);
    put_string(out, k);
    out += ": ";
  }
  void value(const char* k, const Value& v) {
    key(k);
    put(out, v, level + 1);
  }
  void raw(const char* k, const std::string& text) {
    key(k);
    out += text;
  }
  std::string close() {
    out += "\n";
    indent(out, level);
    out += "}";
    return out;
  }
};

// ---- parsing (RFC 8259, as serde_json reads it) --------------------------------------------

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Parser {
 public:
  explicit Parser(const std::string& text) : t_(text) {}

  Value document() {
    Value v = value(0);
    ws();
    if (p_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) {
    throw ParseError(std::string(what) + " at byte " + std::to_string(p_));
  }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r'))
      ++p_;
  }
  bool take(const char* lit) {
    const size_t n = std::char_traits<char>::length(lit);
    if (t_.compare(p_, n, lit) != 0) return false;
    p_ += n;
    return true;
  }
  Value value(int depth) {
    if (depth > 128) fail("recursion limit exceeded");
    ws();
    if (p_ >= t_.size()) fail("EOF while parsing a value");
    const char c = t_[p_];
    if (c == '{') return object(depth);
    if (c == '[') return array(depth);
    if (c == '"') return Value::string(string());
    if (take("true")) return Value::boolean(true);
    if (take("false")) return Value::boolean(false);
    if (take("null")) return Value::null();
    if (c == '-' || (c >= '0' && c <= '9')) return number();
    fail("expected value");
  }
  Value object(int depth) {
    ++p_;
    Value v = Value::object();
    ws();
    if (p_ < t_.size() && t_[p_] == '}') {
      ++p_;
      return v;
    }
    for (;;) {
      ws();
      if (p_ >= t_.size() || t_[p_] != '"') fail("key must be a string");
      std::string k = string();
      ws();
      if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
      ++p_;
      v.o[k] = value(depth + 1);  // a repeated key keeps its last value, as a map insert does
      ws();
      if (p_ < t_.size() && t_[p_] == ',') {
        ++p_;
        continue;
      }
      if (p_ < t_.size() && t_[p_] == '}') {
        ++p_;
        return v;
      }
      fail("expected ',' or '}'");
    }
  }
  Value array(int depth) {
    ++p_;
    Value v;
    v.kind = Value::Array;
    ws();
    if (p_ < t_.size() && t_[p_] == ']') {
      ++p_;
      return v;
    }
    for (;;) {
      v.a.push_back(value(depth + 1));
      ws();
      if (p_ < t_.size() && t_[p_] == ',') {
        ++p_;
        continue;
      }
      if (p_ < t_.size() && t_[p_] == ']') {
        ++p_;
        return v;
      }
      fail("expected ',' or ']'");
    }
  }
  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (p_ + 4 > t_.size()) fail("EOF while parsing a string");
    uint32_t v = 0;
    for (int q = 0; q < 4; ++q) {
      const char c = t_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else fail("invalid escape");
    }
    return v;
  }
  std::string string() {
    ++p_;
    std::string out;
    for (;;) {
      if (p_ >= t_.size()) fail("EOF while parsing a string");
      const unsigned char c = (unsigned char)t_[p_++];
      if (c == '"') return out;
      if (c < 0x20) fail("control character in a string");
      if (c != '\\') {
        out += (char)c;
        continue;
      }
      if (p_ >= t_.size()) fail("EOF while parsing a string");
      const char e = t_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (!take("\\u")) fail("lone leading surrogate in hex escape");
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("lone leading surrogate in hex escape");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            fail("lone trailing surrogate in hex escape");
          }
          utf8(out, cp);
          break;
        }
        default: fail("invalid escape");
      }
    }
  }
  Value number() {
    const size_t start = p_;
    const bool neg = t_[p_] == '-';
    if (neg) ++p_;
    if (p_ >= t_.size() || t_[p_] < '0' || t_[p_] > '9') fail("invalid number");
    if (t_[p_] == '0' && p_ + 1 < t_.size() && t_[p_ + 1] >= '0' && t_[p_ + 1] <= '9')
      fail("invalid number");
    while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    bool is_float = false;
    if (p_ < t_.size() && t_[p_] == '.') {
      is_float = true;
      ++p_;
      if (p_ >= t_.size() || t_[p_] < '0' || t_[p_] > '9') fail("invalid number");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    }
    if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
      is_float = true;
      ++p_;
      if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
      if (p_ >= t_.size() || t_[p_] < '0' || t_[p_] > '9') fail("invalid number");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
    }
    const std::string text = t_.substr(start, p_ - start);
    if (!is_float) {
      // integers that fit u64 (or i64 when negative) stay integers; -0 is the float -0.0
      const char* b = text.data() + (neg ? 1 : 0);
      uint64_t mag = 0;
      const auto r = std::from_chars(b, text.data() + text.size(), mag);
      if (r.ec == std::errc()) {
        Value v;
        if (!neg) {
          v.kind = Value::PosInt;
          v.u = mag;
          return v;
        }
        if (mag == 0) return Value::number(-0.0);
        if (mag <= (uint64_t)INT64_MAX + 1) {
          v.kind = Value::NegInt;
          v.i = mag == (uint64_t)INT64_MAX + 1 ? INT64_MIN : -(int64_t)mag;
          return v;
        }
      }
    }
    const double d = std::strtod(text.c_str(), nullptr);
    if (!std::isfinite(d)) fail("number out of range");
    Value v;
    v.kind = Value::Float;
    v.f = d;
    return v;
  }
};

inline Value parse(const std::string& text) { return Parser(text).document(); }

// f64::total_cmp's key: the bits, with the magnitude bits flipped for negative values.
inline int64_t total_order(double v) {
  int64_t b;
  memcpy(&b, &v, sizeof b);
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

}  // namespace wc::json
