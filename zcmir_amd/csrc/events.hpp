// Progress events of long operations (searches, climbs) as JSON lines, queued until the caller
// drains them with zc_events (the reference's event kinds and fields: src/events.zig:1-86,
// emitted at src/search.zig:655-660, 739-833 and src/gclimb.zig:446-486):
//   {"kind":"log","text":...}
//   {"kind":"trail","label":...,"score":...,"fwd":...,"inv":...|null,"H":[9],"cps":null}
//   {"kind":"pose","H":[9],"cps":null}
//   {"kind":"sweep","thetas":[...],"sigmas":[...],"wrap":...,"best":[...],"top":[[it,is],...],
//    "finalists":[...],"result":[th,s]|null,"cands":[{"H":[9],"score":...,"peak":...,"th":...,"s":...}],
//    "applied":0|-1}
#pragma once
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace zc {

// A JSON object written field by field.
class JsonObj {
 public:
  JsonObj& num(const char* k, double v) { return raw(k, fmt_num(v)); }
  JsonObj& integer(const char* k, long long v) { return raw(k, std::to_string(v)); }
  JsonObj& boolean(const char* k, bool v) { return raw(k, v ? "true" : "false"); }
  JsonObj& str(const char* k, const std::string& v) { return raw(k, quote(v)); }
  JsonObj& null(const char* k) { return raw(k, "null"); }
  JsonObj& nums(const char* k, const double* v, size_t n) {
    std::string a = "[";
    for (size_t i = 0; i < n; ++i) {
      if (i) a += ',';
      a += fmt_num(v[i]);
    }
    return raw(k, a + "]");
  }
  JsonObj& nums(const char* k, const std::vector<double>& v) { return nums(k, v.data(), v.size()); }
  JsonObj& mat(const char* k, const Mat3& H) { return nums(k, H.data(), 9); }
  // a value already in JSON form (an array or object)
  JsonObj& raw(const char* k, const std::string& json) {
    body_ += body_.empty() ? "{" : ",";
    body_ += quote(k) + ":" + json;
    return *this;
  }
  std::string done() const { return body_.empty() ? "{}" : body_ + "}"; }

  // non-finite numbers become null (JSON has no NaN or infinity)
  static std::string fmt_num(double v) {
    if (!std::isfinite(v)) return "null";
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
  }
  static std::string quote(const std::string& s) {
    std::string o = "\"";
    for (const unsigned char c : s) {
      switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
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

 private:
  std::string body_;
};

class Events {
 public:
  void emit(const JsonObj& o) {
    q_ += o.done();
    q_ += '\n';
  }
  void log(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
    char text[768];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(text, sizeof text, fmt, ap);
    va_end(ap);
    emit(JsonObj().str("kind", "log").str("text", text));
  }
  // inv: NaN when the score is one-directional
  void trail(const char* label, double score, double fwd, double inv, const Mat3& H) {
    JsonObj o;
    o.str("kind", "trail").str("label", label).num("score", score).num("fwd", fwd).num("inv", inv).mat("H", H);
    emit(o.null("cps"));
  }
  void pose(const Mat3& H) { emit(JsonObj().str("kind", "pose").mat("H", H).null("cps")); }

  // Whole lines, at most cap bytes, removed from the queue; returns the bytes copied (0 when the
  // queue is empty or its first line does not fit).
  size_t take(char* buf, size_t cap) {
    size_t end = 0;
    for (size_t i = q_.find('\n'); i != std::string::npos && i + 1 <= cap; i = q_.find('\n', i + 1)) end = i + 1;
    if (end && buf) std::memcpy(buf, q_.data(), end);
    q_.erase(0, end);
    return end;
  }
  size_t pending() const { return q_.size(); }
  void clear() { q_.clear(); }

 private:
  std::string q_;
};

}  // namespace zc
