// Engine settings: every option name of the reference engine with its default
// (settings.zig:14-166), so a caller's JSON round-trips unchanged. Set by a JSON object
// (zc_configure): keys are field names, enums their names as strings, absent keys keep their
// value, unknown keys are an error (settings.zig:174-191). Only the SMI hot-path keys change what
// this library computes (metric, smi_features, exact, calibrated, symmetric, map_res,
// min_overlap, half, fft_sizes, clahe/band/invert for preprocessing checks, sw_*); the others
// are stored and reported.
#pragma once
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace zc {

enum class Kind { Bool, U32, F64, Enum };

struct SettingDef {
  const char* name;
  Kind kind;
  const char* def;       // default as JSON text (enums: the bare name)
  const char* choices;   // enums: names separated by '|'
};

// clang-format off
static const SettingDef kSettings[] = {
    {"metric", Kind::Enum, "smi", "smi|smi_edge|e4|lmax|ncc"},
    {"smi_features", Kind::Enum, "laplacian", "laplacian|three|cubic"},
    {"exact", Kind::Bool, "true", nullptr},
    {"calibrated", Kind::Bool, "true", nullptr},
    {"symmetric", Kind::Bool, "true", nullptr},
    {"group", Kind::Enum, "homography", "affine|homography"},
    {"map_res", Kind::U32, "0", nullptr},
    {"min_overlap", Kind::F64, "0.05", nullptr},
    {"rs_smin", Kind::F64, "0.2", nullptr},
    {"rs_smax", Kind::F64, "5", nullptr},
    {"ffd", Kind::Bool, "false", nullptr},
    {"ffd_grid", Kind::U32, "4", nullptr},
    {"ffd_source", Kind::Bool, "false", nullptr},
    {"ffd_stiffness", Kind::F64, "0.15", nullptr},
    {"hop", Kind::Bool, "true", nullptr},
    {"hop_fm", Kind::Bool, "false", nullptr},
    {"g_a0", Kind::F64, "0.08", nullptr},
    {"g_decay", Kind::F64, "0.5", nullptr},
    {"g_steps", Kind::U32, "7", nullptr},
    {"g_iters", Kind::U32, "100", nullptr},
    {"clahe", Kind::Bool, "true", nullptr},
    {"clahe_grid", Kind::U32, "4", nullptr},
    {"clahe_bins", Kind::U32, "64", nullptr},
    {"band", Kind::Bool, "false", nullptr},
    {"bp_fine", Kind::F64, "1", nullptr},
    {"bp_coarse", Kind::F64, "6", nullptr},
    {"invert", Kind::Bool, "false", nullptr},
    {"half", Kind::Bool, "true", nullptr},
    {"fft_sizes", Kind::Enum, "ladder", "ladder|compact"},
    // this library's: time the column-FFT tilings of a length at its first use, keep the fastest
    {"fft_autotune", Kind::Bool, "true", nullptr},
    {"detector", Kind::Enum, "pos_gift", "pos_gift|gls_mift"},
    {"n_octaves", Kind::U32, "3", nullptr},
    {"max_points", Kind::U32, "2000", nullptr},
    {"min_contrast", Kind::F64, "0.01", nullptr},
    {"tau", Kind::F64, "0.8", nullptr},
    {"radius", Kind::F64, "36", nullptr},
    {"nt", Kind::F64, "0", nullptr},
    {"auto_nt", Kind::Bool, "true", nullptr},
    {"second_ori", Kind::Bool, "true", nullptr},
    {"inlier_px", Kind::F64, "10", nullptr},
    {"scale_lo", Kind::F64, "0.2", nullptr},
    {"scale_hi", Kind::F64, "5", nullptr},
    {"max_aniso", Kind::F64, "5", nullptr},
    {"max_persp", Kind::F64, "3", nullptr},
    {"n_trials", Kind::U32, "100000", nullptr},
    {"seed", Kind::U32, "12345", nullptr},
    {"n_sigma", Kind::U32, "4", nullptr},
    {"n_angle", Kind::U32, "6", nullptr},
    {"n_r", Kind::U32, "3", nullptr},
    {"match_method", Kind::Enum, "lofsc", "lofsc|prosac|magsac"},
    {"per_octave", Kind::Bool, "true", nullptr},
    {"mutual", Kind::Bool, "false", nullptr},
    {"pg_n_octaves", Kind::U32, "3", nullptr},
    {"pg_max_points", Kind::U32, "5000", nullptr},
    {"pg_min_contrast", Kind::F64, "0.01", nullptr},
    {"pg_p1", Kind::F64, "10", nullptr},
    {"pg_search", Kind::Bool, "true", nullptr},
    {"pg_corners", Kind::Bool, "true", nullptr},
    {"pg_desc_pow", Kind::F64, "1.5", nullptr},
    {"pg_n_orient", Kind::U32, "6", nullptr},
    {"pg_n_rings", Kind::U32, "3", nullptr},
    {"pg_n_scales", Kind::U32, "4", nullptr},
    {"pg_min_wl", Kind::F64, "3", nullptr},
    {"pg_mult", Kind::F64, "1.6", nullptr},
    {"pg_sigma_onf", Kind::F64, "0.75", nullptr},
    {"pg_pc_k", Kind::F64, "1", nullptr},
    {"pg_cutoff", Kind::F64, "0.5", nullptr},
    {"pg_pc_g", Kind::F64, "3", nullptr},
    {"pg_gsig", Kind::F64, "3", nullptr},
    {"pg_pos_p1", Kind::F64, "20", nullptr},
    {"pg_pos_k", Kind::U32, "20", nullptr},
    {"pg_ratio", Kind::F64, "0.6", nullptr},
    {"pg_pos", Kind::Bool, "true", nullptr},
    {"pg_pos_search_px", Kind::F64, "10", nullptr},
    {"pg_pos_px", Kind::F64, "3", nullptr},
    {"pg_released", Kind::Bool, "false", nullptr},
    {"search_mode", Kind::Enum, "sweep", "sweep|cloud"},
    {"sw_grid", Kind::U32, "128", nullptr},
    {"sw_nth", Kind::U32, "90", nullptr},
    {"sw_rot", Kind::F64, "180", nullptr},
    {"sw_rot_range", Kind::Bool, "false", nullptr},
    {"sw_rot_lo", Kind::F64, "-180", nullptr},
    {"sw_rot_hi", Kind::F64, "180", nullptr},
    {"sw_shear", Kind::Bool, "false", nullptr},
    {"sw_shear_max", Kind::F64, "0.1", nullptr},
    {"sw_nshear", Kind::U32, "3", nullptr},
    {"sw_aniso", Kind::Bool, "false", nullptr},
    {"sw_aniso_max", Kind::F64, "0.1", nullptr},
    {"sw_naniso", Kind::U32, "3", nullptr},
    {"sw_ns", Kind::U32, "20", nullptr},
    {"sw_smin", Kind::F64, "0.5", nullptr},
    {"sw_smax", Kind::F64, "2", nullptr},
    {"sw_k", Kind::U32, "32", nullptr},
    {"sw_fin", Kind::U32, "4", nullptr},
    {"sw_rel", Kind::Bool, "false", nullptr},
    // not a reference option: the sweep at one canvas cell per fixed pixel (sw_grid ignored)
    {"sw_full", Kind::Bool, "false", nullptr},
    // this library's: replay the sweep's launches from a captured HIP graph (eager while profiling)
    {"sweep_graph", Kind::Bool, "true", nullptr},
    // this library's: sweep planes of 1536, 2304 or 3072 points (2^a·3^b) where they are
    // shorter than the next power of two (false: powers of two only); results do not depend on it
    {"sweep_smooth", Kind::Bool, "true", nullptr},
    // this library's: run Gauss-Newton batches (stacks, the search's refinement) on the device, one
    // captured HIP graph per iteration, no host round trip (false: the host-driven loop)
    {"gn_graph", Kind::Bool, "true", nullptr},
    // this library's: run the climb's decisions on the device, one captured HIP graph per
    // iteration read back once per iteration (false: the host-driven loop)
    {"climb_graph", Kind::Bool, "true", nullptr},
    {"hho_auto", Kind::Bool, "true", nullptr},
    {"hho_n", Kind::U32, "100", nullptr},
    {"hho_t", Kind::U32, "20", nullptr},
    {"hho_sg", Kind::F64, "0.01", nullptr},
    {"hho_sh", Kind::F64, "0.01", nullptr},
    {"hho_p", Kind::F64, "0.0001", nullptr},
    {"hho_tx", Kind::F64, "0.01", nullptr},
    {"hho_th", Kind::F64, "0.01", nullptr},
    {"hho_phi", Kind::F64, "0.01", nullptr},
    {"hho_seed", Kind::U32, "1", nullptr},
    {"hho_rel", Kind::Bool, "true", nullptr},
};
// clang-format on

// A value: numbers as double, bools as 0/1, enums as strings.
struct SVal {
  double num = 0;
  std::string str;
};

class Settings {
 public:
  Settings() {
    for (const auto& d : kSettings) {
      SVal v;
      if (d.kind == Kind::Enum) v.str = d.def;
      else if (d.kind == Kind::Bool) v.num = std::strcmp(d.def, "true") == 0 ? 1 : 0;
      else v.num = std::stod(d.def);
      vals_[d.name] = v;
    }
  }

  bool b(const char* k) const { return vals_.at(k).num != 0; }
  uint32_t u(const char* k) const { return (uint32_t)vals_.at(k).num; }
  double f(const char* k) const { return vals_.at(k).num; }
  const std::string& e(const char* k) const { return vals_.at(k).str; }
  void set_bool(const char* k, bool v) { vals_[k].num = v ? 1 : 0; }

  // Apply a JSON object of overrides; all-or-nothing.
  void apply(const char* json, size_t len) {
    Parser p{json, json + len};
    std::map<std::string, SVal> next = vals_;
    p.ws();
    p.expect('{');
    p.ws();
    if (p.peek() == '}') {
      ++p.s;
      return;
    }
    while (true) {
      p.ws();
      const std::string key = p.string();
      p.ws();
      p.expect(':');
      p.ws();
      const SettingDef* d = find(key);
      if (!d) throw std::runtime_error("UnknownSetting: " + key);
      SVal v;
      if (d->kind == Kind::Enum) {
        v.str = p.string();
        if (key == "smi_features" && v.str == "none") v.str = "three";
        if (!has_choice(d->choices, v.str)) throw std::runtime_error("InvalidEnumTag: " + key + "=" + v.str);
      } else if (d->kind == Kind::Bool) {
        v.num = p.boolean() ? 1 : 0;
      } else {
        v.num = p.number();
        if (d->kind == Kind::U32 && (v.num < 0 || v.num != (double)(uint32_t)v.num))
          throw std::runtime_error("InvalidNumber: " + key);
      }
      next[key] = v;
      p.ws();
      if (p.peek() == ',') {
        ++p.s;
        continue;
      }
      p.expect('}');
      break;
    }
    vals_ = std::move(next);
  }

  std::string to_json() const {
    std::string o = "{";
    bool first = true;
    for (const auto& d : kSettings) {
      if (!first) o += ",";
      first = false;
      o += "\"";
      o += d.name;
      o += "\":";
      const SVal& v = vals_.at(d.name);
      if (d.kind == Kind::Enum) o += "\"" + v.str + "\"";
      else if (d.kind == Kind::Bool) o += v.num != 0 ? "true" : "false";
      else if (d.kind == Kind::U32) o += std::to_string((uint32_t)v.num);
      else {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", v.num);
        std::string s = buf;
        if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
        o += s;
      }
    }
    return o + "}";
  }

 private:
  std::map<std::string, SVal> vals_;

  static const SettingDef* find(const std::string& k) {
    for (const auto& d : kSettings)
      if (k == d.name) return &d;
    return nullptr;
  }
  static bool has_choice(const char* choices, const std::string& v) {
    std::string c = choices;
    size_t a = 0;
    while (a <= c.size()) {
      size_t b = c.find('|', a);
      if (b == std::string::npos) b = c.size();
      if (c.substr(a, b - a) == v) return true;
      a = b + 1;
    }
    return false;
  }

  struct Parser {
    const char* s;
    const char* end;
    char peek() const { return s < end ? *s : '\0'; }
    void ws() {
      while (s < end && (*s == ' ' || *s == '\n' || *s == '\t' || *s == '\r')) ++s;
    }
    void expect(char c) {
      if (peek() != c) throw std::runtime_error(std::string("SyntaxError: expected '") + c + "'");
      ++s;
    }
    std::string string() {
      expect('"');
      std::string o;
      while (s < end && *s != '"') {
        if (*s == '\\' && s + 1 < end) ++s;
        o += *s++;
      }
      expect('"');
      return o;
    }
    bool boolean() {
      if (end - s >= 4 && std::strncmp(s, "true", 4) == 0) {
        s += 4;
        return true;
      }
      if (end - s >= 5 && std::strncmp(s, "false", 5) == 0) {
        s += 5;
        return false;
      }
      throw std::runtime_error("UnexpectedToken: expected a bool");
    }
    double number() {
      const char* b = s;
      while (s < end && (std::strchr("+-0123456789.eE", *s) != nullptr)) ++s;
      if (s == b) throw std::runtime_error("UnexpectedToken: expected a number");
      return std::stod(std::string(b, s));
    }
  };
};

}  // namespace zc
