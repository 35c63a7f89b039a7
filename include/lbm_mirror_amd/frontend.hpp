// frontend.hpp — the reference drivers' unit and configuration front end, so
// their parameter sets are consumed unchanged by C++ drivers of this engine:
//
//   lbmx::converter<T>     converter.hpp:15-93       SI <-> lattice units, tau from nu
//   lbmx::args             abstract/args.hpp:43-340  "--name value" pairs + positionals
//   lbmx::load_lbm_config  config.hpp:14-51          {delta_x, kinematic_viscosity, delta_t}
//   lbmx::init_lbm_env     environment.hpp:9-56      ~/.lbm/data/<name>
//   lbmx::output_directory environment.hpp:58-89
//
// Header only, C++17, host code.  Arithmetic happens in T with one rounding
// per operation in the reference's order (compile without FMA contraction,
// e.g. -std=c++17 or -ffp-contract=off), so tau and lattice quantities equal
// the reference driver's bit for bit.  Failures come back as an error with
// the reference's category and message text.
#ifndef LBM_MIRROR_AMD_FRONTEND_HPP
#define LBM_MIRROR_AMD_FRONTEND_HPP

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <string_view>
#include <variant>
#include <vector>

namespace lbmx {

// An error with the reference's category ("Not found", "Invalid state",
// "Malformed JSON", "JSON nesting is too deep") and message.
struct fe_error {
  std::string category, message;
};

template <class V>
struct fe_result {
  bool ok = false;
  V value{};
  fe_error error{};
  static fe_result good(V v) { return fe_result{true, std::move(v), {}}; }
  static fe_result bad(std::string c, std::string m = {}) { return fe_result{false, V{}, {std::move(c), std::move(m)}}; }
};

// ---- converter.hpp ------------------------------------------------------------------
template <class T>
class converter {
 public:
  static constexpr T inv_cs2 = T(3);  // df_info<T, Desc>::inv_cs2 of D2Q9, D3Q19, D3Q27

  converter(T delta_x, T delta_t) : dx_(delta_x), dt_(delta_t) {}

  T delta_x() const { return dx_ * T(1.0); }
  T delta_t() const { return dt_ * T(1.0); }
  T delta_v() const { return (dx_ / dt_) * T(1.0); }
  T delta_a() const { return (dx_ / (dt_ * dt_)) * T(1.0); }
  T meter_si_to_lbm(T m) const { return m / dx_; }
  T meter_lbm_to_si(T m) const { return m * dx_; }
  T second_si_to_lbm(T s) const { return s / dt_; }
  T velocity_si_to_lbm(T v) const { return (v / dx_) * dt_; }
  T acceleration_si_to_lbm(T a) const { return (a / dx_) * (dt_ * dt_); }
  T kinematic_viscosity_si_to_lbm(T nu) const { return (nu / (dx_ * dx_)) * dt_; }
  T kinematic_viscosity_lbm_to_tau(T nu_lbm) const {
    const T scaled = inv_cs2 * nu_lbm;
    return scaled + T(0.5);
  }
  T kinematic_viscosity_si_to_tau(T nu) const { return kinematic_viscosity_lbm_to_tau(kinematic_viscosity_si_to_lbm(nu)); }

 private:
  T dx_, dt_;
};

namespace fe_detail {

inline bool strtod_full(std::string_view view, double& out) {
  std::string buf{view};
  char* end = nullptr;
  out = std::strtod(buf.c_str(), &end);
  return !(end == buf.c_str() || *end != '\0');
}

// std::from_chars semantics: optional '-' (signed only), ASCII digits, in range
template <class I>
inline bool parse_integer(std::string_view v, I& out) {
  bool neg = false;
  size_t k = 0;
  if (std::is_signed<I>::value && !v.empty() && v[0] == '-') {
    neg = true;
    k = 1;
  }
  if (k >= v.size()) return false;
  unsigned __int128 acc = 0;
  const unsigned __int128 lim = neg ? (unsigned __int128)(1) << 63
                                    : (std::is_signed<I>::value ? ((unsigned __int128)(1) << 63) - 1
                                                                : ~uint64_t(0));
  for (; k < v.size(); ++k) {
    if (v[k] < '0' || v[k] > '9') return false;
    acc = acc * 10 + unsigned(v[k] - '0');
    if (acc > lim) return false;
  }
  out = neg ? I(-(__int128)acc) : I(acc);
  return true;
}

}  // namespace fe_detail

// ---- abstract/args.hpp --------------------------------------------------------------
// A runtime mirror of sch::Args<Struct<Member<...>...>, Tuple<...>>: named
// members and positionals of kinds f64, f32, u64, i64, string, bool.
class args {
 public:
  enum class kind { f64, f32, u64, i64, str, boolean };
  using value = std::variant<double, float, uint64_t, int64_t, std::string, bool>;

  args& named(std::string name, kind k, value dflt = {}) {
    order_.push_back(name);
    named_[name] = {k, coerce(k, dflt)};
    return *this;
  }
  args& positional(kind k) {
    pos_kinds_.push_back(k);
    return *this;
  }

  // parse_args (args.hpp:287-340)
  fe_result<int> parse(int argc, char** argv) {
    if (argc <= 0) return fe_result<int>::bad("Invalid state", "No arguments at all, not even argv[0]");
    program_ = argv[0];
    positionals_.assign(pos_kinds_.size(), value{});
    size_t pos = 0;
    for (int i = 1; i < argc; ++i) {
      std::string_view view{argv[i]};
      if (view.size() >= 2 && view.substr(0, 2) == "--") {
        view.remove_prefix(2);
        ++i;
        if (i >= argc) return fe_result<int>::bad("Invalid state", "Missing value for --" + std::string{view});
        auto it = named_.find(std::string{view});
        if (it == named_.end()) return fe_result<int>::bad("Not found", "Unknown argument --" + std::string{view});
        auto r = decode(it->second.first, argv[i]);
        if (!r.ok) return fe_result<int>::bad(r.error.category, r.error.message);
        it->second.second = r.value;
        continue;
      }
      if (pos >= pos_kinds_.size())
        return fe_result<int>::bad("Invalid state",
                                   "Too many positional arguments, did not expect '" + std::string{view} + "'");
      auto r = decode(pos_kinds_[pos], view);
      if (!r.ok) return fe_result<int>::bad(r.error.category, r.error.message);
      positionals_[pos++] = r.value;
    }
    if (pos != pos_kinds_.size())
      return fe_result<int>::bad("Invalid state", "Expected " + std::to_string(pos_kinds_.size()) +
                                                      " positional arguments, got " + std::to_string(pos));
    return fe_result<int>::good(0);
  }

  const std::string& program() const { return program_; }
  template <class V>
  V get(const std::string& name) const {
    return std::get<V>(named_.at(name).second);
  }
  template <class V>
  V positional_value(size_t i) const {
    return std::get<V>(positionals_.at(i));
  }

 private:
  static value coerce(kind k, const value& v) {
    switch (k) {
      case kind::f64: return std::holds_alternative<double>(v) ? v : value{0.0};
      case kind::f32: return std::holds_alternative<float>(v) ? v : value{0.0f};
      case kind::u64: return std::holds_alternative<uint64_t>(v) ? v : value{uint64_t(0)};
      case kind::i64: return std::holds_alternative<int64_t>(v) ? v : value{int64_t(0)};
      case kind::str: return std::holds_alternative<std::string>(v) ? v : value{std::string{}};
      case kind::boolean: return std::holds_alternative<bool>(v) ? v : value{false};
    }
    return v;
  }

  static fe_result<value> decode(kind k, std::string_view view) {
    using R = fe_result<value>;
    if (k == kind::str) return R::good(std::string{view});
    if (k == kind::boolean) {
      if (view == "true" || view == "1" || view == "yes" || view == "on") return R::good(true);
      if (view == "false" || view == "0" || view == "no" || view == "off") return R::good(false);
      return R::bad("Invalid state", "Expected a boolean, got '" + std::string{view} + "'");
    }
    if (view.empty()) return R::bad("Invalid state", "Can't decode empty data.");
    if (k == kind::f64 || k == kind::f32) {
      double d = 0.0;
      if (!fe_detail::strtod_full(view, d))
        return R::bad("Invalid state", "Expected a number, got '" + std::string{view} + "'");
      return k == kind::f32 ? R::good(static_cast<float>(d)) : R::good(d);
    }
    if (k == kind::u64) {
      uint64_t u = 0;
      if (!fe_detail::parse_integer(view, u))
        return R::bad("Invalid state", "Expected an integer, got '" + std::string{view} + "'");
      return R::good(u);
    }
    int64_t s = 0;
    if (!fe_detail::parse_integer(view, s))
      return R::bad("Invalid state", "Expected an integer, got '" + std::string{view} + "'");
    return R::good(s);
  }

  std::vector<std::string> order_;
  std::map<std::string, std::pair<kind, value>> named_;
  std::vector<kind> pos_kinds_;
  std::vector<value> positionals_;
  std::string program_;
};

// ---- config.hpp: the reference's JSON grammar (abstract/json.hpp:96-420) ---------------
namespace fe_detail {

class json {
 public:
  explicit json(std::string_view s, uint64_t max_depth = 64) : s_(s), max_(max_depth) {}
  fe_error err;

  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
  }
  bool at_end() const { return p_ >= s_.size(); }
  bool fail(const char* m, const char* cat = "Malformed JSON") {
    err = {cat, m};
    return false;
  }
  bool peek(char& c) {
    ws();
    if (at_end()) return fail("Unexpected end of input");
    c = s_[p_];
    return true;
  }
  bool expect(char c) {
    char d;
    if (!peek(d)) return false;
    if (d != c) return fail("Unexpected character");
    ++p_;
    return true;
  }
  bool consume_if(char c) {
    ws();
    if (!at_end() && s_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  bool enter() {
    if (++depth_ > max_) return fail("", "JSON nesting is too deep");
    return true;
  }
  void leave() {
    if (depth_ > 0) --depth_;
  }
  bool hex4(uint32_t& v) {
    if (p_ + 4 > s_.size()) return fail("Truncated \\u escape");
    v = 0;
    for (int i = 0; i < 4; ++i) {
      const char h = s_[p_ + i];
      uint32_t d;
      if (h >= '0' && h <= '9') d = uint32_t(h - '0');
      else if (h >= 'a' && h <= 'f') d = uint32_t(h - 'a') + 10;
      else if (h >= 'A' && h <= 'F') d = uint32_t(h - 'A') + 10;
      else return fail("Bad hex digit");
      v = (v << 4) | d;
    }
    p_ += 4;
    return true;
  }
  static void utf8(std::string& o, uint32_t cp) {
    if (cp < 0x80u) {
      o.push_back(char(cp));
    } else if (cp < 0x800u) {
      o.push_back(char(0xC0u | (cp >> 6)));
      o.push_back(char(0x80u | (cp & 0x3Fu)));
    } else if (cp < 0x10000u) {
      o.push_back(char(0xE0u | (cp >> 12)));
      o.push_back(char(0x80u | ((cp >> 6) & 0x3Fu)));
      o.push_back(char(0x80u | (cp & 0x3Fu)));
    } else {
      o.push_back(char(0xF0u | (cp >> 18)));
      o.push_back(char(0x80u | ((cp >> 12) & 0x3Fu)));
      o.push_back(char(0x80u | ((cp >> 6) & 0x3Fu)));
      o.push_back(char(0x80u | (cp & 0x3Fu)));
    }
  }
  bool string(std::string& out) {
    if (!expect('"')) return false;
    out.clear();
    while (true) {
      if (at_end()) return fail("Unterminated string");
      const char c = s_[p_++];
      if (c == '"') return true;
      if (c != '\\') {
        out.push_back(c);
        continue;
      }
      if (at_end()) return fail("Dangling escape");
      const char e = s_[p_++];
      switch (e) {
        case '"': out.push_back('"'); break;
        case '\\': out.push_back('\\'); break;
        case '/': out.push_back('/'); break;
        case 'b': out.push_back('\b'); break;
        case 'f': out.push_back('\f'); break;
        case 'n': out.push_back('\n'); break;
        case 'r': out.push_back('\r'); break;
        case 't': out.push_back('\t'); break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800u && cp <= 0xDBFFu) {
            if (p_ + 1 < s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
              p_ += 2;
              uint32_t lo;
              if (!hex4(lo)) return false;
              cp = (lo >= 0xDC00u && lo <= 0xDFFFu) ? 0x10000u + ((cp - 0xD800u) << 10) + (lo - 0xDC00u) : 0xFFFDu;
            } else {
              cp = 0xFFFDu;
            }
          } else if (cp >= 0xDC00u && cp <= 0xDFFFu) {
            cp = 0xFFFDu;
          }
          utf8(out, cp);
          break;
        }
        default: return fail("Unknown escape");
      }
    }
  }
  bool number_token(std::string_view& tok) {
    ws();
    const size_t start = p_;
    if (!at_end() && (s_[p_] == '-' || s_[p_] == '+')) ++p_;
    while (!at_end() && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    if (!at_end() && s_[p_] == '.') {
      ++p_;
      while (!at_end() && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    }
    if (!at_end() && (s_[p_] == 'e' || s_[p_] == 'E')) {
      ++p_;
      if (!at_end() && (s_[p_] == '-' || s_[p_] == '+')) ++p_;
      while (!at_end() && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
    }
    if (p_ == start) return fail("Expected a number");
    tok = s_.substr(start, p_ - start);
    return true;
  }
  bool boolean() {
    ws();
    if (s_.compare(p_, 4, "true") == 0) {
      p_ += 4;
      return true;
    }
    if (s_.compare(p_, 5, "false") == 0) {
      p_ += 5;
      return true;
    }
    return fail("Expected a boolean");
  }
  bool null() {
    ws();
    if (s_.compare(p_, 4, "null") == 0) {
      p_ += 4;
      return true;
    }
    return false;
  }
  bool skip_value() {
    char c;
    if (!peek(c)) return false;
    if (!enter()) return false;
    bool ok = true;
    if (c == '{' || c == '[') {
      const char close = c == '{' ? '}' : ']';
      ++p_;
      if (!consume_if(close)) {
        while (true) {
          std::string key;
          if (c == '{' && (!string(key) || !expect(':'))) {
            ok = false;
            break;
          }
          if (!skip_value()) {
            ok = false;
            break;
          }
          if (consume_if(close)) break;
          if (!expect(',')) {
            ok = false;
            break;
          }
        }
      }
    } else if (c == '"') {
      std::string s;
      ok = string(s);
    } else if (null()) {
    } else if (c == 't' || c == 'f') {
      ok = boolean();
    } else {
      std::string_view t;
      ok = number_token(t);
    }
    leave();
    return ok;
  }
  bool real(double& v) {
    std::string_view tok;
    if (!number_token(tok)) return false;
    std::string buf{tok};
    char* end = nullptr;
    v = std::strtod(buf.c_str(), &end);
    if (end == buf.c_str()) return fail("Bad floating point");
    return true;
  }

 private:
  std::string_view s_;
  size_t p_ = 0;
  uint64_t depth_ = 0, max_;
};

}  // namespace fe_detail

template <class T>
struct lbm_config {
  T delta_x{}, kinematic_viscosity{}, delta_t{};
};

// load_lbm_config<T, Desc> (config.hpp:22-50): members not given stay 0,
// unknown members are skipped, trailing content is an error.
template <class T>
fe_result<lbm_config<T>> load_lbm_config(const std::string& file_name) {
  using R = fe_result<lbm_config<T>>;
  std::ifstream file{file_name, std::ios::binary};
  if (!file.is_open()) return R::bad("Not found", "Couldn't open file");
  std::string src{std::istreambuf_iterator<char>{file}, std::istreambuf_iterator<char>{}};
  if (file.bad()) return R::bad("Invalid state", "Couldn't read file");
  fe_detail::json p{src};
  lbm_config<T> c{};
  auto failed = [&]() { return R::bad(p.err.category, p.err.message); };
  if (!p.enter() || !p.expect('{')) return failed();
  if (!p.consume_if('}')) {
    while (true) {
      std::string key;
      if (!p.string(key) || !p.expect(':')) return failed();
      double v = 0.0;
      if (key == "delta_x" || key == "kinematic_viscosity" || key == "delta_t") {
        if (!p.real(v)) return failed();
        (key == "delta_x" ? c.delta_x : key == "delta_t" ? c.delta_t : c.kinematic_viscosity) = static_cast<T>(v);
      } else if (!p.skip_value()) {
        return failed();
      }
      if (p.consume_if('}')) break;
      if (!p.expect(',')) return failed();
    }
  }
  p.leave();
  p.ws();
  if (!p.at_end()) return R::bad("Malformed JSON", "Trailing content after the JSON value");
  return R::good(c);
}

// ---- environment.hpp ----------------------------------------------------------------
struct environment {
  std::string name;
  std::filesystem::path lbm_dir, data_dir, name_dir;
};

namespace fe_detail {
inline bool located_data_dir(const std::filesystem::path& lbm_dir, std::filesystem::path& out) {
  namespace fs = std::filesystem;
  const auto loc = lbm_dir / "data_dir_location.txt";
  std::error_code ec;
  if (fs::exists(loc, ec) && fs::is_regular_file(loc, ec)) {
    std::ifstream f{loc};
    if (f.is_open()) {
      std::stringstream b;
      b << f.rdbuf();
      out = fs::path{b.str()};
      return true;
    }
  }
  return false;
}
}  // namespace fe_detail

// init_lbm_env (environment.hpp:18-56) for a driver named `name`
inline fe_result<environment> init_lbm_env(const std::string& name) {
  const char* home = std::getenv("HOME");
  if (!home) return fe_result<environment>::bad("Not found", "Couldn't find home dir");
  environment env;
  env.lbm_dir = std::filesystem::path{home} / ".lbm";
  env.data_dir = env.lbm_dir / "data";
  std::filesystem::path located;
  if (fe_detail::located_data_dir(env.lbm_dir, located)) env.data_dir = located;
  env.name = name;
  env.name_dir = env.data_dir / name;
  return fe_result<environment>::good(env);
}

// output_directory (environment.hpp:63-89)
inline fe_result<std::filesystem::path> output_directory() {
  const char* home = std::getenv("HOME");
  if (!home) return fe_result<std::filesystem::path>::bad("Not found", "Couldn't find home dir");
  const auto lbm_dir = std::filesystem::path{home} / ".lbm";
  std::filesystem::path located;
  if (fe_detail::located_data_dir(lbm_dir, located)) return fe_result<std::filesystem::path>::good(located);
  return fe_result<std::filesystem::path>::good(lbm_dir / "data");
}

}  // namespace lbmx

#endif
