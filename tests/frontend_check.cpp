// Drives include/lbm_mirror_amd/frontend.hpp for tests/test_frontend.py: one
// case per stdin line (tab-separated; argv words separated by 0x1f), one
// result line per case on stdout, numbers as C99 hex floats.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "lbm_mirror_amd/frontend.hpp"

static std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t p = 0;
  while (true) {
    size_t q = s.find(sep, p);
    out.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return out;
}

static std::string hx(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%a", v);
  return b;
}

template <class T>
static double convert(double dx, double dt, int what, double v) {
  lbmx::converter<T> c{T(dx), T(dt)};
  const T x = T(v);
  switch (what) {
    case 0: return c.meter_si_to_lbm(x);
    case 1: return c.meter_lbm_to_si(x);
    case 2: return c.second_si_to_lbm(x);
    case 3: return c.velocity_si_to_lbm(x);
    case 4: return c.acceleration_si_to_lbm(x);
    case 5: return c.kinematic_viscosity_si_to_lbm(x);
    case 6: return c.kinematic_viscosity_si_to_tau(x);
    case 7: return c.kinematic_viscosity_lbm_to_tau(x);
    case 8: return c.delta_x();
    case 9: return c.delta_t();
    case 10: return c.delta_v();
    case 11: return c.delta_a();
    case 12: return c.kinematic_viscosity_si_to_tau(x);
  }
  return 0.0;
}

int main() {
  using K = lbmx::args::kind;
  std::string line;
  while (std::getline(std::cin, line)) {
    auto f = split(line, '\t');
    const std::string& mode = f[0];
    if (mode == "conv") {
      const int st = std::stoi(f[1]);
      const double dx = std::strtod(f[2].c_str(), nullptr), dt = std::strtod(f[3].c_str(), nullptr);
      const int what = std::stoi(f[4]);
      const double v = std::strtod(f[5].c_str(), nullptr);
      std::cout << hx(st == 0 ? convert<double>(dx, dt, what, v) : convert<float>(dx, dt, what, v)) << "\n";
    } else if (mode == "args") {
      auto words = split(f[1], '\x1f');
      std::vector<char*> argv;
      for (auto& w : words) argv.push_back(w.data());
      lbmx::args a;
      a.named("name", K::str, std::string{"default_name"})
          .named("time_steps", K::u64, uint64_t(256))
          .named("sub_steps", K::u64, uint64_t(16))
          .named("axis_x", K::f64)
          .named("visc", K::f32)
          .named("shift", K::i64)
          .named("coupling", K::str)
          .named("verbose", K::boolean)
          .positional(K::u64)
          .positional(K::f64);
      auto r = a.parse(int(argv.size()), argv.data());
      if (!r.ok) {
        std::cout << "err\x1f" << r.error.category << "\x1f" << r.error.message << "\n";
        continue;
      }
      std::cout << "ok\x1f" << hx(double(a.get<uint64_t>("time_steps"))) << "\x1f"
                << hx(double(a.get<uint64_t>("sub_steps"))) << "\x1f" << hx(a.get<double>("axis_x")) << "\x1f"
                << hx(double(a.get<float>("visc"))) << "\x1f" << hx(double(a.get<int64_t>("shift"))) << "\x1f"
                << hx(a.get<bool>("verbose") ? 1.0 : 0.0) << "\x1f" << hx(double(a.positional_value<uint64_t>(0)))
                << "\x1f" << hx(a.positional_value<double>(1)) << "\x1f" << a.get<std::string>("name") << "\x1f"
                << a.get<std::string>("coupling") << "\x1f" << a.program() << "\n";
    } else if (mode == "cfg") {
      const int st = std::stoi(f[1]);
      auto out = [&](auto r) {
        if (!r.ok) {
          std::cout << "err\x1f" << r.error.category << "\x1f" << r.error.message << "\n";
        } else {
          std::cout << "ok\x1f" << hx(double(r.value.delta_x)) << "\x1f" << hx(double(r.value.kinematic_viscosity))
                    << "\x1f" << hx(double(r.value.delta_t)) << "\n";
        }
      };
      if (st == 0) out(lbmx::load_lbm_config<double>(f[2]));
      else out(lbmx::load_lbm_config<float>(f[2]));
    } else if (mode == "env") {
      auto e = lbmx::init_lbm_env(f[1]);
      auto o = lbmx::output_directory();
      if (!e.ok || !o.ok) {
        std::cout << "err\x1f" << e.error.category << "\x1f" << e.error.message << "\n";
        continue;
      }
      std::cout << "ok\x1f" << e.value.lbm_dir.string() << "\x1f" << e.value.data_dir.string() << "\x1f"
                << e.value.name_dir.string() << "\x1f" << o.value.string() << "\n";
    }
  }
  return 0;
}
