// poiseuille_2d — the reference's 2D channel driver (examples/stdkr_2d_bgk:
// init.hpp flag layout, step.hpp:21-67 component switch) written against the
// C++ component mirror.  D2Q9 FP64, parabolic Zou/He velocity inlet, Zou/He
// pressure outlet (rho = 1), bounce-back walls.
//
// usage: poiseuille_2d [nx ny steps u_max tau out.bin]
// Writes density then velocity (doubles, x-fastest) to out.bin and prints a
// JSON line with the run's MLUPS.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/lbm_mirror_amd/component.hpp"

using namespace lbm_amd;
using T = sch::Float64;
using Desc = sch::D2Q9;

int main(int argc, char** argv) {
  const uint64_t nx = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 256;
  const uint64_t ny = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 64;
  const uint64_t steps = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 20000;
  const double u_max = argc > 4 ? std::atof(argv[4]) : 0.05;
  const double tau = argc > 5 ? std::atof(argv[5]) : 0.8;
  const char* out = argc > 6 ? argv[6] : nullptr;
  try {
    lattice<T, Desc> lat({nx, ny, 1});
    // setup_initial_conditions: walls everywhere, fluid inside, inflow and
    // outflow columns without the corner rows
    std::vector<uint8_t> info(nx * ny, 1);
    for (uint64_t y = 1; y + 1 < ny; ++y)
      for (uint64_t x = 0; x < nx; ++x) info[x + nx * y] = x == 0 ? 3 : (x == nx - 1 ? 4 : 2);
    lat.set_flags(info);
    component<T, Desc, cmpt::BounceBack> bb;
    component<T, Desc, cmpt::BGK> collision{tau};
    component<T, Desc, cmpt::ZouHeVelocityX<true>> flow_in;
    component<T, Desc, cmpt::ZouHePressureX<false>> flow_out{1.0};
    lat.bind(1, bb, none{});
    lat.bind(2, none{}, collision);
    lat.bind(3, flow_in, collision);
    lat.bind(4, flow_out, collision);
    // per-cell inlet velocity: parabola with the walls half-way
    std::vector<double> u(2 * nx * ny, 0.0);
    for (uint64_t y = 1; y + 1 < ny; ++y) {
      const double yy = double(y);
      u[nx * y] = u_max * 4.0 * (yy - 0.5) * (double(ny) - 1.5 - yy) / ((double(ny) - 2.0) * (double(ny) - 2.0));
    }
    lat.set_boundary_velocity(u);
    lat.init_equilibrium(1.0, {0.0, 0.0, 0.0});
    lat.step(0, steps);
    lat.sync();
    const double ms = lat.last_step_ms();
    std::vector<double> rho, vel;
    lat.macros(rho, vel);
    if (out) {
      FILE* fp = std::fopen(out, "wb");
      if (!fp) return 2;
      std::fwrite(rho.data(), sizeof(double), rho.size(), fp);
      std::fwrite(vel.data(), sizeof(double), vel.size(), fp);
      std::fclose(fp);
    }
    const double mlups = double(nx * ny) * double(steps) / (ms * 1e-3) / 1e6;
    std::printf("{\"example\": \"poiseuille_2d\", \"nx\": %llu, \"ny\": %llu, \"steps\": %llu, "
                "\"ms\": %.3f, \"MLUPS\": %.1f, \"ux_centre\": %.17g}\n",
                (unsigned long long)nx, (unsigned long long)ny, (unsigned long long)steps, ms, mlups,
                vel[(3 * nx / 4) + nx * (ny / 2)]);
  } catch (const error& e) {
    std::fprintf(stderr, "lbm error %d: %s\n", int(e.status), e.what());
    return 1;
  }
  return 0;
}
