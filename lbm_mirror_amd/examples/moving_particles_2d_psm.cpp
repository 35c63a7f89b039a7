// moving_particles_2d_psm — the reference's particle-laden channel driver
// (examples/moving_poiseulle_particles_2d_psm_gpu/sim.cpp) on this engine:
// a 5120x256 D2Q9 FP32 channel, bounce-back walls, PSM fluid (tau 0.8),
// Zou/He pressure inflow ramping 1 -> 1.01 over 8 steps and outflow 1 (with
// BGK), one disc of radius 0.1*256 and density 1 at (0.25, 0.5) of the
// channel, DEM wall contact {1e-2, 0.3, 0, 0}, 16 sub-steps by default.
// Command line as the reference's (--name, --sub_steps) through the front
// end's argument parser; output in init_lbm_env's name_dir: the initial state
// as VTK, the macro struct (velocity, density, porosity, force) as VTK every
// 32 steps, and one "Particle: x, y; ax, ay;step" line per step on stdout.
//
// Additions of this port: --time_steps (the reference runs 32*2048) and
// --vtk_every (32).  The initial-state file holds info and dfs: the
// reference's field struct also carries dfs_old (equal to dfs at that point)
// and the HLBM sums particle_N / particle_D (zero in a PSM run), which this
// engine does not keep as fields.
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "../../include/lbm_mirror_amd/component.hpp"
#include "../../include/lbm_mirror_amd/frontend.hpp"

using namespace lbm_amd;
using T = sch::Float32;
using Desc = sch::D2Q9;

constexpr uint64_t kDimY = 256;        // sim.cpp:13-14
constexpr uint64_t kDimX = kDimY * 20;

int main(int argc, char** argv) {
  using K = lbmx::args::kind;
  lbmx::args args;
  args.named("name", K::str, std::string{"moving_poiseulle_particles_2d_psm_gpu"})
      .named("sub_steps", K::u64, uint64_t(16))
      .named("time_steps", K::u64, uint64_t(32 * 2048))
      .named("vtk_every", K::u64, uint64_t(32));
  auto parsed = args.parse(argc, argv);
  if (!parsed.ok) {
    std::fprintf(stderr, "%s: %s\n", parsed.error.category.c_str(), parsed.error.message.c_str());
    return 1;
  }
  auto env = lbmx::init_lbm_env(args.get<std::string>("name"));
  if (!env.ok) {
    std::fprintf(stderr, "%s: %s\n", env.error.category.c_str(), env.error.message.c_str());
    return 1;
  }
  const std::filesystem::path out_dir = env.value.name_dir;
  std::error_code ec;
  std::filesystem::create_directories(out_dir, ec);
  if (ec) {
    std::fprintf(stderr, "Could not create output directory\n");
    return 1;
  }
  const uint64_t steps = args.get<uint64_t>("time_steps");
  const uint64_t every = args.get<uint64_t>("vtk_every");
  const std::string dir = out_dir.string();
  try {
    lattice<T, Desc> lat({kDimX, kDimY, 1});
    lbm_ctx* ctx = lat.handle();
    // init (sim.cpp:70-112): walls, interior fluid, the x faces in- and
    // outflow without their corners
    std::vector<uint8_t> info(lat.cells(), 1);
    for (uint64_t y = 0; y < kDimY; ++y)
      for (uint64_t x = 0; x < kDimX; ++x) {
        const bool iy = y > 0 && y + 1 < kDimY;
        if (!iy) continue;
        info[x + kDimX * y] = x == 0 ? 3 : (x + 1 == kDimX ? 4 : 2);
      }
    lat.set_flags(info);
    // step (sim.cpp:222-262)
    component<T, Desc, cmpt::BGK> bgk{0.8};
    component<T, Desc, cmpt::Psm> collision{0.8};
    component<T, Desc, cmpt::BounceBack> bb;
    component<T, Desc, cmpt::ZouHeHorizontal<true>> flow_in{1.0, 0.01 / 8.0, 1.01, 8};
    component<T, Desc, cmpt::ZouHeHorizontal<false>> flow_out{1.0};
    lat.bind(1, bb, none{});
    lat.bind(2, none{}, collision);
    lat.bind(3, flow_in, bgk);
    lat.bind(4, flow_out, bgk);
    // rho 1, u 0, f = eq everywhere; porosity 1, force 0 (sim.cpp:114-140)
    lat.init_equilibrium(1.0, {0.0, 0.0, 0.0});
    check(lbm_write_vtk(ctx, dir.c_str(), "initial_state", 0, "info,dfs"));
    // the disc (sim.cpp:142-165): r = dim_y * 0.1f, density 1, at rest
    const double radius = double(float(kDimY) * 0.1f);
    lbm_particle p{};
    p.position[0] = p.position_old[0] = 0.25 * double(kDimX);
    p.position[1] = p.position_old[1] = 0.5 * double(kDimY);
    p.radius = radius;
    p.density = 1.0;
    check(lbm_spheroid_mass(2, T::id, radius, 1.0, &p.mass, &p.inertia));
    lbm_particle_params prm{};
    prm.k_n = 1.0e-2;  // dem_params{1.0e-2, 0.3, 0.0, 0.0} (sim.cpp:207)
    prm.zeta_n = 0.3;
    prm.walls = 1;     // ParticleContact with the channel extent
    prm.porosity_mode = LBM_POROSITY_SCATTER;
    prm.sub_steps = args.get<uint64_t>("sub_steps");
    prm.cell_capacity = 4;
    prm.eps = 1.5;
    prm.max_radius = radius;
    prm.coupling = LBM_COUPLING_PSM;
    check(lbm_particles_set(ctx, &p, 1, &prm));
    for (uint64_t i = 0; i < steps; ++i) {
      lat.step(i, 1);
      if (every && i % every == 0) check(lbm_write_vtk(ctx, dir.c_str(), "m", i, "velocity,density,porosity,force"));
      lbm_particle q{};
      check(lbm_particles_get(ctx, &q, 1));
      // std::cout's defaults for the float values: %g, 6 significant digits
      std::printf("Particle: %g, %g; %g, %g;%llu\n", q.position[0], q.position[1], q.acceleration[0],
                  q.acceleration[1], (unsigned long long)i);
    }
  } catch (const error& e) {
    std::fprintf(stderr, "lbm error %d: %s\n", int(e.status), e.what());
    return 1;
  }
  return 0;
}
