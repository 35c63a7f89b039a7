// vawt_3d — the reference's vertical-axis rotor driver
// (examples/vawt_3d_hlbm_gpu/sim.cpp) on this engine: its command line
// (--name --time_steps --sub_steps --axis_x --axis_y --axis_z) through the
// front end's argument parser, tau from the kinematic viscosity through the
// converter, the output directory from init_lbm_env, the component mirror for
// the flag switch, and the C ABI for the pinned HLBM sphere that spins about
// the given axis (ParticleAxisIntegrate).  Like the reference it writes
// <name_dir>/particle.csv (one row per step) and the macro struct
// (velocity, density, force, porosity) as VTK.
//
// Additions of this port: --vtk_every N (0: only the final field; the
// reference writes every step, N = 1) and --stable_faces true (the
// normal-only D3Q27 Zou/He faces; the reference's faces return 4/3 of the
// tangential momentum each step and these channels diverge after ~100 steps).
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/lbm_mirror_amd/component.hpp"
#include "../../include/lbm_mirror_amd/frontend.hpp"

using namespace lbm_amd;
using T = sch::Float32;
using Desc = sch::D3Q27;

constexpr uint64_t kDim = 32;  // sim.cpp:17-19

int main(int argc, char** argv) {
  using K = lbmx::args::kind;
  lbmx::args args;
  args.named("name", K::str, std::string{"vawt_3d_hlbm_gpu"})
      .named("time_steps", K::u64, uint64_t(256))
      .named("sub_steps", K::u64, uint64_t(16))
      .named("axis_x", K::f64, 0.0)
      .named("axis_y", K::f64, 1.0)
      .named("axis_z", K::f64, 0.0)
      .named("vtk_every", K::u64, uint64_t(1))
      .named("stable_faces", K::boolean, false);
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
    std::fprintf(stderr, "could not create %s\n", out_dir.string().c_str());
    return 1;
  }
  const uint64_t steps = args.get<uint64_t>("time_steps");
  const uint64_t sub_steps = args.get<uint64_t>("sub_steps");
  const uint64_t vtk_every = args.get<uint64_t>("vtk_every");
  // converter{delta_x 1, delta_t 1}; tau from nu = 0.1 (sim.cpp:295)
  const lbmx::converter<float> conv{1.0f, 1.0f};
  const double tau = conv.kinematic_viscosity_si_to_tau(0.1f);
  try {
    lattice<T, Desc> lat({kDim, kDim, kDim});
    lbm_ctx* ctx = lat.handle();
    const uint64_t N = lat.cells();
    // setup_initial_conditions (sim.cpp:100-180): walls, interior fluid, the x
    // faces in- and outflow without their edges; rho 1, u = (0.01, 0, 0) inside
    std::vector<uint8_t> info(N, 1);
    std::vector<double> rho(N, 1.0), u(3 * N, 0.0);
    for (uint64_t z = 0; z < kDim; ++z)
      for (uint64_t y = 0; y < kDim; ++y)
        for (uint64_t x = 0; x < kDim; ++x) {
          const uint64_t c = x + kDim * (y + kDim * z);
          const bool iy = y > 0 && y + 1 < kDim, iz = z > 0 && z + 1 < kDim;
          if (iy && iz) info[c] = x == 0 ? 3 : (x + 1 == kDim ? 4 : 2);
          if (iy && iz && x > 0 && x + 1 < kDim) u[c] = 0.01;
        }
    lat.set_flags(info);
    const bool stable = args.get<bool>("stable_faces");
    component<T, Desc, cmpt::BounceBack> bb;
    component<T, Desc, cmpt::Hlbm> collision{tau};
    component<T, Desc, cmpt::ZouHeHorizontal<true>> flow_in{1.0, 0.001 / 256.0, 1.001, 256};
    component<T, Desc, cmpt::ZouHeHorizontal<false>> flow_out{1.0};
    flow_in.normal_only(stable);
    flow_out.normal_only(stable);
    lat.bind(1, bb, none{});
    lat.bind(2, none{}, collision);
    lat.bind(3, flow_in, collision);
    lat.bind(4, flow_out, collision);
    check(lbm_init_equilibrium_field(ctx, rho.data(), u.data()));
    // the sphere: r = 0.15 dim, density 1, at (0.3, 0.5, 0.5) dim, at rest,
    // pinned, spinning about the axis; no contact pass (sim.cpp:190-230, 266-285)
    const double radius = 0.15 * double(kDim);
    lbm_particle p{};
    const double at[3] = {0.3 * kDim, 0.5 * kDim, 0.5 * kDim};
    for (int d = 0; d < 3; ++d) p.position[d] = p.position_old[d] = at[d];
    p.radius = radius;
    p.density = 1.0;
    check(lbm_spheroid_mass(3, T::id, radius, 1.0, &p.mass, &p.inertia));
    lbm_particle_params prm{};
    prm.sub_steps = sub_steps;
    prm.porosity_mode = LBM_POROSITY_SCATTER;
    prm.cell_capacity = 4;
    prm.eps = 1.5;
    prm.max_radius = radius;
    prm.coupling = LBM_COUPLING_HLBM;
    prm.fixed_axis = 1;
    prm.spin_axis[0] = args.get<double>("axis_x");
    prm.spin_axis[1] = args.get<double>("axis_y");
    prm.spin_axis[2] = args.get<double>("axis_z");
    check(lbm_particles_set(ctx, &p, 1, &prm));
    // the axis the rows are projected on, normalised in T (sim.cpp:485-494)
    float ax[3] = {float(prm.spin_axis[0]), float(prm.spin_axis[1]), float(prm.spin_axis[2])};
    float dot = 0.0f;
    for (float a : ax) dot = dot + a * a;
    double axis_hat[3];
    for (int k = 0; k < 3; ++k) axis_hat[k] = dot > 0.0f ? double(ax[k] / std::sqrt(dot)) : double(ax[k]);

    std::ofstream csv{out_dir / "particle.csv"};
    if (!csv.is_open()) {
      std::fprintf(stderr, "could not open particle.csv\n");
      return 1;
    }
    csv << "step,theta,omega,alpha,tau,drag,"
        << "x,y,z,vx,vy,vz,fx,fy,fz,"
        << "rot_x,rot_y,rot_z,omega_x,omega_y,omega_z,"
        << "alpha_x,alpha_y,alpha_z,tau_x,tau_y,tau_z\n";
    const double ss = double(sub_steps == 0 ? 1 : sub_steps);
    const std::string dir = out_dir.string();
    const char* members = "velocity,density,force,porosity";
    for (uint64_t t = 0; t < steps; ++t) {
      if (vtk_every && t % vtk_every == 0) check(lbm_write_vtk(ctx, dir.c_str(), "t", t, members));
      lat.step(t, 1);
      lbm_particle q{};
      check(lbm_particles_get(ctx, &q, 1));
      double th = 0, om = 0, al = 0, tq = 0;
      for (int k = 0; k < 3; ++k) {
        th += q.rotation[k] * axis_hat[k];
        om += (q.rotation[k] - q.rotation_old[k]) * ss * axis_hat[k];
        al += q.angular_acceleration[k] * axis_hat[k];
        tq += q.torque[k] * axis_hat[k];
      }
      csv << t << "," << th << "," << om << "," << al << "," << tq << "," << q.force[0];
      for (int k = 0; k < 3; ++k) csv << "," << q.position[k];
      for (int k = 0; k < 3; ++k) csv << "," << (q.position[k] - q.position_old[k]) * ss;
      for (int k = 0; k < 3; ++k) csv << "," << q.force[k];
      for (int k = 0; k < 3; ++k) csv << "," << q.rotation[k];
      for (int k = 0; k < 3; ++k) csv << "," << (q.rotation[k] - q.rotation_old[k]) * ss;
      for (int k = 0; k < 3; ++k) csv << "," << q.angular_acceleration[k];
      for (int k = 0; k < 3; ++k) csv << "," << q.torque[k];
      csv << "\n";
    }
    check(lbm_write_vtk(ctx, dir.c_str(), "t", steps, members));
    lbm_particle q{};
    check(lbm_particles_get(ctx, &q, 1));
    std::printf("{\"example\": \"vawt_3d\", \"steps\": %llu, \"tau\": %.9g, \"out\": \"%s\", "
                "\"rotation\": [%.17g, %.17g, %.17g]}\n",
                (unsigned long long)steps, tau, dir.c_str(), q.rotation[0], q.rotation[1], q.rotation[2]);
  } catch (const error& e) {
    std::fprintf(stderr, "lbm error %d: %s\n", int(e.status), e.what());
    return 1;
  }
  return 0;
}
