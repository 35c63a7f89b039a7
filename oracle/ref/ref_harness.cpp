// ref_harness.cpp — calls the reference's own host-side functions, compiled
// in place from /root/reference (see Makefile), so the oracle can be checked
// against the real arithmetic.  TEST INFRASTRUCTURE ONLY: built into
// oracle/_ref/, loaded by tests/ only.
//
// Covered (all on the reference's host `data<>` types, the way its own tests
// in modules/core/tests call them):
//   equilibrium           modules/core/c++/equilibrium.hpp:7-48
//   compute_rho_u         modules/core/c++/macroscopic.hpp:10-35
//   particle_porosity     modules/core/c++/particle/porosity.hpp:56-89
//   particle_aabb         modules/core/c++/particle/aabb.hpp:41-99
//   create_spheroid_particle_group mass/inertia  particle/particle.hpp:19-116
//   narrow_phase / particle_contact_response / accumulate_particle_contacts /
//   accumulate_box_wall_contacts  particle/contact.hpp:124-512
//   reset_particle_accumulators / integrate_particle / integrate_particle_fixed_axis
//                                                      particle/integrate.hpp
//   make_cell_list / build_cell_list / cell_porosity  particle/cell_list.hpp,
//                                                       particle/couple.hpp
//   write_vtk_file / write_csv_file   write_vtk.hpp, write_csv.hpp
//   converter, parse_args, load_lbm_config, init_lbm_env, output_directory
//                                     converter.hpp, abstract/args.hpp, config.hpp,
//                                     environment.hpp
//   codec<Struct<Chunk...>, enc::Forst>::encode   abstract/forst.hpp, forst_chunk.hpp
// The fluid components (BGK, BGKGuo, Stream, Zou/He, Psm, the PSM gather)
// take Chunk field views that only the SYCL backend provides
// (modules/sycl/c++/data.hpp:323-396); AdaptiveCpp is absent, so they are not
// built here.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>

#include "modules/core/c++/equilibrium.hpp"
#include "modules/core/c++/macroscopic.hpp"
#include "modules/core/c++/particle/contact.hpp"
#include "modules/core/c++/particle/couple.hpp"
#include "modules/core/c++/particle/integrate.hpp"
#include "modules/core/c++/particle/particle.hpp"
#include "modules/core/c++/particle/porosity.hpp"
#include "modules/core/c++/abstract/args.hpp"
#include "modules/core/c++/config.hpp"
#include "modules/core/c++/converter.hpp"
#include "modules/core/c++/environment.hpp"
#include "modules/core/c++/forst_chunk.hpp"
#include "modules/core/c++/write_csv.hpp"
#include "modules/core/c++/write_vtk.hpp"

#include "../../include/lbm_mirror_amd/lbm.h"

namespace kl = kel::lbm;
namespace sch = kel::lbm::sch;

namespace {

template <class T, class Desc>
void eq_t(double rho_in, const double* u, double* out) {
  kl::data<sch::Scalar<T>> rho;
  rho.at({}).set(rho_in);
  kl::data<sch::Vector<T, Desc::D>> vel;
  for (uint64_t d = 0; d < Desc::D; ++d) vel.at({{d}}).set(u[d]);
  auto eq = kl::equilibrium<T, Desc>(rho, vel);
  for (uint64_t i = 0; i < Desc::Q; ++i) out[i] = eq.at(i).get();
}

template <class T, class Desc>
void rhou_t(const double* f, double* rho_out, double* u_out) {
  kl::data<sch::FixedArray<T, Desc::Q>> dfs;
  for (uint64_t i = 0; i < Desc::Q; ++i) dfs.at({i}).set(f[i]);
  kl::data<sch::Scalar<T>> rho;
  kl::data<sch::Vector<T, Desc::D>> vel;
  kl::compute_rho_u<T, Desc>(dfs, rho, vel);
  *rho_out = rho.at({}).get();
  for (uint64_t d = 0; d < Desc::D; ++d) u_out[d] = vel.at({{d}}).get();
}

template <class T, uint64_t D>
double porosity_t(const double* rel, double r, double eps_in) {
  kl::data<sch::Vector<T, D>> v;
  for (uint64_t d = 0; d < D; ++d) v.at({{d}}).set(rel[d]);
  kl::data<sch::Scalar<T>> rad, eps;
  rad.at({}).set(r);
  eps.at({}).set(eps_in);
  auto p = kl::particle_porosity<T, D, 1u, kl::por::ParticleSpheroid<T>>::calculate(v, rad, eps);
  return p.at({}).get();
}

template <class T, uint64_t D, uint64_t PC>
kl::data<sch::ParticleGroup<T, D, PC, kl::coll::Spheroid<T>>> make_group(const lbm_particle* p) {
  kl::data<sch::LbmMeter<sch::Scalar<T>>> r;
  kl::data<sch::Scalar<T>> dens;
  r.at({}).set(p[0].radius);
  dens.at({}).set(p[0].density);
  auto g = kl::create_spheroid_particle_group<T, D, PC>(r, dens, {4u});
  auto& parts = g.template get<"particles">();
  for (uint64_t k = 0; k < PC; ++k) {
    auto& q = parts.at({k});
    auto& rb = q.template get<"rigid_body">();
    for (uint64_t d = 0; d < D; ++d) {
      rb.template get<"position">().at({{d}}).set(p[k].position[d]);
      rb.template get<"position_old">().at({{d}}).set(p[k].position_old[d]);
      rb.template get<"acceleration">().at({{d}}).set(p[k].acceleration[d]);
      rb.template get<"force">().at({{d}}).set(p[k].force[d]);
    }
    if constexpr (D == 2) {
      rb.template get<"rotation">().at({}).set(p[k].rotation[0]);
      rb.template get<"rotation_old">().at({}).set(p[k].rotation_old[0]);
      rb.template get<"angular_acceleration">().at({}).set(p[k].angular_acceleration[0]);
      rb.template get<"torque">().at({}).set(p[k].torque[0]);
    } else {
      for (uint64_t d = 0; d < 3; ++d) {
        rb.template get<"rotation">().at({{d}}).set(p[k].rotation[d]);
        rb.template get<"rotation_old">().at({{d}}).set(p[k].rotation_old[d]);
        rb.template get<"angular_acceleration">().at({{d}}).set(p[k].angular_acceleration[d]);
        rb.template get<"torque">().at({{d}}).set(p[k].torque[d]);
      }
    }
    rb.template get<"mass">().at({}).set(p[k].mass);
    rb.template get<"inertia">().at({}).set(p[k].inertia);
    rb.template get<"density">().at({}).set(p[k].density);
    q.template get<"collision">().template get<"radius">().at({}).set(p[k].radius);
  }
  return g;
}

template <class T, uint64_t D, class G>
void read_group(const G& g, lbm_particle* p, uint64_t PC) {
  auto& parts = g.template get<"particles">();
  for (uint64_t k = 0; k < PC; ++k) {
    auto& q = parts.at({k});
    auto& rb = q.template get<"rigid_body">();
    std::memset(&p[k], 0, sizeof(lbm_particle));
    for (uint64_t d = 0; d < D; ++d) {
      p[k].position[d] = rb.template get<"position">().at({{d}}).get();
      p[k].position_old[d] = rb.template get<"position_old">().at({{d}}).get();
      p[k].acceleration[d] = rb.template get<"acceleration">().at({{d}}).get();
      p[k].force[d] = rb.template get<"force">().at({{d}}).get();
    }
    if constexpr (D == 2) {
      p[k].rotation[0] = rb.template get<"rotation">().at({}).get();
      p[k].rotation_old[0] = rb.template get<"rotation_old">().at({}).get();
      p[k].angular_acceleration[0] = rb.template get<"angular_acceleration">().at({}).get();
      p[k].torque[0] = rb.template get<"torque">().at({}).get();
    } else {
      for (uint64_t d = 0; d < 3; ++d) {
        p[k].rotation[d] = rb.template get<"rotation">().at({{d}}).get();
        p[k].rotation_old[d] = rb.template get<"rotation_old">().at({{d}}).get();
        p[k].angular_acceleration[d] = rb.template get<"angular_acceleration">().at({{d}}).get();
        p[k].torque[d] = rb.template get<"torque">().at({{d}}).get();
      }
    }
    p[k].mass = rb.template get<"mass">().at({}).get();
    p[k].inertia = rb.template get<"inertia">().at({}).get();
    p[k].density = rb.template get<"density">().at({}).get();
    p[k].radius = q.template get<"collision">().template get<"radius">().at({}).get();
  }
}

// One DEM phase sequence on a group: optional reset, contacts for every
// particle (plus box walls), then optional integration — the order of the
// reference's own dynamics tests (contact_dynamics.cpp:44-61).
template <class T, uint64_t D, uint64_t PC>
void dem_t(lbm_particle* p, const lbm_particle_params* prm, const uint64_t* extent, int reset,
           int contact, int integrate) {
  auto g = make_group<T, D, PC>(p);
  auto& parts = g.template get<"particles">();
  kl::dem_params<T> dem{prm->k_n, prm->zeta_n, prm->gamma_t, prm->mu};
  kl::data<sch::FixedArray<sch::UInt64, D>> ext;
  for (uint64_t d = 0; d < D; ++d) ext.at({d}).set(extent[d]);
  if (reset)
    for (uint64_t k = 0; k < PC; ++k) kl::reset_particle_accumulators<T, D>(parts.at({k}));
  if (contact) {
    for (uint64_t k = 0; k < PC; ++k) {
      kl::accumulate_particle_contacts<T, D>(parts, k, dem, {prm->sub_steps});
      if (prm->walls) kl::accumulate_box_wall_contacts<T, D>(parts.at({k}), ext, dem, {prm->sub_steps});
    }
  }
  if (integrate && prm->fixed_axis) {
    // what component<..., cmpt::ParticleAxisIntegrate> does (integrate.hpp:298-327)
    kl::data<kl::impl::spin_axis_schema<T, D>> axis;
    if constexpr (D == 2) {
      axis.at({}).set(prm->spin_axis[0]);
    } else {
      for (uint64_t d = 0; d < 3; ++d) axis.at({{d}}).set(prm->spin_axis[d]);
    }
    const auto unit = kl::impl::normalise_spin_axis<T, D>(axis);
    for (uint64_t k = 0; k < PC; ++k) kl::integrate_particle_fixed_axis<T, D>(parts.at({k}), unit, {prm->sub_steps});
  } else if (integrate) {
    kl::data<sch::LbmAcceleration<sch::Vector<T, D>>> acc;
    for (uint64_t d = 0; d < D; ++d) acc.at({{d}}).set(prm->body_acceleration[d]);
    for (uint64_t k = 0; k < PC; ++k) kl::integrate_particle<T, D>(parts.at({k}), acc, {prm->sub_steps});
  }
  read_group<T, D>(g, p, PC);
}

template <class T, uint64_t D>
int dem_n(uint32_t n, lbm_particle* p, const lbm_particle_params* prm, const uint64_t* ext,
          int reset, int contact, int integrate) {
  switch (n) {
    case 1: dem_t<T, D, 1>(p, prm, ext, reset, contact, integrate); return 0;
    case 2: dem_t<T, D, 2>(p, prm, ext, reset, contact, integrate); return 0;
    case 3: dem_t<T, D, 3>(p, prm, ext, reset, contact, integrate); return 0;
    case 4: dem_t<T, D, 4>(p, prm, ext, reset, contact, integrate); return 0;
  }
  return -1;
}

// cell_porosity over every cell of a 2D lattice, capacity 4
template <class T, uint64_t PC>
uint64_t couple_t(const lbm_particle* p, const uint64_t* extent, double max_radius, double eps_in,
                  double* out) {
  auto g = make_group<T, 2, PC>(p);
  auto& parts = g.template get<"particles">();
  kl::data<sch::FixedArray<sch::UInt64, 2>> ext;
  for (uint64_t d = 0; d < 2; ++d) ext.at({d}).set(extent[d]);
  auto cl = kl::make_cell_list<T, 2, 4u>(ext, max_radius);
  kl::build_cell_list<T, 2, 4u>(parts, cl);
  kl::data<sch::Scalar<T>> eps;
  eps.at({}).set(eps_in);
  for (uint64_t y = 0; y < extent[1]; ++y)
    for (uint64_t x = 0; x < extent[0]; ++x) {
      kl::data<sch::FixedArray<sch::UInt64, 2>> cell{{x, y}};
      out[x + extent[0] * y] = kl::cell_porosity<T, 2, 4u>(parts, cl, cell, eps).at({}).get();
    }
  return cl.template get<"overflow">().at({0u}).get();
}


// ---- result writers ---------------------------------------------------------
// The drivers' field structs (e.g. examples/stokes_drag_particle_2d_hlbm_gpu/
// sim.cpp:28-52): macros as ghost-0 chunks, populations as a ghost-1 chunk.
template <typename T, uint64_t D, uint64_t Q, uint64_t... Dims>
struct out_schema {
  using Vel = sch::Chunk<sch::Vector<T, D>, 0u, Dims...>;
  using Rho = sch::Chunk<sch::Scalar<T>, 0u, Dims...>;
  using Info = sch::Chunk<sch::UInt8, 0u, Dims...>;
  using Df = sch::Chunk<sch::FixedArray<T, Q>, 1u, Dims...>;
  using Full = sch::Struct<sch::Member<Vel, "velocity">, sch::Member<Rho, "density">,
                           sch::Member<Vel, "force">, sch::Member<Rho, "porosity">,
                           sch::Member<Info, "info">, sch::Member<Df, "dfs">>;
  using Macro = sch::Struct<sch::Member<Vel, "velocity">, sch::Member<Rho, "density">>;
};

template <uint64_t... Dims>
kl::data<sch::FixedArray<sch::UInt64, sizeof...(Dims)>> cell_index(uint64_t c) {
  constexpr uint64_t ext[] = {Dims...};
  kl::data<sch::FixedArray<sch::UInt64, sizeof...(Dims)>> idx;
  for (uint64_t d = 0; d < sizeof...(Dims); ++d) {
    idx.at({d}).set(c % ext[d]);
    c /= ext[d];
  }
  return idx;
}

struct out_in {
  const double *vel, *rho, *force, *por, *dfs;
  const uint8_t* info;
};

template <typename T, uint64_t D, uint64_t Q, bool Full, uint64_t... Dims, typename S>
void fill_fields(kl::data<S>& f, const out_in& in) {
  constexpr uint64_t N = (Dims * ...);
  for (uint64_t c = 0; c < N; ++c) {
    auto idx = cell_index<Dims...>(c);
    for (uint64_t d = 0; d < D; ++d) f.template get<"velocity">().at(idx).at({{d}}).set(in.vel[d * N + c]);
    f.template get<"density">().at(idx).at({}).set(in.rho[c]);
    if constexpr (Full) {
      for (uint64_t d = 0; d < D; ++d) f.template get<"force">().at(idx).at({{d}}).set(in.force[d * N + c]);
      f.template get<"porosity">().at(idx).at({}).set(in.por[c]);
      f.template get<"info">().at(idx).set(in.info[c]);
      for (uint64_t i = 0; i < Q; ++i) f.template get<"dfs">().at(idx).at({i}).set(in.dfs[i * N + c]);
    }
  }
}

// what: 0 VTK, 1 CSV (files under out_dir as the reference names them),
// 2 Forst encoding into buf; returns the byte count for 2, 0 otherwise, -1 on error.
template <typename T, uint64_t D, uint64_t Q, uint64_t... Dims>
int64_t write_t(int full, const out_in& in, const char* out_dir, const char* name, uint64_t t, int what,
                uint8_t* buf, uint64_t cap) {
  auto run = [&](auto* tag, auto full_c) -> int64_t {
    using S = std::remove_pointer_t<decltype(tag)>;
    auto f = std::make_unique<kl::data<S>>();
    fill_fields<T, D, Q, decltype(full_c)::value, Dims...>(*f, in);
    if (what == 0) return kl::write_vtk_file(std::filesystem::path{out_dir}, name, t, *f).is_value() ? 0 : -1;
    if (what == 1) return kl::write_csv_file(std::filesystem::path{out_dir}, name, t, *f).is_value() ? 0 : -1;
    kl::codec<S, kl::enc::Forst> codec;
    kl::data<S, kl::enc::Forst> bytes;
    if (!codec.encode(*f, bytes).is_value()) return -1;
    const uint64_t n = bytes.size().get();
    if (n > cap) return -1;
    std::memcpy(buf, bytes.raw(), n);
    return int64_t(n);
  };
  using OS = out_schema<T, D, Q, Dims...>;
  return full ? run((typename OS::Full*)nullptr, std::true_type{})
              : run((typename OS::Macro*)nullptr, std::false_type{});
}

template <typename T>
int64_t write_dims(int shape, int full, const out_in& in, const char* dir, const char* name, uint64_t t, int what,
                   uint8_t* buf, uint64_t cap) {
  switch (shape) {
    case 0: return write_t<T, 2, 9, 5u, 3u>(full, in, dir, name, t, what, buf, cap);
    case 1: return write_t<T, 3, 19, 3u, 4u, 2u>(full, in, dir, name, t, what, buf, cap);
    case 2: return write_t<T, 3, 19, 128u, 64u, 32u>(full, in, dir, name, t, what, buf, cap);
  }
  return -1;
}


// ---- unit / config front end ---------------------------------------------------
template <typename T>
double convert_t(double dx, double dt, int what, double v) {
  kl::converter<T> conv{{dx}, {dt}};
  switch (what) {
    case 0: return conv.meter_si_to_lbm(kl::data<sch::SiMeter<T>>{v}).handle().get();
    case 1: return conv.meter_lbm_to_si(kl::data<sch::LbmMeter<T>>{v}).handle().get();
    case 2: return conv.second_si_to_lbm(kl::data<sch::SiSecond<T>>{v}).handle().get();
    case 3: return conv.velocity_si_to_lbm(kl::data<sch::SiVelocity<T>>{v}).handle().get();
    case 4: return conv.acceleration_si_to_lbm(kl::data<sch::SiAcceleration<T>>{v}).handle().get();
    case 5: return conv.kinematic_viscosity_si_to_lbm(kl::data<sch::SiKinematicViscosity<T>>{v}).handle().get();
    case 6: return conv.template kinematic_viscosity_si_to_tau<sch::D2Q9>(kl::data<sch::SiKinematicViscosity<T>>{v}).handle().get();
    case 7: return conv.template kinematic_viscosity_lbm_to_tau<sch::D2Q9>(kl::data<sch::LbmKinematicViscosity<T>>{v}).handle().get();
    case 8: return conv.delta_x().handle().get();
    case 9: return conv.delta_t().handle().get();
    case 10: return conv.delta_v().handle().get();
    case 11: return conv.delta_a().handle().get();
    case 12: return conv.template kinematic_viscosity_si_to_tau<sch::D3Q27>(kl::data<sch::SiKinematicViscosity<T>>{v}).handle().get();
  }
  return 0.0;
}

// The drivers' argument struct (vawt_3d_hlbm_gpu/sim.cpp:44-52 and
// dkt's coupling string) plus an integer positional pair, as the
// reference's own args tests use (tests/args.cpp:33-58).
using RefArgs = sch::Args<
    sch::Struct<sch::Member<sch::String, "name">, sch::Member<sch::UInt64, "time_steps">,
                sch::Member<sch::UInt64, "sub_steps">, sch::Member<sch::Float64, "axis_x">,
                sch::Member<sch::Float32, "visc">, sch::Member<sch::Int64, "shift">,
                sch::Member<sch::String, "coupling">, sch::Member<sch::Bool, "verbose">>,
    sch::Tuple<sch::UInt64, sch::Float64>>;

void put_text(char* out, uint64_t cap, std::string_view v) {
  if (!out || cap == 0) return;
  const uint64_t n = std::min<uint64_t>(cap - 1, v.size());
  std::memcpy(out, v.data(), n);
  out[n] = '\0';
}
}  // namespace

extern "C" {

int ref_equilibrium(int lat, int st, double rho, const double* u, double* eq) {
  if (lat == LBM_D2Q9) st == LBM_F64 ? eq_t<sch::Float64, sch::D2Q9>(rho, u, eq) : eq_t<sch::Float32, sch::D2Q9>(rho, u, eq);
  else if (lat == LBM_D3Q19) st == LBM_F64 ? eq_t<sch::Float64, sch::D3Q19>(rho, u, eq) : eq_t<sch::Float32, sch::D3Q19>(rho, u, eq);
  else if (lat == LBM_D3Q27) st == LBM_F64 ? eq_t<sch::Float64, sch::D3Q27>(rho, u, eq) : eq_t<sch::Float32, sch::D3Q27>(rho, u, eq);
  else return -1;
  return 0;
}

int ref_rho_u(int lat, int st, const double* f, double* rho, double* u) {
  if (lat == LBM_D2Q9) st == LBM_F64 ? rhou_t<sch::Float64, sch::D2Q9>(f, rho, u) : rhou_t<sch::Float32, sch::D2Q9>(f, rho, u);
  else if (lat == LBM_D3Q19) st == LBM_F64 ? rhou_t<sch::Float64, sch::D3Q19>(f, rho, u) : rhou_t<sch::Float32, sch::D3Q19>(f, rho, u);
  else if (lat == LBM_D3Q27) st == LBM_F64 ? rhou_t<sch::Float64, sch::D3Q27>(f, rho, u) : rhou_t<sch::Float32, sch::D3Q27>(f, rho, u);
  else return -1;
  return 0;
}

double ref_porosity(int D, int st, const double* rel, double r, double eps) {
  if (D == 2) return st == LBM_F64 ? porosity_t<sch::Float64, 2>(rel, r, eps) : porosity_t<sch::Float32, 2>(rel, r, eps);
  return st == LBM_F64 ? porosity_t<sch::Float64, 3>(rel, r, eps) : porosity_t<sch::Float32, 3>(rel, r, eps);
}

int ref_aabb(int st, const lbm_particle* p, const uint64_t* extent, uint64_t* a, uint64_t* b) {
  auto run = [&](auto zero) {
    using T = decltype(zero);
    using PG = sch::ParticleGroup<T, 2, 1, kl::coll::Spheroid<T>>;
    auto g = make_group<T, 2, 1>(p);
    kl::data<sch::FixedArray<sch::UInt64, 2>> meta{{extent[0], extent[1]}};
    auto eo = kl::particle_aabb<PG>::calculate(g, {{0u}}, meta);
    if (!eo.is_value()) return -1;
    auto& v = eo.get_value();
    for (uint64_t d = 0; d < 2; ++d) {
      a[d] = v.template get<"a">().at({d}).get();
      b[d] = v.template get<"b">().at({d}).get();
    }
    return 0;
  };
  return st == LBM_F64 ? run(sch::Float64{}) : run(sch::Float32{});
}

int ref_spheroid_mass(int D, int st, double r, double rho, double* m, double* I) {
  lbm_particle p{};
  p.radius = r;
  p.density = rho;
  auto run = [&](auto zero, auto dims) {
    using T = decltype(zero);
    constexpr uint64_t DD = decltype(dims)::value;
    kl::data<sch::LbmMeter<sch::Scalar<T>>> rr;
    kl::data<sch::Scalar<T>> dens;
    rr.at({}).set(r);
    dens.at({}).set(rho);
    auto g = kl::create_spheroid_particle_group<T, DD, 1>(rr, dens, {4u});
    auto& rb = g.template get<"particles">().at({0u}).template get<"rigid_body">();
    *m = rb.template get<"mass">().at({}).get();
    *I = rb.template get<"inertia">().at({}).get();
  };
  using D2 = std::integral_constant<uint64_t, 2>;
  using D3 = std::integral_constant<uint64_t, 3>;
  if (D == 2) st == LBM_F64 ? run(sch::Float64{}, D2{}) : run(sch::Float32{}, D2{});
  else st == LBM_F64 ? run(sch::Float64{}, D3{}) : run(sch::Float32{}, D3{});
  return 0;
}

int ref_dem(int D, int st, uint32_t n, lbm_particle* p, const lbm_particle_params* prm,
            const uint64_t* extent, int reset, int contact, int integrate) {
  if (D == 2) return st == LBM_F64 ? dem_n<sch::Float64, 2>(n, p, prm, extent, reset, contact, integrate)
                                   : dem_n<sch::Float32, 2>(n, p, prm, extent, reset, contact, integrate);
  return st == LBM_F64 ? dem_n<sch::Float64, 3>(n, p, prm, extent, reset, contact, integrate)
                       : dem_n<sch::Float32, 3>(n, p, prm, extent, reset, contact, integrate);
}

int64_t ref_couple(int st, uint32_t n, const lbm_particle* p, const uint64_t* extent,
                   double max_radius, double eps, double* out) {
  auto run = [&](auto zero) -> int64_t {
    using T = decltype(zero);
    switch (n) {
      case 1: return int64_t(couple_t<T, 1>(p, extent, max_radius, eps, out));
      case 2: return int64_t(couple_t<T, 2>(p, extent, max_radius, eps, out));
      case 3: return int64_t(couple_t<T, 3>(p, extent, max_radius, eps, out));
      case 4: return int64_t(couple_t<T, 4>(p, extent, max_radius, eps, out));
    }
    return -1;
  };
  return st == LBM_F64 ? run(sch::Float64{}) : run(sch::Float32{});
}

// Result writers on the reference's own field structs.  shape 0: 2D 5x3
// (velocity D=2, dfs Q=9); 1: 3D 3x4x2 (Q=19); 2: 3D 128x64x32 (timing).
// full 1: {velocity, density, force, porosity, info, dfs}; 0: {velocity,
// density}.  Inputs are SoA doubles (cell x-fastest), cast to the storage type.
int64_t ref_write_fields(int shape, int st, int full, const double* vel, const double* rho,
                         const double* force, const double* por, const uint8_t* info, const double* dfs,
                         const char* out_dir, const char* name, uint64_t t, int what, uint8_t* buf,
                         uint64_t cap, double* seconds) {
  out_in in{vel, rho, force, por, dfs, info};
  auto t0 = std::chrono::steady_clock::now();
  int64_t r = st == LBM_F64 ? write_dims<sch::Float64>(shape, full, in, out_dir, name, t, what, buf, cap)
                            : write_dims<sch::Float32>(shape, full, in, out_dir, name, t, what, buf, cap);
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return r;
}

double ref_convert(int st, double dx, double dt, int what, double v) {
  return st == LBM_F64 ? convert_t<sch::Float64>(dx, dt, what, v) : convert_t<sch::Float32>(dx, dt, what, v);
}

// parse_args on RefArgs.  Returns 0 on success, else 1 + the error's
// description and message in err.  nums: time_steps, sub_steps, axis_x,
// visc, shift, verbose, pos0, pos1; text: "name\ncoupling\nprogram".
int ref_parse_args(int argc, char** argv, double* nums, char* text, uint64_t text_cap, char* err,
                   uint64_t err_cap) {
  kl::data<RefArgs> a;
  auto& n = a.template get<"args">();
  n.template get<"name">().set(std::string{"default_name"});
  n.template get<"time_steps">().set(256u);
  n.template get<"sub_steps">().set(16u);
  auto eov = kl::parse_args(a, argc, argv);
  if (eov.is_error()) {
    auto& e = eov.get_error();
    std::string m = std::string{e.get_category()} + "|" + std::string{e.get_message()};
    put_text(err, err_cap, m);
    return 1;
  }
  nums[0] = double(n.template get<"time_steps">().get());
  nums[1] = double(n.template get<"sub_steps">().get());
  nums[2] = n.template get<"axis_x">().get();
  nums[3] = double(n.template get<"visc">().get());
  nums[4] = double(n.template get<"shift">().get());
  nums[5] = n.template get<"verbose">().get() ? 1.0 : 0.0;
  auto& pos = a.template get<"positionals">();
  nums[6] = double(pos.template get<0u>().get());
  nums[7] = pos.template get<1u>().get();
  std::string t = std::string{n.template get<"name">().stl_view()} + "\n" +
                  std::string{n.template get<"coupling">().stl_view()} + "\n" +
                  std::string{a.template get<"program">().stl_view()};
  put_text(text, text_cap, t);
  return 0;
}

// load_lbm_config (config.hpp:22-50): 0 and the three values, else 1 and the
// error's category|message.
int ref_load_config(const char* path, int st, double* out, char* err, uint64_t err_cap) {
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    auto eo = kl::load_lbm_config<T, sch::D2Q9>(path);
    if (eo.is_error()) {
      auto& e = eo.get_error();
      put_text(err, err_cap, std::string{e.get_category()} + "|" + std::string{e.get_message()});
      return 1;
    }
    auto& c = eo.get_value();
    out[0] = double(c.template get<"delta_x">().get());
    out[1] = double(c.template get<"kinematic_viscosity">().get());
    out[2] = double(c.template get<"delta_t">().get());
    return 0;
  };
  return st == LBM_F64 ? run(sch::Float64{}) : run(sch::Float32{});
}

// init_lbm_env (environment.hpp:18-56) for an argument struct named `name`,
// and output_directory (:63-89): lbm_dir, data_dir, name_dir, output_dir
// separated by the unit separator 0x1f.
int ref_lbm_env(const char* name, char* out, uint64_t cap, char* err, uint64_t err_cap) {
  kl::data<RefArgs> a;
  a.template get<"args">().template get<"name">().set(std::string{name});
  auto eo = kl::init_lbm_env(a);
  if (eo.is_error()) {
    put_text(err, err_cap, std::string{eo.get_error().get_category()} + "|" +
                               std::string{eo.get_error().get_message()});
    return 1;
  }
  auto od = kl::output_directory();
  if (od.is_error()) return 2;
  auto& e = eo.get_value();
  put_text(out, cap, e.lbm_dir.string() + "\x1f" + e.data_dir.string() + "\x1f" + e.name_dir.string() + "\x1f" +
                         od.get_value().string());
  return 0;
}

}  // extern "C"
