// component.hpp — C++ host mirror of the reference's component vocabulary
// (keldu/lbm_mirror modules/core/c++/component.hpp:35-36, collision.hpp,
// boundary.hpp, psm.hpp) over the C ABI in lbm.h.
//
// The reference builds `component<T, Desc, cmpt::Kind, encode::Sycl>` on the
// host and calls `apply(field, macros, index, t)` per cell inside a
// `parallel_for` with a `switch(info)` (poiseulle_2d_gpu.cpp:158-235).  Here
// the same component objects are built the same way, bound to a flag once
// (`lattice.bind(flag, boundary, collision)`), and `lattice.step(t0, n)` runs
// whole steps on the GPU.  Header-only C++17; no HIP or torch types appear.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm.h"

namespace lbm_amd {

namespace sch {  // lattice and precision tags with the reference's names
struct D2Q9 { static constexpr int32_t id = LBM_D2Q9; static constexpr int D = 2, Q = 9; };
struct D3Q19 { static constexpr int32_t id = LBM_D3Q19; static constexpr int D = 3, Q = 19; };
struct D3Q27 { static constexpr int32_t id = LBM_D3Q27; static constexpr int D = 3, Q = 27; };
struct Float64 { static constexpr int32_t id = LBM_F64; };
struct Float32 { static constexpr int32_t id = LBM_F32; };
struct Float16Shifted { static constexpr int32_t id = LBM_F16S; };
struct BFloat16Shifted { static constexpr int32_t id = LBM_BF16S; };
}  // namespace sch

namespace cmpt {  // component kinds, reference names
struct BGK {};
struct BGKGuo {};
struct Psm {};
struct Hlbm {};
struct FpLbm {};
struct BounceBack {};
template <uint64_t I> struct AntiBounceBack {};
template <bool East> struct ZouHePressureX {};
template <bool East> using ZouHeHorizontal = ZouHePressureX<East>;
template <bool East> struct ZouHeVelocityX {};
struct Equilibrium {};
}  // namespace cmpt

class error : public std::runtime_error {
 public:
  error(lbm_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  lbm_status status;
};

inline void check(lbm_status s) {
  if (s != LBM_OK) throw error(s, lbm_last_error());
}

template <typename T, typename Desc, typename Kind>
class component;

// ---- collisions ---------------------------------------------------------
template <typename T, typename Desc>
class component<T, Desc, cmpt::BGK> {
 public:
  explicit component(double relaxation) : tau_(relaxation) {}
  void fill(lbm_op_desc& d) const { d.collision = LBM_COLL_BGK; d.tau = tau_; }
 private:
  double tau_;
};

template <typename T, typename Desc>
class component<T, Desc, cmpt::BGKGuo> {
 public:
  explicit component(double relaxation, std::array<double, 3> global_force = {0, 0, 0},
                     bool cell_force = false)
      : tau_(relaxation), g_(global_force), cell_(cell_force) {}
  void fill(lbm_op_desc& d) const {
    d.collision = LBM_COLL_BGK_GUO;
    d.tau = tau_;
    for (int k = 0; k < 3; ++k) d.vec[k] = g_[k];
    if (cell_) d.flags |= LBM_OPF_GUO_CELL_FORCE;
  }
 private:
  double tau_;
  std::array<double, 3> g_;
  bool cell_;
};

// cmpt::Hlbm (hlbm.hpp:62-105): BGK towards eq(rho, u*B + N(1-B)/D); the
// particle side is selected with lbm_particle_params::coupling = LBM_COUPLING_HLBM.
template <typename T, typename Desc>
class component<T, Desc, cmpt::Hlbm> {
 public:
  explicit component(double relaxation) : tau_(relaxation) {}
  void fill(lbm_op_desc& d) const {
    d.collision = LBM_COLL_HLBM;
    d.tau = tau_;
  }
 private:
  double tau_;
};

// cmpt::FpLbm (fplbm.hpp:60-133): Guo forcing with the cell force field;
// the particle side is lbm_particle_params::coupling = LBM_COUPLING_FPLBM.
template <typename T, typename Desc>
class component<T, Desc, cmpt::FpLbm> {
 public:
  explicit component(double relaxation) : tau_(relaxation) {}
  void fill(lbm_op_desc& d) const {
    d.collision = LBM_COLL_FPLBM;
    d.tau = tau_;
  }
 private:
  double tau_;
};

template <typename T, typename Desc>
class component<T, Desc, cmpt::Psm> {
 public:
  explicit component(double relaxation) : tau_(relaxation) {}
  // cmpt::OneParticleAt before the collision (particle/particle_opa.hpp)
  component& stamp_disc(std::array<double, 3> pos, double rad, double eps) {
    stamp_ = true; pos_ = pos; rad_ = rad; eps_ = eps; return *this;
  }
  // cmpt::ForceGather after the collision (momentum_gather.hpp)
  component& force_gather() { gather_ = true; return *this; }
  void fill(lbm_op_desc& d) const {
    d.collision = LBM_COLL_PSM;
    d.tau = tau_;
    if (stamp_) {
      d.flags |= LBM_OPF_PSM_STAMP_DISC;
      d.disc[0] = pos_[0]; d.disc[1] = pos_[1]; d.disc[2] = pos_[2]; d.disc[3] = rad_; d.disc[4] = eps_;
    }
    if (gather_) d.flags |= LBM_OPF_PSM_FORCE_GATHER;
  }
 private:
  double tau_;
  bool stamp_ = false, gather_ = false;
  std::array<double, 3> pos_{};
  double rad_ = 0, eps_ = 1.5;
};

// ---- boundaries -----------------------------------------------------------
template <typename T, typename Desc>
class component<T, Desc, cmpt::BounceBack> {
 public:
  void fill(lbm_op_desc& d) const { d.boundary = LBM_BC_BOUNCE_BACK; }
};

template <typename T, typename Desc>
class component<T, Desc, cmpt::AntiBounceBack<0u>> {
 public:
  void fill(lbm_op_desc& d) const { d.boundary = LBM_BC_ANTI_BOUNCE_BACK; }
};

template <typename T, typename Desc, bool East>
class component<T, Desc, cmpt::ZouHePressureX<East>> {
 public:
  // a constant set-point, or the drivers' ramp rho(t) = t < n ? r0 + s*t : r1
  explicit component(double rho) : r0_(rho), s_(0), r1_(rho), n_(0) {}
  component(double rho_start, double slope, double rho_final, uint64_t ramp_steps)
      : r0_(rho_start), s_(slope), r1_(rho_final), n_(ramp_steps) {}
  // D3Q27 only: LBM_OPF_ZOUHE_NORMAL_ONLY, the reconstruction without its
  // tangential terms (no reference counterpart; keeps 3D channels stable)
  component& normal_only(bool on = true) {
    flags_ = on ? LBM_OPF_ZOUHE_NORMAL_ONLY : 0u;
    return *this;
  }
  void fill(lbm_op_desc& d) const {
    d.boundary = LBM_BC_ZOUHE_PRESSURE_X;
    d.east = East ? 1 : 0;
    d.rho_start = r0_; d.rho_slope = s_; d.rho_final = r1_; d.ramp_steps = n_;
    d.flags |= flags_;
  }
 private:
  double r0_, s_, r1_;
  uint64_t n_;
  uint32_t flags_ = 0;
};

template <typename T, typename Desc, bool East>
class component<T, Desc, cmpt::ZouHeVelocityX<East>> {
 public:
  void fill(lbm_op_desc& d) const { d.boundary = LBM_BC_ZOUHE_VELOCITY_X; d.east = East ? 1 : 0; }
};

template <typename T, typename Desc>
class component<T, Desc, cmpt::Equilibrium> {
 public:
  component(double rho, std::array<double, 3> u) : rho_(rho), u_(u) {}
  void fill(lbm_op_desc& d) const {
    d.boundary = LBM_BC_EQUILIBRIUM;
    d.rho_eq = rho_;
    for (int k = 0; k < 3; ++k) d.vec[k] = u_[k];
  }
 private:
  double rho_;
  std::array<double, 3> u_;
};

struct none {
  void fill(lbm_op_desc&) const {}
};

// ---- the lattice ----------------------------------------------------------
template <typename T, typename Desc>
class lattice {
 public:
  explicit lattice(std::array<uint64_t, 3> extent, int32_t device = 0) {
    check(lbm_ctx_create(&ctx_, Desc::id, T::id, extent.data(), device));
    n_ = extent[0] * extent[1] * (Desc::D == 3 ? extent[2] : 1);
  }
  ~lattice() { lbm_ctx_destroy(ctx_); }
  lattice(const lattice&) = delete;
  lattice& operator=(const lattice&) = delete;

  uint64_t cells() const { return n_; }
  void set_flags(const std::vector<uint8_t>& info) { check(lbm_set_flags(ctx_, info.data())); }
  void set_periodic(uint32_t axes) { check(lbm_set_periodic(ctx_, axes)); }

  // what the drivers' `case flag:` does: boundary, then collision
  template <typename B, typename C>
  void bind(uint8_t flag, const B& boundary, const C& collision) {
    lbm_op_desc d{};
    d.rho_start = d.rho_final = 1.0;
    d.rho_eq = 1.0;
    d.tau = 1.0;
    boundary.fill(d);
    collision.fill(d);
    check(lbm_set_operator(ctx_, flag, &d));
  }

  void set_boundary_velocity(const std::vector<double>& u) { check(lbm_set_boundary_velocity(ctx_, u.data())); }
  void init_equilibrium(double rho, std::array<double, 3> u) { check(lbm_init_equilibrium(ctx_, rho, u.data())); }
  void set_populations(const std::vector<double>& f) { check(lbm_set_populations(ctx_, f.data())); }
  std::vector<double> populations() {
    std::vector<double> f(size_t(Desc::Q) * n_);
    check(lbm_get_populations(ctx_, f.data()));
    return f;
  }
  void macros(std::vector<double>& rho, std::vector<double>& u) {
    rho.resize(n_);
    u.resize(size_t(Desc::D) * n_);
    check(lbm_get_macros(ctx_, rho.data(), u.data()));
  }
  void step(uint64_t t0, uint64_t n) { check(lbm_step(ctx_, t0, n)); }
  void sync() { check(lbm_sync(ctx_)); }
  double last_step_ms() {
    double ms = 0;
    check(lbm_last_step_ms(ctx_, &ms));
    return ms;
  }
  lbm_ctx* handle() { return ctx_; }

 private:
  lbm_ctx* ctx_ = nullptr;
  uint64_t n_ = 0;
};

}  // namespace lbm_amd
