// lbm_oracle.cpp — CPU restatement of the reference's hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker
// (or as the timed CPU baseline).  The product (liblbm_mirror_amd.so) never
// links, loads or calls it.
//
// What it restates (keldu/lbm_mirror, paths relative to the reference root):
//   lattice tables        modules/core/c++/descriptor/{d2q9,d3q19,d3q27}.hpp
//   equilibrium           modules/core/c++/equilibrium.hpp:7-48
//   compute_rho_u         modules/core/c++/macroscopic.hpp:10-35
//   BGK / BGKGuo          modules/core/c++/collision.hpp:19-97, 99-205
//   Stream                modules/core/c++/stream.hpp:19-58
//   BounceBack / ABB / Equilibrium / Zou-He pressure (D2Q9, D3Q27) / Zou-He
//   velocity (D2Q9)       modules/core/c++/boundary.hpp:32-805
//   PsmReset / Psm / PsmOneParticleGather   modules/core/c++/psm.hpp:20-446
//   PSM (cell force)      modules/core/c++/component/psm.hpp:16-80
//   HlbmReset / Hlbm / HlbmOneParticleGather / HlbmOneParticleStamp
//                         modules/core/c++/hlbm.hpp:28-49, 62-105, 111-256, 264-308;
//                         driver order examples/poiseulle_particles_2d_gpu/step.hpp:21-43
//   FpLbmReset / FpLbm / FpLbmOneParticle   modules/core/c++/fplbm.hpp:21-58, 60-133, 135-247;
//                         compute_rho_momentum  macroscopic.hpp:41-60;
//                         driver order examples/moving_poiseulle_particles_2d_fplbm_gpu/sim.cpp:185-258
//   ForceGather           modules/core/c++/momentum_gather.hpp:15-52
//   OneParticleAt         modules/core/c++/particle/particle_opa.hpp:14-46
//   particle porosity     modules/core/c++/particle/porosity.hpp:56-89
//   particle AABB         modules/core/c++/particle/aabb.hpp:41-84
//   DEM contact           modules/core/c++/particle/contact.hpp:76-567
//   integrate / reset     modules/core/c++/particle/integrate.hpp:16-83,195-296
//   ParticleAxisIntegrate modules/core/c++/particle/integrate.hpp:101-189,298-327
//                         (normalise_spin_axis, math::normalize abstract/math.hpp:70-87)
//   Verlet step           modules/core/c++/particle/particle.hpp:171-211
//   spheroid mass         modules/core/c++/particle/particle.hpp:79-101
//   cell list / couple    modules/core/c++/particle/cell_list.hpp:23-206,
//                         modules/core/c++/particle/couple.hpp:26-105
//   driver step order     examples/poiseulle_2d_gpu/poiseulle_2d_gpu.cpp:158-399,
//                         examples/moving_poiseulle_particles_2d_psm_gpu/sim.cpp:176-280
//
// Arithmetic follows the reference operation by operation: every binary
// operation of a `data<T>` rounds to T (abstract/data.hpp:170-188), sums run
// in the reference's order, and expressions that the reference writes with
// double literals on float operands are written with the same shapes here so
// that C++ promotes them the same way (equilibrium bracket, DEM damping,
// contact point, disc inertia).  Build with -ffp-contract=off.
//
// Two switches depart from the reference on purpose, off by default, for
// comparing the GPU at tighter tolerances than the reference's own
// arithmetic allows: SimBase::sum64 applies the particle force/torque sums
// as FP64 sums rounded to T once (the GPU's arithmetic) instead of the
// reference's serial T sums (in FP64 the same sums in the same order, a CPU
// test holds them bitwise equal; the FP64 sums are pinned by brute force);
// SimBase::series_sin takes the porosity ramp's sine from the engine's
// series instead of the C library (pinned to one ulp of the correctly
// rounded sine, and equal to it after rounding to float, by a CPU test).
//
// Layout is the reference's: two AoS population buffers ("dfs_old", "dfs")
// with a one-cell ghost ring that stays zero (chunk.hpp:25-35), selected by
// time-step parity (collision.hpp:61-64, stream.hpp:28-31).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "../include/lbm_mirror_amd/lbm.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- lattices
struct D2Q9 {
  static constexpr int D = 2, Q = 9;
  static constexpr int c[9][3] = {{0, 0, 0},  {-1, 0, 0}, {1, 0, 0},
                                  {0, -1, 0}, {0, 1, 0},  {-1, -1, 0},
                                  {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}};
  static constexpr double w[9] = {4. / 9.,  1. / 9.,  1. / 9.,
                                  1. / 9.,  1. / 9.,  1. / 36.,
                                  1. / 36., 1. / 36., 1. / 36.};
  static constexpr int opp[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
};
struct D3Q19 {
  static constexpr int D = 3, Q = 19;
  static constexpr int c[19][3] = {
      {0, 0, 0},   {-1, 0, 0},  {1, 0, 0},   {0, -1, 0}, {0, 1, 0},
      {0, 0, -1},  {0, 0, 1},   {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
      {1, 1, 0},   {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1}, {1, 0, 1},
      {0, -1, -1}, {0, 1, -1},  {0, -1, 1},  {0, 1, 1}};
  static constexpr double w[19] = {
      1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
      1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
      1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
  static constexpr int opp[19] = {0,  2,  1,  4,  3,  6,  5,  10, 9, 8,
                                  7,  14, 13, 12, 11, 18, 17, 16, 15};
};
struct D3Q27 {
  static constexpr int D = 3, Q = 27;
  static constexpr int c[27][3] = {
      {0, 0, 0},    {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {-1, -1, 0},
      {1, -1, 0},   {0, 1, 0},   {-1, 1, 0},  {1, 1, 0},   {0, 0, -1},
      {-1, 0, -1},  {1, 0, -1},  {0, -1, -1}, {-1, -1, -1}, {1, -1, -1},
      {0, 1, -1},   {-1, 1, -1}, {1, 1, -1},  {0, 0, 1},   {-1, 0, 1},
      {1, 0, 1},    {0, -1, 1},  {-1, -1, 1}, {1, -1, 1},  {0, 1, 1},
      {-1, 1, 1},   {1, 1, 1}};
  static constexpr double w[27] = {
      8. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  1. / 54.,  1. / 54.,
      2. / 27.,  1. / 54.,  1. / 54.,  2. / 27.,  1. / 54.,  1. / 54.,
      1. / 54.,  1. / 216., 1. / 216., 1. / 54.,  1. / 216., 1. / 216.,
      2. / 27.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 216., 1. / 216.,
      1. / 54.,  1. / 216., 1. / 216.};
  static constexpr int opp[27] = {0,  2,  1,  6,  8,  7,  3,  5,  4,
                                  18, 20, 19, 24, 26, 25, 21, 23, 22,
                                  9,  11, 10, 15, 17, 16, 12, 14, 13};
};

// ------------------------------------------------------- moment functions
// compute_rho_u (macroscopic.hpp:10-35): ρ and u from 0 in population order,
// u_d += T(c_jd)·f_j, then u_d /= ρ.
template <class T, class L>
void rho_u(const T* f, T& rho, T* u) {
  rho = T(0);
  for (int d = 0; d < L::D; ++d) u[d] = T(0);
  for (int j = 0; j < L::Q; ++j) {
    rho = T(rho + f[j]);
    for (int d = 0; d < L::D; ++d) u[d] = T(u[d] + T(T(L::c[j][d]) * f[j]));
  }
  for (int d = 0; d < L::D; ++d) u[d] = T(u[d] / rho);
}

// equilibrium (equilibrium.hpp:7-48).  w_i and inv_cs2 are stored as T
// (descriptor/d2q9.hpp:31-56); the bracket has double literals, so for
// T = float it is evaluated in double with vc·vc rounded to float first,
// and the product is narrowed once by set().
template <class T, class L>
void equilibrium(T rho, const T* u, T* eq) {
  const T inv_cs2 = T(3.0);
  T vv = T(0.0);
  for (int j = 0; j < L::D; ++j) vv = T(vv + T(u[j] * u[j]));
  for (int i = 0; i < L::Q; ++i) {
    T vc = T(0);
    for (int j = 0; j < L::D; ++j) vc = T(vc + T(u[j] * T(L::c[i][j])));
    T a = T(vc * inv_cs2);
    const T wi = T(L::w[i]);
    T wr = T(wi * rho);
    T aa = T(a * a);
    eq[i] = T(wr * (1.0 + a - inv_cs2 * 0.5 * vv + aa * 0.5));
  }
}

// ------------------------------------------------------------ particles
template <class T>
struct Part {
  T pos[3], pos_old[3], rot[3], rot_old[3], acc[3], ang_acc[3];
  T force[3], torque[3];
  T mass, inertia, density, radius;
};

// sin on [0, pi/2] by series (x + x^3 P(x^2) up to x^17 below pi/4,
// cos(pi/2 - x) = 1 + z Q(z) up to z^8 above, pi/2 - x with pi/2 in two
// doubles; coefficients the doubles nearest (-1)^k/n!).  The engine's own
// sine for the porosity ramp; used here only where SimBase::series_sin is
// set, to compare the GPU's porosity bit for bit.  Within one ulp of the
// correctly rounded sin on the ramp's range (tests/test_oracle.py).
inline double series_sin_q1(double x) {
  static const double kS[8] = {-1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0,
                               1.0 / 6227020800.0, -1.0 / 1307674368000.0, 1.0 / 355687428096000.0};
  static const double kC[8] = {-1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0,
                               1.0 / 479001600.0, -1.0 / 87178291200.0, 1.0 / 20922789888000.0};
  if (x > 0.78539816339744831) {
    const double r = (1.5707963267948966 - x) + 6.123233995736766e-17;
    const double z = r * r;
    double q = kC[7];
    for (int k = 6; k >= 0; --k) q = kC[k] + z * q;
    return 1.0 + z * q;
  }
  const double z = x * x;
  double p = kS[7];
  for (int k = 6; k >= 0; --k) p = kS[k] + z * p;
  return x + x * (z * p);
}

// particle_porosity<..., ParticleSpheroid>::calculate (porosity.hpp:56-89)
template <class T, int D>
T porosity(const T* rel, T rad, T eps, bool series_sin = false) {
  T s2 = T(0);
  for (int d = 0; d < D; ++d) s2 = T(s2 + T(rel[d] * rel[d]));
  T s = std::sqrt(s2);
  T eps_h = T(eps / 2);
  T lo = T(rad - eps_h), hi = T(rad + eps_h);
  if (s <= lo) return T(0);
  if (s >= hi) return T(1);
  T inner = (M_PI / (2 * eps)) * T(s - lo);
  T si = series_sin ? T(series_sin_q1(static_cast<double>(inner))) : std::sin(inner);
  return T(si * si);
}

// particle_aabb::try_calculate (aabb.hpp:41-84): [a, b) per axis.
template <class T, int D>
void aabb(const T* pos, T r, const uint64_t* extent, uint64_t* a, uint64_t* b) {
  for (int d = 0; d < D; ++d) {
    T lower = (pos[d] >= r) ? T(pos[d] - r) : T(0);
    a[d] = static_cast<uint64_t>(lower);
    T upper = T(pos[d] + r);
    uint64_t bi = static_cast<uint64_t>(T(upper + T(1)));
    b[d] = bi < extent[d] ? bi : extent[d];
  }
}

inline uint64_t sub_steps_of(uint64_t s) { return s == 0 ? 1 : s; }

// impl::safe_normalize (contact.hpp:82-98)
template <class T, int D>
bool safe_normalize(const T* in, T* out) {
  T l2 = T(0);
  for (int d = 0; d < D; ++d) l2 = T(l2 + T(in[d] * in[d]));
  if (!(l2 > 0)) return false;
  T inv = T(1) / std::sqrt(l2);
  for (int d = 0; d < D; ++d) out[d] = T(in[d] * inv);
  return true;
}

template <class T, int D>
struct Manifold {
  T depth;
  T n[3];
  T point[3];
};

// cross products (abstract/math.hpp:205-245); 2D returns the scalar in r[0].
template <class T, int D>
void cross_rf(const T* r, const T* f, T* out) {
  if (D == 2) {
    out[0] = T(T(r[0] * f[1]) - T(r[1] * f[0]));
  } else {
    out[0] = T(T(r[1] * f[2]) - T(r[2] * f[1]));
    out[1] = T(T(r[2] * f[0]) - T(r[0] * f[2]));
    out[2] = T(T(r[0] * f[1]) - T(r[1] * f[0]));
  }
}

// narrow_phase<Spheroid>::contact (contact.hpp:124-184), canonical (lo, hi)
template <class T, int D>
bool narrow(const Part<T>& lo, const Part<T>& hi, Manifold<T, D>& m) {
  T d[3];
  for (int i = 0; i < D; ++i) d[i] = T(lo.pos[i] - hi.pos[i]);
  T len2 = T(0);
  for (int i = 0; i < D; ++i) len2 = T(len2 + T(d[i] * d[i]));
  T r_lo = lo.radius, r_hi = hi.radius;
  T sum_r = T(r_lo + r_hi);
  if (!(len2 < T(sum_r * sum_r))) return false;
  T len = std::sqrt(len2);
  T depth = T(sum_r - len);
  T cap = (r_lo < r_hi) ? r_lo : r_hi;
  if (depth > cap) depth = cap;
  if (!safe_normalize<T, D>(d, m.n)) {
    for (int i = 0; i < D; ++i) m.n[i] = T(0);
    m.n[0] = T(1);
  }
  m.depth = depth;
  for (int i = 0; i < D; ++i) m.point[i] = T(hi.pos[i] + m.n[i] * (r_hi - depth * 0.5));
  return true;
}

// impl::surface_velocity (contact.hpp:271-311)
template <class T, int D>
void surface_velocity(const Part<T>& p, const T* r, T steps, T* v) {
  for (int i = 0; i < D; ++i) v[i] = T(T(p.pos[i] - p.pos_old[i]) * steps);
  if (D == 2) {
    T om = T(T(p.rot[0] - p.rot_old[0]) * steps);
    T tang[2] = {T(-T(om * r[1])), T(om * r[0])};
    for (int i = 0; i < 2; ++i) v[i] = T(v[i] + tang[i]);
  } else {
    T om[3];
    for (int i = 0; i < 3; ++i) om[i] = T(T(p.rot[i] - p.rot_old[i]) * steps);
    T tang[3];
    cross_rf<T, 3>(om, r, tang);
    for (int i = 0; i < 3; ++i) v[i] = T(v[i] + tang[i]);
  }
}

struct Dem {
  double k_n, zeta_n, gamma_t, mu;
};

// impl::soft_shell_response (contact.hpp:204-264).  The damping and normal
// force lines carry double literals; for T = float they are double values.
template <class T, int D>
void soft_shell(const Manifold<T, D>& m, const T* r_self, const T* v_rel, T m_eff,
                const Dem& dem, T* force, T* torque) {
  const T* n = m.n;
  T depth = m.depth;
  T k_n = T(dem.k_n), zeta = T(dem.zeta_n), gamma_t = T(dem.gamma_t), mu = T(dem.mu);
  auto g_n = 2.0 * zeta * std::sqrt(T(m_eff * k_n));
  T v_n = T(0);
  for (int i = 0; i < D; ++i) v_n = T(v_n + T(v_rel[i] * n[i]));
  auto f_n = T(k_n * depth) - g_n * v_n;
  if (f_n < 0) f_n = 0;
  T v_t[3];
  for (int i = 0; i < D; ++i) v_t[i] = T(v_rel[i] - T(v_n * n[i]));
  T t_hat[3];
  bool have_t = safe_normalize<T, D>(v_t, t_hat);
  T vt2 = T(0);
  for (int i = 0; i < D; ++i) vt2 = T(vt2 + T(v_t[i] * v_t[i]));
  T v_t_len = std::sqrt(vt2);
  T f_t = T(gamma_t * v_t_len);
  auto f_t_cap = mu * f_n;
  if (f_t > f_t_cap) f_t = f_t_cap;
  for (int i = 0; i < D; ++i) {
    auto v = f_n * n[i];
    if (have_t) v -= T(f_t * t_hat[i]);
    force[i] = T(v);
  }
  cross_rf<T, D>(r_self, force, torque);
}

// particle_contact_response (contact.hpp:321-364)
template <class T, int D>
void pair_response(Manifold<T, D> m, const Part<T>& self, const Part<T>& other,
                   bool self_is_lo, const Dem& dem, uint64_t sub_steps, T* force,
                   T* torque) {
  T steps = T(sub_steps_of(sub_steps));
  if (!self_is_lo)
    for (int i = 0; i < D; ++i) m.n[i] = T(-m.n[i]);
  T r_s[3], r_o[3];
  for (int i = 0; i < D; ++i) {
    r_s[i] = T(m.point[i] - self.pos[i]);
    r_o[i] = T(m.point[i] - other.pos[i]);
  }
  T v_s[3], v_o[3], v_rel[3];
  surface_velocity<T, D>(self, r_s, steps, v_s);
  surface_velocity<T, D>(other, r_o, steps, v_o);
  for (int i = 0; i < D; ++i) v_rel[i] = T(v_s[i] - v_o[i]);
  T m_eff = T(T(self.mass * other.mass) / T(self.mass + other.mass));
  soft_shell<T, D>(m, r_s, v_rel, m_eff, dem, force, torque);
}

// broadphase_collision_check (particle.hpp:247-279)
template <class T, int D>
bool broadphase(const Part<T>& l, const Part<T>& r) {
  T n2 = T(0);
  for (int i = 0; i < D; ++i) {
    T dd = T(l.pos[i] - r.pos[i]);
    n2 = T(n2 + T(dd * dd));
  }
  T rl = l.radius, rr = r.radius;
  T rab2 = T(T(T(rl * rl) + T(rr * rr)) + T(T(rr * rl) * T(2.0)));
  return n2 < rab2;
}

template <class T, int D>
void add_response(Part<T>& self, const T* f, const T* tq) {
  for (int i = 0; i < D; ++i) self.force[i] = T(self.force[i] + f[i]);
  if (D == 2) self.torque[0] = T(self.torque[0] + tq[0]);
}

// accumulate_particle_contacts (contact.hpp:466-512), reading `ps` (the
// pre-contact state of every particle) and adding into `self`.
template <class T, int D>
void particle_contacts(const std::vector<Part<T>>& ps, uint32_t si, Part<T>& self,
                       const Dem& dem, uint64_t sub_steps) {
  const Part<T>& me = ps[si];
  for (uint32_t j = 0; j < ps.size(); ++j) {
    if (j == si) continue;
    const Part<T>& other = ps[j];
    if (!broadphase<T, D>(me, other)) continue;
    bool self_is_lo = si < j;
    const Part<T>& lo = self_is_lo ? me : other;
    const Part<T>& hi = self_is_lo ? other : me;
    Manifold<T, D> m;
    if (!narrow<T, D>(lo, hi, m)) continue;
    T f[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    pair_response<T, D>(m, me, other, self_is_lo, dem, sub_steps, f, tq);
    add_response<T, D>(self, f, tq);
  }
}

// accumulate_box_wall_contacts (contact.hpp:407-456)
template <class T, int D>
void wall_contacts(const Part<T>& me, Part<T>& self, const uint64_t* extent,
                   const Dem& dem, uint64_t sub_steps) {
  T r = me.radius;
  for (int ax = 0; ax < D; ++ax) {
    T x = me.pos[ax];
    T hi_bound = static_cast<T>(extent[ax]);
    for (int side = 0; side < 2; ++side) {
      T depth = (side == 0) ? T(r - x) : T(x - T(hi_bound - r));
      if (!(depth > 0)) continue;
      if (depth > r) depth = r;
      Manifold<T, D> m;
      for (int i = 0; i < D; ++i) {
        m.n[i] = T(0);
        m.point[i] = me.pos[i];
      }
      m.n[ax] = (side == 0) ? T(1.0) : T(-1.0);
      m.point[ax] = (side == 0) ? T(0.0) : hi_bound;
      m.depth = depth;
      // wall_contact_response (contact.hpp:374-398)
      T steps = T(sub_steps_of(sub_steps));
      T r_s[3], v_rel[3];
      for (int i = 0; i < D; ++i) r_s[i] = T(m.point[i] - me.pos[i]);
      surface_velocity<T, D>(me, r_s, steps, v_rel);
      T f[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
      soft_shell<T, D>(m, r_s, v_rel, me.mass, dem, f, tq);
      add_response<T, D>(self, f, tq);
    }
  }
}

// verlet sub-steps shared by both integrators (particle.hpp:171-211)
template <class T, int D>
void verlet_steps(Part<T>& p, uint64_t sub_steps);

// normalise_spin_axis (integrate.hpp:115-131): the sign in 2D; the axis over
// its length in T in 3D (math::normalize: dot from zero, a zero vector kept)
template <class T, int D>
void spin_axis(const double* in, T* out) {
  out[0] = out[1] = out[2] = T(0);
  if (D == 2) {
    out[0] = T(in[0]) < T(0) ? T(-1) : T(1);
    return;
  }
  T a[3] = {T(in[0]), T(in[1]), T(in[2])};
  T dot = T(0);
  for (int i = 0; i < 3; ++i) dot = T(dot + T(a[i] * a[i]));
  if (dot <= T(0)) {
    for (int i = 0; i < 3; ++i) out[i] = a[i];
    return;
  }
  T len = T(std::sqrt(dot));
  for (int i = 0; i < 3; ++i) out[i] = T(a[i] / len);
}

// integrate_particle_fixed_axis (integrate.hpp:143-190): acceleration 0
// (pinned), angular acceleration axis * (axis . torque) / I
template <class T, int D>
void integrate_fixed_axis(Part<T>& p, const T* axis, uint64_t sub_steps) {
  for (int i = 0; i < D; ++i) p.acc[i] = T(0);
  if (D == 2) {
    p.ang_acc[0] = T(T(axis[0] * p.torque[0]) / p.inertia);
  } else {
    T tau = T(0);
    for (int i = 0; i < 3; ++i) tau = T(tau + T(axis[i] * p.torque[i]));
    T alpha = T(tau / p.inertia);
    for (int i = 0; i < 3; ++i) p.ang_acc[i] = T(axis[i] * alpha);
  }
  verlet_steps<T, D>(p, sub_steps);
}

// integrate_particle + verlet_step_lambda (integrate.hpp:45-83,
// particle.hpp:171-211)
// keep_ang: FpLbmOneParticle sets only the linear acceleration (fplbm.hpp:233-236)
template <class T, int D>
void integrate(Part<T>& p, const T* body_acc, uint64_t sub_steps, bool keep_ang = false) {
  for (int i = 0; i < D; ++i) p.acc[i] = T(T(p.force[i] / p.mass) + body_acc[i]);
  if (D == 2 && !keep_ang) p.ang_acc[0] = T(p.torque[0] / p.inertia);
  verlet_steps<T, D>(p, sub_steps);
}

template <class T, int D>
void verlet_steps(Part<T>& p, uint64_t sub_steps) {
  uint64_t steps = sub_steps_of(sub_steps);
  T ts = T(T(1) / static_cast<T>(steps));
  const int R = (D == 2) ? 1 : 3;
  for (uint64_t s = 0; s < steps; ++s) {
    T ts2 = T(ts * ts);
    T pn[3];
    for (int i = 0; i < D; ++i) pn[i] = T(T(T(p.pos[i] * T(2.0)) - p.pos_old[i]) + T(p.acc[i] * ts2));
    for (int i = 0; i < D; ++i) {
      p.pos_old[i] = p.pos[i];
      p.pos[i] = pn[i];
    }
    T rn[3];
    for (int i = 0; i < R; ++i) rn[i] = T(T(T(p.rot[i] * T(2.0)) - p.rot_old[i]) + T(p.ang_acc[i] * ts2));
    for (int i = 0; i < R; ++i) {
      p.rot_old[i] = p.rot[i];
      p.rot[i] = rn[i];
    }
  }
}

// create_spheroid_particle_group mass/inertia (particle.hpp:79-101)
template <class T, int D>
void spheroid_mass(T r, T rho, T& m_out, T& i_out) {
  constexpr T pi = T(3.14159265358979323846);
  if (D == 2) {
    auto m = T(T(T(pi * r) * r) * rho);
    m_out = m;
    i_out = T(0.5 * m * r * r);
  } else {
    auto m = (4.0 / 3.0) * pi * r * r * r * rho;
    m_out = T(m);
    i_out = T(0.4 * m * r * r);
  }
}

// ------------------------------------------------------------ cell list
template <class T, int D>
struct CellList {
  T bin = T(1);
  uint64_t grid[3] = {1, 1, 1};
  std::vector<uint64_t> count, slot;
  uint64_t cap = 4, overflow = 0;

  void make(const uint64_t* extent, T max_radius, uint64_t capacity) {
    cap = capacity;
    T b = T(max_radius * 2);
    if (!(b > 1)) b = T(1);
    bin = b;
    uint64_t cells = 1;
    for (int i = 0; i < D; ++i) {
      T e = static_cast<T>(extent[i]);
      grid[i] = static_cast<uint64_t>(T(e / bin)) + 1u;
      cells *= grid[i];
    }
    count.assign(cells, 0);
    slot.assign(cells * cap, 0);
  }
  uint64_t bin_of(const T* pos) const {
    uint64_t flat = 0, stride = 1;
    for (int i = 0; i < D; ++i) {
      T x = T(pos[i] / bin);
      int64_t b = static_cast<int64_t>(x);
      if (b < 0) b = 0;
      if (static_cast<uint64_t>(b) >= grid[i]) b = static_cast<int64_t>(grid[i]) - 1;
      flat += static_cast<uint64_t>(b) * stride;
      stride *= grid[i];
    }
    return flat;
  }
  void build(const std::vector<Part<T>>& ps) {
    std::fill(count.begin(), count.end(), 0);
    overflow = 0;
    for (uint64_t p = 0; p < ps.size(); ++p) {
      uint64_t b = bin_of(ps[p].pos);
      uint64_t c = count[b];
      if (c >= cap) {
        ++overflow;
        continue;
      }
      slot[b * cap + c] = p;
      count[b] = c + 1;
    }
  }
  template <class F>
  void for_each_candidate(const uint64_t* cell, F&& f) const {
    int64_t base[3];
    for (int i = 0; i < D; ++i) base[i] = static_cast<int64_t>(T(static_cast<T>(cell[i]) / bin));
    uint64_t total = 1;
    for (int i = 0; i < D; ++i) total *= 3;
    for (uint64_t off = 0; off < total; ++off) {
      uint64_t rem = off, flat = 0, stride = 1;
      bool inside = true;
      for (int i = 0; i < D; ++i) {
        int64_t delta = static_cast<int64_t>(rem % 3u) - 1;
        rem /= 3u;
        int64_t b = base[i] + delta;
        int64_t g = static_cast<int64_t>(grid[i]);
        if (b < 0 || b >= g) {
          inside = false;
          break;
        }
        flat += static_cast<uint64_t>(b) * stride;
        stride *= static_cast<uint64_t>(g);
      }
      if (!inside) continue;
      for (uint64_t s = 0; s < count[flat]; ++s) f(slot[flat * cap + s]);
    }
  }
};

// ------------------------------------------------------------ simulation
struct SimBase {
  // Particle force and torque sums: false (default) the reference's serial
  // sums in T (psm.hpp:404-443, hlbm.hpp:234,249-252, fplbm.hpp:224-236);
  // true the FP64 sums of the same terms, rounded to T once (the GPU's
  // arithmetic, which sums in FP64 by a tree; the order of an FP64 sum
  // changes its last bits only).
  bool sum64 = false;
  // The porosity ramp's sine: false (default) the C library's std::sin, as
  // the reference calls it (porosity.hpp:82); true the engine's series
  // (series_sin_q1), so the GPU's porosity can be compared bit for bit.
  bool series_sin = false;
  virtual ~SimBase() {}
  virtual int set_flags(const uint8_t*) = 0;
  virtual void set_periodic(uint32_t) = 0;
  virtual void set_operator(uint8_t, const lbm_op_desc*) = 0;
  virtual void set_boundary_velocity(const double*) = 0;
  virtual void set_force(const double*) = 0;
  virtual void get_force(double*) const = 0;
  virtual void set_porosity(const double*) = 0;
  virtual void get_porosity(double*) const = 0;
  virtual void set_populations(uint64_t, const double*) = 0;
  virtual void get_populations(uint64_t, double*) const = 0;
  virtual void init_equilibrium(uint64_t, double, const double*) = 0;
  virtual void get_macros(double*, double*) const = 0;
  virtual void set_macros(const double*, const double*) = 0;
  virtual void step(uint64_t, uint64_t, int) = 0;
  virtual int particles_set(const lbm_particle*, uint32_t, const lbm_particle_params*) = 0;
  virtual void particles_get(lbm_particle*, uint32_t) const = 0;
  virtual void psm_couple(uint64_t, double*, double*, double*) = 0;
  virtual void dem_contact() = 0;
  virtual void integrate_all() = 0;
  virtual uint64_t overflow() const = 0;
  virtual void get_hlbm(double*, double*) const = 0;
};

template <class T, class L>
struct Sim final : SimBase {
  static constexpr int D = L::D, Q = L::Q;
  uint64_t n[3];
  uint64_t gn[3];  // ghost-padded extents (2D: gn[2] = 1)
  uint64_t N, G;
  std::vector<T> buf[2];  // [0] "dfs_old", [1] "dfs"
  std::vector<uint8_t> info;
  std::vector<T> rho, vel, force, por, bc_vel;
  std::vector<T> pN, pD;  // HLBM particle velocity sums N (D per cell) and weights D
  uint32_t periodic = 0;
  lbm_op_desc ops[256];
  bool has_op[256];
  // particles
  std::vector<Part<T>> parts;
  lbm_particle_params pp;
  CellList<T, D> cl;
  std::vector<double> F64, T64, ABS;

  explicit Sim(const uint64_t* ext) {
    for (int d = 0; d < 3; ++d) n[d] = (d < D) ? ext[d] : 1;
    for (int d = 0; d < 3; ++d) gn[d] = (d < D) ? n[d] + 2 : 1;
    N = n[0] * n[1] * n[2];
    G = gn[0] * gn[1] * gn[2];
    buf[0].assign(G * Q, T(0));
    buf[1].assign(G * Q, T(0));
    info.assign(N, 0);
    rho.assign(N, T(0));
    vel.assign(N * D, T(0));
    force.assign(N * D, T(0));
    por.assign(N, T(1));
    pN.assign(N * D, T(0));
    pD.assign(N, T(0));
    bc_vel.assign(N * D, T(0));
    std::memset(ops, 0, sizeof(ops));
    std::memset(has_op, 0, sizeof(has_op));
    std::memset(&pp, 0, sizeof(pp));
  }

  uint64_t cell(const uint64_t* x) const { return x[0] + n[0] * (x[1] + n[1] * x[2]); }
  uint64_t gcell(const uint64_t* x) const {  // interior coordinate → ghost-padded
    uint64_t g[3];
    for (int d = 0; d < 3; ++d) g[d] = (d < D) ? x[d] + 1 : 0;
    return g[0] + gn[0] * (g[1] + gn[1] * g[2]);
  }
  // Neighbour x + s·c (s = ±1) as a ghost-padded index: outside the lattice
  // it is the zero ghost ring, or wraps along periodic axes.
  bool neighbour(const uint64_t* x, const int* c, int s, uint64_t* out) const {
    for (int d = 0; d < 3; ++d) {
      if (d >= D) {
        out[d] = 0;
        continue;
      }
      int64_t v = static_cast<int64_t>(x[d]) + s * c[d];
      if (v < 0 || v >= static_cast<int64_t>(n[d])) {
        if (periodic & (1u << d)) {
          v = (v + static_cast<int64_t>(n[d])) % static_cast<int64_t>(n[d]);
        } else {
          return false;
        }
      }
      out[d] = static_cast<uint64_t>(v);
    }
    return true;
  }
  void coords(uint64_t c, uint64_t* x) const {
    x[0] = c % n[0];
    x[1] = (c / n[0]) % n[1];
    x[2] = c / (n[0] * n[1]);
  }
  T* A(uint64_t t) { return buf[(t % 2 == 0) ? 0 : 1].data(); }
  const T* A(uint64_t t) const { return buf[(t % 2 == 0) ? 0 : 1].data(); }
  T* B(uint64_t t) { return buf[(t % 2 == 0) ? 1 : 0].data(); }

  int set_flags(const uint8_t* f) override {
    info.assign(f, f + N);
    return 0;
  }
  void set_periodic(uint32_t m) override { periodic = m; }
  void set_operator(uint8_t flag, const lbm_op_desc* op) override {
    ops[flag] = *op;
    has_op[flag] = true;
  }
  void set_boundary_velocity(const double* u) override {
    for (uint64_t c = 0; c < N; ++c)
      for (int d = 0; d < D; ++d) bc_vel[c * D + d] = T(u[d * N + c]);
  }
  // fields at the boundary are d*N + cell; stored cell*D + d
  void set_force(const double* f) override {
    for (uint64_t c = 0; c < N; ++c)
      for (int d = 0; d < D; ++d) force[c * D + d] = T(f[d * N + c]);
  }
  void get_force(double* f) const override {
    for (uint64_t c = 0; c < N; ++c)
      for (int d = 0; d < D; ++d) f[d * N + c] = force[c * D + d];
  }
  void set_porosity(const double* p) override {
    for (uint64_t c = 0; c < N; ++c) por[c] = T(p[c]);
  }
  void get_porosity(double* p) const override {
    for (uint64_t c = 0; c < N; ++c) p[c] = por[c];
  }
  void set_populations(uint64_t t, const double* f) override {
    T* a = A(t);
    uint64_t x[3];
    for (uint64_t c = 0; c < N; ++c) {
      coords(c, x);
      uint64_t g = gcell(x);
      for (int i = 0; i < Q; ++i) a[g * Q + i] = T(f[i * N + c]);
    }
  }
  void get_populations(uint64_t t, double* f) const override {
    const T* a = A(t);
    uint64_t x[3];
    for (uint64_t c = 0; c < N; ++c) {
      coords(c, x);
      uint64_t g = gcell(x);
      for (int i = 0; i < Q; ++i) f[i * N + c] = a[g * Q + i];
    }
  }
  void init_equilibrium(uint64_t t, double r, const double* u) override {
    T rr = T(r), uu[3] = {T(u[0]), T(u[1]), T(u[2])}, eq[27];
    equilibrium<T, L>(rr, uu, eq);
    T* a = A(t);
    uint64_t x[3];
    for (uint64_t c = 0; c < N; ++c) {
      coords(c, x);
      uint64_t g = gcell(x);
      for (int i = 0; i < Q; ++i) a[g * Q + i] = eq[i];
      rho[c] = rr;
      for (int d = 0; d < D; ++d) vel[c * D + d] = uu[d];
    }
  }
  void get_macros(double* r, double* u) const override {
    for (uint64_t c = 0; c < N; ++c) {
      if (r) r[c] = rho[c];
      if (u)
        for (int d = 0; d < D; ++d) u[d * N + c] = vel[c * D + d];
    }
  }
  void set_macros(const double* r, const double* u) override {
    for (uint64_t c = 0; c < N; ++c) {
      if (r) rho[c] = T(r[c]);
      if (u)
        for (int d = 0; d < D; ++d) vel[c * D + d] = T(u[d * N + c]);
    }
  }

  // --- boundary components -------------------------------------------------
  // ZouHePressureX<Dir> for D2Q9 (boundary.hpp:175-227)
  static void zouhe_p_d2q9(T* f, bool dir, T rs) {
    T rvx;
    if (dir) {
      T S = T(T(T(f[0] + f[3]) + f[4]) + T(T(T(f[1] + f[5]) + f[7]) * T(2u)));
      rvx = T(rs - S);
    } else {
      T S = T(T(T(f[0] + f[3]) + f[4]) + T(T(T(f[2] + f[6]) + f[8]) * T(2u)));
      rvx = T(S - rs);
    }
    const T c23 = T(2.0 / 3.0), c16 = T(1.0 / 6.0), h = T(0.5);
    if (dir) {
      f[2] = T(f[1] + T(c23 * rvx));
      f[6] = T(T(f[5] + T(c16 * rvx)) + T(h * T(f[3] - f[4])));
      f[8] = T(T(f[7] + T(c16 * rvx)) + T(h * T(f[4] - f[3])));
    } else {
      f[1] = T(f[2] - T(c23 * rvx));
      f[5] = T(T(f[6] - T(c16 * rvx)) + T(h * T(f[4] - f[3])));
      f[7] = T(T(f[8] - T(c16 * rvx)) + T(h * T(f[3] - f[4])));
    }
  }
  // ZouHePressureX<East> for D3Q27 (boundary.hpp:229-756); AtXMax = !East.
  // normal_only: LBM_OPF_ZOUHE_NORMAL_ONLY, the tangential terms dropped
  // (no reference counterpart)
  static void zouhe_p_d3q27(T* f, bool east, T rs, bool normal_only = false) {
    const bool at_max = !east;
    const T two = T(2.0);
    T S0 = T(T(T(T(T(T(T(T(f[0] + f[3]) + f[6]) + f[9]) + f[12]) + f[15]) + f[18]) + f[21]) + f[24]);
    T S, ux, uy, uz;
    if (at_max) {
      T k = T(T(T(T(T(T(T(T(f[2] + f[5]) + f[8]) + f[11]) + f[14]) + f[17]) + f[20]) + f[23]) + f[26]);
      S = T(S0 + T(two * k));
      ux = T(S - rs);
      T y0 = T(T(T(T(T(T(-f[3]) + f[6]) - f[12]) + f[15]) - f[21]) + f[24]);
      T y1 = T(T(T(T(T(T(-f[5]) + f[8]) - f[14]) + f[17]) - f[23]) + f[26]);
      uy = T(y0 + T(two * y1));
      T z0 = T(T(T(T(T(T(-f[9]) - f[12]) - f[15]) + f[18]) + f[21]) + f[24]);
      T z1 = T(T(T(T(T(T(-f[11]) - f[14]) - f[17]) + f[20]) + f[23]) + f[26]);
      uz = T(z0 + T(two * z1));
    } else {
      T k = T(T(T(T(T(T(T(T(f[1] + f[4]) + f[7]) + f[10]) + f[13]) + f[16]) + f[19]) + f[22]) + f[25]);
      S = T(S0 + T(two * k));
      ux = T(rs - S);
      T y0 = T(T(T(T(T(T(-f[3]) + f[6]) - f[12]) + f[15]) - f[21]) + f[24]);
      T y1 = T(T(T(T(T(T(-f[4]) + f[7]) - f[13]) + f[16]) - f[22]) + f[25]);
      uy = T(y0 + T(two * y1));
      T z0 = T(T(T(T(T(T(-f[9]) - f[12]) - f[15]) + f[18]) + f[21]) + f[24]);
      T z1 = T(T(T(T(T(T(-f[10]) - f[13]) - f[16]) + f[19]) + f[22]) + f[25]);
      uz = T(z0 + T(two * z1));
    }
    if (normal_only) uy = uz = T(0);
    const T c49 = T(4.0 / 9.0), c19 = T(1.0 / 9.0), c136 = T(1.0 / 36.0);
    if (at_max) {
      f[1] = T(f[2] - T(c49 * ux));
      f[4] = T(f[5] - T(c19 * T(ux + uy)));
      f[7] = T(f[8] + T(c19 * T(T(-ux) + uy)));
      f[10] = T(f[11] - T(c19 * T(ux + uz)));
      f[13] = T(f[14] - T(c136 * T(T(ux + uy) + uz)));
      f[16] = T(f[17] + T(c136 * T(T(T(-ux) + uy) + uz)));
      f[19] = T(f[20] + T(c19 * T(T(-ux) + uz)));
      f[22] = T(f[23] + T(c136 * T(T(T(-ux) - uy) + uz)));
      f[25] = T(f[26] + T(c136 * T(T(T(-ux) + uy) + uz)));
    } else {
      f[2] = T(f[1] + T(c49 * ux));
      f[5] = T(f[4] + T(c19 * T(ux - uy)));
      f[8] = T(f[7] + T(c19 * T(ux + uy)));
      f[11] = T(f[10] + T(c19 * T(ux - uz)));
      f[14] = T(f[13] + T(c136 * T(T(ux - uy) - uz)));
      f[17] = T(f[16] + T(c136 * T(T(ux + uy) - uz)));
      f[20] = T(f[19] + T(c19 * T(ux + uz)));
      f[23] = T(f[22] + T(c136 * T(T(ux - uy) + uz)));
      f[26] = T(f[25] + T(c136 * T(T(ux + uy) + uz)));
    }
  }
  // ZouHeVelocityX<East>, D2Q9 indices (boundary.hpp:759-805)
  static void zouhe_u_d2q9(T* f, bool east, const T* u) {
    T ds = east ? T(T(f[1] + f[5]) + f[7]) : T(T(f[2] + f[6]) + f[8]);
    T r = T(T(T(T(f[0] + f[4]) + f[3]) + T(ds * T(2))) / T(u[0] + T(1.0)));
    const T c23 = T(2.0 / 3.0), c16 = T(1.0 / 6.0), h = T(0.5f);
    if (east) {
      f[2] = T(f[1] + T(T(c23 * r) * u[0]));
      f[6] = T(T(f[5] + T(T(c16 * r) * u[0])) + T(T(T(f[3] - f[4]) + T(r * u[1])) * h));
      f[8] = T(T(f[7] + T(T(c16 * r) * u[0])) + T(T(T(f[4] - f[3]) + T(r * u[1])) * h));
    } else {
      f[1] = T(f[2] + T(T(c23 * r) * u[0]));
      f[5] = T(T(f[6] + T(T(c16 * r) * u[0])) + T(T(T(f[4] - f[3]) + T(r * u[1])) * h));
      f[7] = T(T(f[8] + T(T(c16 * r) * u[0])) + T(T(T(f[3] - f[4]) + T(r * u[1])) * h));
    }
  }

  static T omega_of(double tau) { return T(T(1) / T(tau)); }

  // One cell of the collide phase: switch(info) of the drivers.
  void collide_cell(uint64_t c, uint64_t t, T* a) {
    const uint8_t fl = info[c];
    if (fl == 0 || !has_op[fl]) return;
    const lbm_op_desc& op = ops[fl];
    uint64_t x[3];
    coords(c, x);
    T* f = a + gcell(x) * Q;
    switch (op.boundary) {
      case LBM_BC_BOUNCE_BACK: {
        T cpy[27];
        for (int i = 0; i < Q; ++i) cpy[i] = f[i];
        for (int i = 0; i < Q; ++i) f[i] = cpy[L::opp[i]];
        break;
      }
      case LBM_BC_ANTI_BOUNCE_BACK: {  // boundary.hpp:123-131, D2Q9 only
        T cpy[27];
        for (int i = 0; i < Q; ++i) cpy[i] = f[i];
        static const int perm[9] = {0, 1, 2, 4, 3, 7, 8, 5, 6};
        for (int i = 0; i < 9; ++i) f[i] = cpy[perm[i]];
        break;
      }
      case LBM_BC_ZOUHE_PRESSURE_X: {
        double rsd = (op.ramp_steps > 0 && t < op.ramp_steps)
                         ? op.rho_start + op.rho_slope * static_cast<double>(t)
                         : op.rho_final;
        T rs = T(rsd);
        if (Q == 9) zouhe_p_d2q9(f, op.east != 0, rs);
        if (Q == 27) zouhe_p_d3q27(f, op.east != 0, rs, (op.flags & LBM_OPF_ZOUHE_NORMAL_ONLY) != 0);
        break;
      }
      case LBM_BC_ZOUHE_VELOCITY_X: {
        T u[3] = {T(0), T(0), T(0)};
        for (int d = 0; d < D; ++d) u[d] = bc_vel[c * D + d];
        zouhe_u_d2q9(f, op.east != 0, u);
        break;
      }
      case LBM_BC_EQUILIBRIUM: {
        T r = T(op.rho_eq), u[3] = {T(op.vec[0]), T(op.vec[1]), T(op.vec[2])};
        equilibrium<T, L>(r, u, f);
        break;
      }
      default:
        break;
    }
    const T omega = omega_of(op.tau);
    switch (op.collision) {
      case LBM_COLL_BGK: {  // collision.hpp:76-96
        T r, u[3], eq[27];
        rho_u<T, L>(f, r, u);
        rho[c] = r;
        for (int d = 0; d < D; ++d) vel[c * D + d] = u[d];
        equilibrium<T, L>(r, u, eq);
        for (int i = 0; i < Q; ++i) f[i] = T(f[i] + T(omega * T(eq[i] - f[i])));
        break;
      }
      case LBM_COLL_BGK_GUO: {  // collision.hpp:135-204
        T F[3];
        for (int d = 0; d < D; ++d) {
          T cf = (op.flags & LBM_OPF_GUO_CELL_FORCE) ? force[c * D + d] : T(0);
          F[d] = T(cf + T(op.vec[d]));
        }
        T r, u[3], eq[27];
        rho_u<T, L>(f, r, u);
        rho[c] = r;
        for (int d = 0; d < D; ++d) vel[c * D + d] = u[d];
        equilibrium<T, L>(r, u, eq);
        const T ics2 = T(3.0);
        const T relax = T(T(1) - T(T(0.5f) * omega));
        for (int i = 0; i < Q; ++i) {
          T ci[3];
          for (int d = 0; d < D; ++d) ci[d] = T(L::c[i][d]);
          T cu = T(0);
          for (int d = 0; d < D; ++d) cu = T(cu + T(ci[d] * u[d]));
          T s2 = T(T(cu * ics2) * ics2);
          T proj = T(0);
          for (int d = 0; d < D; ++d) {
            T t1 = T(T(ci[d] - u[d]) * ics2);
            T t2 = T(ci[d] * s2);
            proj = T(proj + T(T(t1 + t2) * F[d]));
          }
          T Fi = T(T(L::w[i]) * proj);
          f[i] = T(T(f[i] + T(omega * T(eq[i] - f[i]))) + T(Fi * relax));
        }
        break;
      }
      case LBM_COLL_PSM: {  // psm.hpp:58-89 (+ particle_opa, ForceGather, component/psm)
        if (op.flags & LBM_OPF_PSM_STAMP_DISC) {
          T rel[3];
          for (int d = 0; d < D; ++d) rel[d] = T(static_cast<T>(x[d]) - T(op.disc[d]));
          por[c] = porosity<T, D>(rel, T(op.disc[3]), T(op.disc[4]), this->series_sin);
        }
        T r, u[3], eq[27];
        rho_u<T, L>(f, r, u);
        rho[c] = r;
        for (int d = 0; d < D; ++d) vel[c * D + d] = u[d];
        equilibrium<T, L>(r, u, eq);
        const T B = por[c];
        const T flip = T(T(1.0) - B);
        T cpy[27];
        for (int i = 0; i < Q; ++i) cpy[i] = f[i];
        for (int i = 0; i < Q; ++i) {
          const int o = L::opp[i];
          f[i] = T(T(cpy[i] + T(T(omega * T(eq[i] - cpy[i])) * B)) + T(T(cpy[o] - cpy[i]) * flip));
        }
        if (op.flags & (LBM_OPF_PSM_CELL_FORCE | LBM_OPF_PSM_FORCE_GATHER)) {
          T fc[3] = {T(0), T(0), T(0)};
          for (int i = 0; i < Q; ++i) {
            T diff = T(f[i] - f[L::opp[i]]);
            for (int d = 0; d < D; ++d) fc[d] = T(fc[d] + T(diff * T(L::c[i][d])));
          }
          // component/psm.hpp:75 scales by B; momentum_gather.hpp:47-49 by 1-B
          const T s = (op.flags & LBM_OPF_PSM_CELL_FORCE) ? B : flip;
          for (int d = 0; d < D; ++d) force[c * D + d] = T(fc[d] * s);
        }
        break;
      }
      case LBM_COLL_FPLBM: {  // fplbm.hpp:60-133 (rho, m as FpLbmReset computes them)
        T r = T(0), m[3] = {T(0), T(0), T(0)}, u[3] = {T(0), T(0), T(0)}, eq[27];
        for (int j = 0; j < Q; ++j) {  // compute_rho_momentum (macroscopic.hpp:41-60)
          r = T(r + f[j]);
          for (int d = 0; d < D; ++d) m[d] = T(m[d] + T(T(L::c[j][d]) * f[j]));
        }
        for (int d = 0; d < D; ++d) u[d] = T(m[d] / r);
        rho[c] = r;
        for (int d = 0; d < D; ++d) vel[c * D + d] = m[d];
        equilibrium<T, L>(r, u, eq);
        const T ics2 = T(3.0);
        const T relax = T(T(1) - T(T(0.5f) * omega));
        for (int i = 0; i < Q; ++i) {
          T ci[3];
          for (int d = 0; d < D; ++d) ci[d] = T(L::c[i][d]);
          T cu = T(0);
          for (int d = 0; d < D; ++d) cu = T(cu + T(ci[d] * u[d]));
          T s2 = T(T(cu * ics2) * ics2);
          T proj = T(0);
          for (int d = 0; d < D; ++d) {
            T t1 = T(T(ci[d] - u[d]) * ics2);
            T t2 = T(ci[d] * s2);
            proj = T(proj + T(T(t1 + t2) * force[c * D + d]));
          }
          T Fi = T(T(L::w[i]) * proj);
          f[i] = T(T(f[i] + T(omega * T(eq[i] - f[i]))) + T(Fi * relax));
        }
        break;
      }
      case LBM_COLL_HLBM: {  // hlbm.hpp:62-105
        T r, u[3], eq[27];
        rho_u<T, L>(f, r, u);
        const T B = por[c];
        const T flip = T(T(1.0) - B);
        const T Dw = pD[c];
        for (int d = 0; d < D; ++d) {
          const T Nd = pN[c * D + d];
          const T term = (Dw > 0.0) ? T(T(Nd * flip) / Dw) : Nd;
          u[d] = T(T(u[d] * B) + term);
        }
        rho[c] = r;
        for (int d = 0; d < D; ++d) vel[c * D + d] = u[d];
        equilibrium<T, L>(r, u, eq);
        for (int i = 0; i < Q; ++i) f[i] = T(f[i] + T(omega * T(eq[i] - f[i])));
        break;
      }
      default:
        break;
    }
  }

  // Stream (stream.hpp:19-58): pull from x - c_i for every flag > 0 cell.
  void stream_cell(uint64_t c, const T* a, T* b) {
    if (info[c] == 0) return;
    uint64_t x[3], nb[3];
    coords(c, x);
    T* dst = b + gcell(x) * Q;
    for (int i = 0; i < Q; ++i) {
      if (neighbour(x, L::c[i], -1, nb))
        dst[i] = a[gcell(nb) * Q + i];
      else
        dst[i] = T(0);
    }
  }

  // --- particle coupling ----------------------------------------------------
  // PsmOneParticleGather (psm.hpp:101-445) for particle p on populations `a`.
  // Writes porosity and force fields in serial particle order (the last
  // writer wins, as a serial loop over particles would); returns the FP64
  // sums used by tests next to the reference's T sum.
  void gather(uint32_t p, const T* a, bool write_porosity) {
    Part<T>& P = parts[p];
    const T one = T(1.0);
    T ts = T(one / static_cast<T>(pp.sub_steps));
    T vp[3];
    for (int d = 0; d < D; ++d) vp[d] = T(T(P.pos[d] - P.pos_old[d]) / ts);
    uint64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    aabb<T, D>(P.pos, P.radius, n, lo, hi);
    T fp[3] = {T(0), T(0), T(0)};
    double f64[3] = {0, 0, 0}, t64[3] = {0, 0, 0}, abs64 = 0;
    const T eps = T(pp.eps);
    uint64_t x[3] = {0, 0, 0};
    // iterator<D>::apply: axis 0 outermost (iterator.hpp:24-50)
    for (x[0] = lo[0]; x[0] < hi[0]; ++x[0])
      for (x[1] = lo[1]; x[1] < hi[1]; ++x[1])
        for (x[2] = (D == 3 ? lo[2] : 0); x[2] < (D == 3 ? hi[2] : 1); ++x[2]) {
          const uint64_t c = cell(x);
          T rel[3];
          for (int d = 0; d < D; ++d) rel[d] = T(static_cast<T>(x[d]) - P.pos[d]);
          T B = porosity<T, D>(rel, P.radius, eps, this->series_sin);
          if (write_porosity) por[c] = B;
          if (B >= 1.0) continue;
          const T* f = a + gcell(x) * Q;
          T mom[3] = {T(0.0), T(0.0), T(0.0)};
          for (int i = 1; i < Q; ++i) {
            const int o = L::opp[i];
            T e[3];
            for (int d = 0; d < D; ++d) e[d] = T(L::c[i][d]);
            T upe = T(0);
            for (int d = 0; d < D; ++d) upe = T(upe + T(e[d] * vp[d]));
            uint64_t nb[3];
            T fo = neighbour(x, L::c[i], +1, nb) ? a[gcell(nb) * Q + o] : T(0);
            T added = T(T(f[i] * T(T(1.0) - upe)) + T(fo * T(T(1.0) + upe)));
            for (int d = 0; d < D; ++d) mom[d] = T(mom[d] + T(e[d] * added));
          }
          T flip = T(one - B);
          T fc[3];
          for (int d = 0; d < D; ++d) {
            fc[d] = T(mom[d] * flip);
            force[c * D + d] = fc[d];
            fp[d] = T(fp[d] - fc[d]);
            f64[d] -= static_cast<double>(fc[d]);
            abs64 += std::fabs(static_cast<double>(fc[d]));
          }
          T tq[3] = {T(0), T(0), T(0)};
          cross_rf<T, D>(rel, fc, tq);
          for (int d = 0; d < (D == 2 ? 1 : 3); ++d) t64[d] -= static_cast<double>(tq[d]);
        }
    for (int d = 0; d < D; ++d) P.force[d] = this->sum64 ? T(P.force[d] + T(f64[d])) : T(P.force[d] + fp[d]);
    if (pp.psm_torque) {
      for (int d = 0; d < (D == 2 ? 1 : 3); ++d) P.torque[d] = T(P.torque[d] + T(t64[d]));
    }
    for (int d = 0; d < 3; ++d) {
      F64[p * 3 + d] = f64[d];
      T64[p * 3 + d] = t64[d];
    }
    ABS[p] = abs64;
  }

  // HlbmOneParticleGather (hlbm.hpp:111-256) for particle p: porosity
  // assigned over the AABB, the raw momentum exchange as the cell force (no
  // solid-fraction factor), force and torque sums subtracted from the body.
  void gather_hlbm(uint32_t p, const T* a) {
    Part<T>& P = parts[p];
    const T one = T(1.0);
    uint64_t steps = pp.sub_steps == 0 ? 1 : pp.sub_steps;
    T ts = T(one / static_cast<T>(steps));
    T vp[3];
    for (int d = 0; d < D; ++d) vp[d] = T(T(P.pos[d] - P.pos_old[d]) / ts);
    uint64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    aabb<T, D>(P.pos, P.radius, n, lo, hi);
    T fp[3] = {T(0), T(0), T(0)}, tp[3] = {T(0), T(0), T(0)};
    double f64[3] = {0, 0, 0}, t64[3] = {0, 0, 0}, abs64 = 0;
    const T eps = T(1.5f);  // hlbm.hpp:173
    uint64_t x[3] = {0, 0, 0};
    for (x[0] = lo[0]; x[0] < hi[0]; ++x[0])
      for (x[1] = lo[1]; x[1] < hi[1]; ++x[1])
        for (x[2] = (D == 3 ? lo[2] : 0); x[2] < (D == 3 ? hi[2] : 1); ++x[2]) {
          const uint64_t c = cell(x);
          T rel[3];
          for (int d = 0; d < D; ++d) rel[d] = T(static_cast<T>(x[d]) - P.pos[d]);
          const T B = porosity<T, D>(rel, P.radius, eps, this->series_sin);
          por[c] = B;
          if (B >= 1.0f) continue;
          const T* f = a + gcell(x) * Q;
          T mom[3] = {T(0.0), T(0.0), T(0.0)};
          for (int i = 1; i < Q; ++i) {  // i = 0 adds e_0 = 0
            const int o = L::opp[i];
            T e[3];
            for (int d = 0; d < D; ++d) e[d] = T(L::c[i][d]);
            T upe = T(0);
            for (int d = 0; d < D; ++d) upe = T(upe + T(e[d] * vp[d]));
            uint64_t nb[3];
            T fo = neighbour(x, L::c[i], +1, nb) ? a[gcell(nb) * Q + o] : T(0);
            T added = T(T(f[i] * T(T(1) - upe)) + T(fo * T(T(1) + upe)));
            for (int d = 0; d < D; ++d) mom[d] = T(mom[d] + T(e[d] * added));
          }
          for (int d = 0; d < D; ++d) {
            force[c * D + d] = mom[d];
            fp[d] = T(fp[d] + mom[d]);
            f64[d] -= static_cast<double>(mom[d]);
            abs64 += std::fabs(static_cast<double>(mom[d]));
          }
          T tq[3] = {T(0), T(0), T(0)};
          cross_rf<T, D>(rel, mom, tq);
          for (int d = 0; d < (D == 2 ? 1 : 3); ++d) {
            tp[d] = T(tp[d] + tq[d]);
            t64[d] -= static_cast<double>(tq[d]);
          }
        }
    for (int d = 0; d < D; ++d) P.force[d] = this->sum64 ? T(P.force[d] + T(f64[d])) : T(P.force[d] - fp[d]);
    for (int d = 0; d < (D == 2 ? 1 : 3); ++d)
      P.torque[d] = this->sum64 ? T(P.torque[d] + T(t64[d])) : T(P.torque[d] - tp[d]);
    for (int d = 0; d < 3; ++d) {
      F64[p * 3 + d] = f64[d];
      T64[p * 3 + d] = t64[d];
    }
    ABS[p] = abs64;
  }

  // FpLbmOneParticle (fplbm.hpp:135-247) for particle p (its own AABB and
  // radius): porosity assigned, cell force (u_s - m)(1 - B) with m the
  // momentum of a flag-2 cell (FpLbmReset) or the stored macro velocity of
  // any other cell, body force the negative sum.  Integration follows.
  void gather_fplbm(uint32_t p, const T* a) {
    Part<T>& P = parts[p];
    const T one = T(1.0f), k = T(1.0f), eps = T(1.5f);
    const T sss = static_cast<T>(pp.sub_steps);
    T vs[3];
    for (int d = 0; d < D; ++d) vs[d] = T(T(P.pos[d] - P.pos_old[d]) * sss);
    uint64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    aabb<T, D>(P.pos, P.radius, n, lo, hi);
    T fp[3] = {T(0), T(0), T(0)};
    double f64[3] = {0, 0, 0}, abs64 = 0;
    uint64_t x[3] = {0, 0, 0};
    for (x[0] = lo[0]; x[0] < hi[0]; ++x[0])
      for (x[1] = lo[1]; x[1] < hi[1]; ++x[1])
        for (x[2] = (D == 3 ? lo[2] : 0); x[2] < (D == 3 ? hi[2] : 1); ++x[2]) {
          const uint64_t c = cell(x);
          T rel[3];
          for (int d = 0; d < D; ++d) rel[d] = T(static_cast<T>(x[d]) - P.pos[d]);
          const T B = porosity<T, D>(rel, P.radius, eps, this->series_sin);
          por[c] = B;
          const T flip = T(one - B);
          T mom[3] = {T(0), T(0), T(0)};
          if (info[c] == 2) {
            const T* f = a + gcell(x) * Q;
            for (int j = 0; j < Q; ++j)
              for (int d = 0; d < D; ++d) mom[d] = T(mom[d] + T(T(L::c[j][d]) * f[j]));
          } else {
            for (int d = 0; d < D; ++d) mom[d] = vel[c * D + d];
          }
          for (int d = 0; d < D; ++d) {
            const T fc = T(T(T(vs[d] - mom[d]) * k) * flip);
            force[c * D + d] = fc;
            fp[d] = T(fp[d] - fc);
            f64[d] -= static_cast<double>(fc);
            abs64 += std::fabs(static_cast<double>(fc));
          }
        }
    for (int d = 0; d < D; ++d) P.force[d] = this->sum64 ? T(f64[d]) : fp[d];
    for (int d = 0; d < 3; ++d) {
      F64[p * 3 + d] = f64[d];
      T64[p * 3 + d] = 0.0;
    }
    ABS[p] = abs64;
  }

  // HlbmOneParticleStamp (hlbm.hpp:264-308): after the move, every cell of
  // the post-move AABB gains D += 1 - porosity and N += u_p (1 - porosity).
  void stamp_hlbm() {
    const T one = T(1.0);
    uint64_t steps = pp.sub_steps == 0 ? 1 : pp.sub_steps;
    const T ts = T(one / static_cast<T>(steps));
    for (uint32_t p = 0; p < parts.size(); ++p) {
      const Part<T>& P = parts[p];
      T vp[3];
      for (int d = 0; d < D; ++d) vp[d] = T(T(P.pos[d] - P.pos_old[d]) / ts);
      uint64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      aabb<T, D>(P.pos, P.radius, n, lo, hi);
      uint64_t x[3] = {0, 0, 0};
      for (x[0] = lo[0]; x[0] < hi[0]; ++x[0])
        for (x[1] = lo[1]; x[1] < hi[1]; ++x[1])
          for (x[2] = (D == 3 ? lo[2] : 0); x[2] < (D == 3 ? hi[2] : 1); ++x[2]) {
            const uint64_t c = cell(x);
            const T flip = T(one - por[c]);
            pD[c] = T(pD[c] + flip);
            for (int d = 0; d < D; ++d) pN[c * D + d] = T(pN[c * D + d] + T(vp[d] * flip));
          }
    }
  }

  int particles_set(const lbm_particle* in, uint32_t cnt, const lbm_particle_params* prm) override {
    parts.resize(cnt);
    pp = *prm;
    for (uint32_t p = 0; p < cnt; ++p) {
      Part<T>& q = parts[p];
      for (int d = 0; d < 3; ++d) {
        q.pos[d] = T(in[p].position[d]);
        q.pos_old[d] = T(in[p].position_old[d]);
        q.rot[d] = T(in[p].rotation[d]);
        q.rot_old[d] = T(in[p].rotation_old[d]);
        q.acc[d] = T(in[p].acceleration[d]);
        q.ang_acc[d] = T(in[p].angular_acceleration[d]);
        q.force[d] = T(in[p].force[d]);
        q.torque[d] = T(in[p].torque[d]);
      }
      q.mass = T(in[p].mass);
      q.inertia = T(in[p].inertia);
      q.density = T(in[p].density);
      q.radius = T(in[p].radius);
    }
    F64.assign(cnt * 3, 0);
    T64.assign(cnt * 3, 0);
    ABS.assign(cnt, 0);
    cl.make(n, T(pp.max_radius), pp.cell_capacity > 0 ? pp.cell_capacity : 4);
    return 0;
  }
  void particles_get(lbm_particle* out, uint32_t cnt) const override {
    for (uint32_t p = 0; p < cnt && p < parts.size(); ++p) {
      const Part<T>& q = parts[p];
      std::memset(&out[p], 0, sizeof(lbm_particle));
      for (int d = 0; d < 3; ++d) {
        out[p].position[d] = q.pos[d];
        out[p].position_old[d] = q.pos_old[d];
        out[p].rotation[d] = q.rot[d];
        out[p].rotation_old[d] = q.rot_old[d];
        out[p].acceleration[d] = q.acc[d];
        out[p].angular_acceleration[d] = q.ang_acc[d];
        out[p].force[d] = q.force[d];
        out[p].torque[d] = q.torque[d];
      }
      out[p].mass = q.mass;
      out[p].inertia = q.inertia;
      out[p].density = q.density;
      out[p].radius = q.radius;
    }
  }
  // PsmReset over the lattice, ParticleForceReset + gather for every particle
  // in index order, then (product mode) the cell_porosity composition.
  void couple(uint64_t t) {
    const T* a = A(t);
    if (pp.coupling == LBM_COUPLING_FPLBM) {
      // FpLbmReset (fplbm.hpp:21-58): porosity 1, force 0; the momentum of
      // flag-2 cells is taken from the populations where the gather needs it
      std::fill(por.begin(), por.end(), T(1.0));
      std::fill(force.begin(), force.end(), T(0.0));
      for (auto& q : parts)
        for (int d = 0; d < 3; ++d) {
          q.force[d] = T(0);
          q.torque[d] = T(0);
        }
      for (uint32_t p = 0; p < parts.size(); ++p) gather_fplbm(p, a);
      return;
    }
    if (pp.coupling == LBM_COUPLING_HLBM) {
      // HlbmReset (hlbm.hpp:28-49), ParticleForceReset, gathers in index order
      std::fill(por.begin(), por.end(), T(1.0));
      std::fill(pN.begin(), pN.end(), T(0.0));
      std::fill(pD.begin(), pD.end(), T(0.0));
      for (auto& q : parts)
        for (int d = 0; d < 3; ++d) {
          q.force[d] = T(0);
          q.torque[d] = T(0);
        }
      for (uint32_t p = 0; p < parts.size(); ++p) gather_hlbm(p, a);
      return;
    }
    std::fill(por.begin(), por.end(), T(1.0));  // PsmReset (psm.hpp:20-32)
    const bool product = pp.porosity_mode == LBM_POROSITY_PRODUCT;
    for (uint32_t p = 0; p < parts.size(); ++p) {  // reset_particle_accumulators
      for (int d = 0; d < 3; ++d) {
        parts[p].force[d] = T(0);
        parts[p].torque[d] = T(0);
      }
    }
    for (uint32_t p = 0; p < parts.size(); ++p) gather(p, a, !product);
    if (product) {
      cl.build(parts);
      const T eps = T(pp.eps);
      for (uint64_t c = 0; c < N; ++c) {
        uint64_t x[3];
        coords(c, x);
        T cp[3];
        for (int d = 0; d < D; ++d) cp[d] = static_cast<T>(x[d]);
        T B = T(1);
        cl.for_each_candidate(x, [&](uint64_t p) {
          T rel[3];
          for (int d = 0; d < D; ++d) rel[d] = T(cp[d] - parts[p].pos[d]);
          B = T(B * porosity<T, D>(rel, parts[p].radius, eps, this->series_sin));
        });
        por[c] = B;
      }
    }
  }
  void psm_couple(uint64_t t, double* f64, double* t64, double* abs) override {
    couple(t);
    if (f64) std::copy(F64.begin(), F64.end(), f64);
    if (t64) std::copy(T64.begin(), T64.end(), t64);
    if (abs) std::copy(ABS.begin(), ABS.end(), abs);
  }
  void dem_contact() override {
    const std::vector<Part<T>> snap = parts;  // every particle sees pre-contact state
    Dem dem{pp.k_n, pp.zeta_n, pp.gamma_t, pp.mu};
    for (uint32_t p = 0; p < parts.size(); ++p) {
      particle_contacts<T, D>(snap, p, parts[p], dem, pp.sub_steps);
      if (pp.walls) wall_contacts<T, D>(snap[p], parts[p], n, dem, pp.sub_steps);
    }
  }
  void integrate_all() override {
    // FpLbmOneParticle integrates only when sub_steps > 0 (fplbm.hpp:238-244)
    if (pp.coupling == LBM_COUPLING_FPLBM && pp.sub_steps == 0) return;
    if (pp.fixed_axis) {
      T ax[3];
      spin_axis<T, D>(pp.spin_axis, ax);
      for (auto& p : parts) integrate_fixed_axis<T, D>(p, ax, pp.sub_steps);
      return;
    }
    T g[3] = {T(pp.body_acceleration[0]), T(pp.body_acceleration[1]), T(pp.body_acceleration[2])};
    const bool fp = pp.coupling == LBM_COUPLING_FPLBM;
    for (auto& p : parts) integrate<T, D>(p, g, pp.sub_steps, fp);
  }
  uint64_t overflow() const override { return cl.overflow; }
  void get_hlbm(double* nsum, double* dsum) const override {
    for (uint64_t c = 0; c < N; ++c) {
      if (dsum) dsum[c] = pD[c];
      for (int d = 0; d < D; ++d)
        if (nsum) nsum[uint64_t(d) * N + c] = pN[c * D + d];
    }
  }

  void step(uint64_t t0, uint64_t ns, int threads) override {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (uint64_t t = t0; t < t0 + ns; ++t) {
      if (!parts.empty()) {
        couple(t);
        dem_contact();
        integrate_all();
        if (pp.coupling == LBM_COUPLING_HLBM) stamp_hlbm();
      }
      T* a = A(t);
      T* b = B(t);
      const int64_t NN = static_cast<int64_t>(N);
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < NN; ++c) collide_cell(static_cast<uint64_t>(c), t, a);
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < NN; ++c) stream_cell(static_cast<uint64_t>(c), a, b);
    }
  }
};

template <class L>
SimBase* make_sim(int storage, const uint64_t* ext) {
  if (storage == LBM_F64) return new Sim<double, L>(ext);
  if (storage == LBM_F32) return new Sim<float, L>(ext);
  return nullptr;
}

}  // namespace orc

// ================================================================ C API
using namespace orc;

extern "C" {

int orc_lattice(int lat, int* D, int* Q, int* c, double* w, int* opp) {
  auto fill = [&](auto L) {
    using LL = decltype(L);
    *D = LL::D;
    *Q = LL::Q;
    for (int i = 0; i < LL::Q; ++i) {
      for (int d = 0; d < 3; ++d) c[i * 3 + d] = LL::c[i][d];
      w[i] = LL::w[i];
      opp[i] = LL::opp[i];
    }
  };
  if (lat == LBM_D2Q9) fill(D2Q9{});
  else if (lat == LBM_D3Q19) fill(D3Q19{});
  else if (lat == LBM_D3Q27) fill(D3Q27{});
  else return -1;
  return 0;
}

}  // extern "C"
template <class T, class L>
static void eq_t(double rho, const double* u, double* eq) {
  T uu[3] = {T(u[0]), T(u[1]), T(u[2])}, e[27];
  equilibrium<T, L>(T(rho), uu, e);
  for (int i = 0; i < L::Q; ++i) eq[i] = e[i];
}
template <class T, class L>
static void rhou_t(const double* f, double* rho, double* u) {
  T ff[27], r, uu[3];
  for (int i = 0; i < L::Q; ++i) ff[i] = T(f[i]);
  rho_u<T, L>(ff, r, uu);
  *rho = r;
  for (int d = 0; d < L::D; ++d) u[d] = uu[d];
}
#define ORC_DISPATCH(lat, st, FN, ...)                                   \
  do {                                                                   \
    if ((lat) == LBM_D2Q9 && (st) == LBM_F64) FN<double, D2Q9>(__VA_ARGS__); \
    if ((lat) == LBM_D2Q9 && (st) == LBM_F32) FN<float, D2Q9>(__VA_ARGS__);  \
    if ((lat) == LBM_D3Q19 && (st) == LBM_F64) FN<double, D3Q19>(__VA_ARGS__); \
    if ((lat) == LBM_D3Q19 && (st) == LBM_F32) FN<float, D3Q19>(__VA_ARGS__);  \
    if ((lat) == LBM_D3Q27 && (st) == LBM_F64) FN<double, D3Q27>(__VA_ARGS__); \
    if ((lat) == LBM_D3Q27 && (st) == LBM_F32) FN<float, D3Q27>(__VA_ARGS__);  \
  } while (0)

extern "C" {
void orc_equilibrium(int lat, int st, double rho, const double* u, double* eq) {
  ORC_DISPATCH(lat, st, eq_t, rho, u, eq);
}
void orc_rho_u(int lat, int st, const double* f, double* rho, double* u) {
  ORC_DISPATCH(lat, st, rhou_t, f, rho, u);
}

}  // extern "C"
// Single-cell operators for closed-form tests: op applied to f in place,
// boundary then collision, exactly as the collide phase does it.
template <class T, class L>
static void cell_op_t(const lbm_op_desc* op, uint64_t t, double* f, const double* cell_force,
                      const double* bc_u, double porosity_in, double* rho_out, double* u_out,
                      double* force_out) {
  uint64_t ext[3] = {1, 1, 1};
  Sim<T, L> s(ext);
  uint8_t fl = 2;
  s.set_flags(&fl);
  s.set_operator(2, op);
  T* a = s.A(t);
  uint64_t x[3] = {0, 0, 0};
  T* ff = a + s.gcell(x) * L::Q;
  for (int i = 0; i < L::Q; ++i) ff[i] = T(f[i]);
  for (int d = 0; d < L::D; ++d) {
    s.force[d] = cell_force ? T(cell_force[d]) : T(0);
    s.bc_vel[d] = bc_u ? T(bc_u[d]) : T(0);
  }
  s.por[0] = T(porosity_in);
  s.collide_cell(0, t, a);
  for (int i = 0; i < L::Q; ++i) f[i] = ff[i];
  if (rho_out) *rho_out = s.rho[0];
  for (int d = 0; d < L::D; ++d) {
    if (u_out) u_out[d] = s.vel[d];
    if (force_out) force_out[d] = s.force[d];
  }
}
template <class T, class L>
static void cell_hlbm_t(const lbm_op_desc* op, double* f, double porosity_in, const double* nsum, double dsum,
                        double* rho_out, double* u_out) {
  uint64_t ext[3] = {1, 1, 1};
  Sim<T, L> s(ext);
  uint8_t fl = 2;
  s.set_flags(&fl);
  s.set_operator(2, op);
  T* a = s.A(0);
  uint64_t x[3] = {0, 0, 0};
  T* ff = a + s.gcell(x) * L::Q;
  for (int i = 0; i < L::Q; ++i) ff[i] = T(f[i]);
  for (int d = 0; d < L::D; ++d) s.pN[d] = T(nsum[d]);
  s.pD[0] = T(dsum);
  s.por[0] = T(porosity_in);
  s.collide_cell(0, 0, a);
  for (int i = 0; i < L::Q; ++i) f[i] = ff[i];
  if (rho_out) *rho_out = s.rho[0];
  for (int d = 0; d < L::D; ++d)
    if (u_out) u_out[d] = s.vel[d];
}
extern "C" {
void orc_cell_hlbm(int lat, int st, const lbm_op_desc* op, double* f, double porosity_in, const double* nsum,
                   double dsum, double* rho_out, double* u_out) {
  ORC_DISPATCH(lat, st, cell_hlbm_t, op, f, porosity_in, nsum, dsum, rho_out, u_out);
}
void orc_cell_op(int lat, int st, const lbm_op_desc* op, uint64_t t, double* f,
                 const double* cell_force, const double* bc_u, double porosity_in,
                 double* rho_out, double* u_out, double* force_out) {
  ORC_DISPATCH(lat, st, cell_op_t, op, t, f, cell_force, bc_u, porosity_in, rho_out, u_out,
               force_out);
}

void orc_porosity(int D, int st, const double* rel, double r, double eps, double* out) {
  if (st == LBM_F64) {
    double rr[3] = {rel[0], rel[1], rel[2]};
    *out = (D == 2) ? porosity<double, 2>(rr, r, eps) : porosity<double, 3>(rr, r, eps);
  } else {
    float rr[3] = {float(rel[0]), float(rel[1]), float(rel[2])};
    *out = (D == 2) ? porosity<float, 2>(rr, float(r), float(eps))
                    : porosity<float, 3>(rr, float(r), float(eps));
  }
}
void orc_aabb(int D, int st, const double* pos, double r, const uint64_t* extent, uint64_t* a,
              uint64_t* b) {
  if (st == LBM_F64) {
    if (D == 2) aabb<double, 2>(pos, r, extent, a, b);
    else aabb<double, 3>(pos, r, extent, a, b);
  } else {
    float p[3] = {float(pos[0]), float(pos[1]), float(pos[2])};
    if (D == 2) aabb<float, 2>(p, float(r), extent, a, b);
    else aabb<float, 3>(p, float(r), extent, a, b);
  }
}
void orc_spheroid_mass(int D, int st, double r, double rho, double* m, double* I) {
  if (st == LBM_F64) {
    double mm, ii;
    if (D == 2) spheroid_mass<double, 2>(r, rho, mm, ii);
    else spheroid_mass<double, 3>(r, rho, mm, ii);
    *m = mm;
    *I = ii;
  } else {
    float mm, ii;
    if (D == 2) spheroid_mass<float, 2>(float(r), float(rho), mm, ii);
    else spheroid_mass<float, 3>(float(r), float(rho), mm, ii);
    *m = mm;
    *I = ii;
  }
}

void* orc_sim_create(int lat, int st, const uint64_t* ext) {
  if (lat == LBM_D2Q9) return make_sim<D2Q9>(st, ext);
  if (lat == LBM_D3Q19) return make_sim<D3Q19>(st, ext);
  if (lat == LBM_D3Q27) return make_sim<D3Q27>(st, ext);
  return nullptr;
}
void orc_sim_destroy(void* s) { delete static_cast<SimBase*>(s); }
int orc_sim_set_flags(void* s, const uint8_t* f) { return static_cast<SimBase*>(s)->set_flags(f); }
void orc_sim_set_periodic(void* s, uint32_t m) { static_cast<SimBase*>(s)->set_periodic(m); }
void orc_sim_set_operator(void* s, uint8_t flag, const lbm_op_desc* op) {
  static_cast<SimBase*>(s)->set_operator(flag, op);
}
void orc_sim_set_boundary_velocity(void* s, const double* u) {
  static_cast<SimBase*>(s)->set_boundary_velocity(u);
}
void orc_sim_set_force_field(void* s, const double* f) { static_cast<SimBase*>(s)->set_force(f); }
void orc_sim_get_force_field(void* s, double* f) { static_cast<SimBase*>(s)->get_force(f); }
void orc_sim_set_porosity(void* s, const double* p) { static_cast<SimBase*>(s)->set_porosity(p); }
void orc_sim_get_porosity(void* s, double* p) { static_cast<SimBase*>(s)->get_porosity(p); }
void orc_sim_set_populations(void* s, uint64_t t, const double* f) {
  static_cast<SimBase*>(s)->set_populations(t, f);
}
void orc_sim_get_populations(void* s, uint64_t t, double* f) {
  static_cast<SimBase*>(s)->get_populations(t, f);
}
void orc_sim_init_equilibrium(void* s, uint64_t t, double rho, const double* u) {
  static_cast<SimBase*>(s)->init_equilibrium(t, rho, u);
}
void orc_sim_get_macros(void* s, double* rho, double* u) {
  static_cast<SimBase*>(s)->get_macros(rho, u);
}
void orc_sim_set_macros(void* s, const double* rho, const double* u) {
  static_cast<SimBase*>(s)->set_macros(rho, u);
}
void orc_sim_step(void* s, uint64_t t0, uint64_t n, int threads) {
  static_cast<SimBase*>(s)->step(t0, n, threads);
}
int orc_particles_set(void* s, const lbm_particle* p, uint32_t n, const lbm_particle_params* prm) {
  return static_cast<SimBase*>(s)->particles_set(p, n, prm);
}
void orc_particles_get(void* s, lbm_particle* p, uint32_t n) {
  static_cast<SimBase*>(s)->particles_get(p, n);
}
void orc_psm_couple(void* s, uint64_t t, double* f64, double* t64, double* abs_sum) {
  static_cast<SimBase*>(s)->psm_couple(t, f64, t64, abs_sum);
}
void orc_dem_contact(void* s) { static_cast<SimBase*>(s)->dem_contact(); }
void orc_particle_integrate(void* s) { static_cast<SimBase*>(s)->integrate_all(); }
uint64_t orc_cell_list_overflow(void* s) { return static_cast<SimBase*>(s)->overflow(); }
void orc_sim_get_hlbm(void* s, double* nsum, double* dsum) { static_cast<SimBase*>(s)->get_hlbm(nsum, dsum); }
void orc_sim_set_sum64(void* s, int on) { static_cast<SimBase*>(s)->sum64 = on != 0; }
void orc_sim_set_series_sin(void* s, int on) { static_cast<SimBase*>(s)->series_sin = on != 0; }
double orc_series_sin(double x) { return series_sin_q1(x); }

// narrow_phase<Spheroid>::contact on a canonical (lo, hi) pair.
int orc_narrow(int D, int st, const double* pos_lo, double r_lo, const double* pos_hi, double r_hi,
               double* depth, double* normal, double* point) {
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    Part<T> lo{}, hi{};
    for (int d = 0; d < 3; ++d) {
      lo.pos[d] = T(pos_lo[d]);
      hi.pos[d] = T(pos_hi[d]);
    }
    lo.radius = T(r_lo);
    hi.radius = T(r_hi);
    Manifold<T, 3> m3{};
    Manifold<T, 2> m2{};
    bool hit = (D == 2) ? narrow<T, 2>(lo, hi, m2) : narrow<T, 3>(lo, hi, m3);
    const T* n = (D == 2) ? m2.n : m3.n;
    const T* pt = (D == 2) ? m2.point : m3.point;
    *depth = (D == 2) ? m2.depth : m3.depth;
    for (int d = 0; d < D; ++d) {
      normal[d] = n[d];
      point[d] = pt[d];
    }
    return hit ? 1 : 0;
  };
  return st == LBM_F64 ? run(0.0) : run(0.0f);
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
