// Per-cell operators of the fused stream+collide kernel.
//
// Each function works on one cell's populations held in registers.  The
// arithmetic keeps the reference's operation order so that FP64 and FP32
// results are bitwise reproducible (every `data<T>` operation rounds to T,
// modules/core/c++/abstract/data.hpp:170-188); the library is built with
// -ffp-contract=off.  Products with a lattice velocity component c ∈ {−1,0,1}
// are written as add/subtract/skip, which is exact in IEEE arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.hpp"
#include "../../include/lbm_mirror_amd/lbm.h"

namespace lbmx {

// Device copy of one flag's operator (lbm_op_desc); T-valued parameters are
// already rounded to T on the host and carried as double.
struct OpDev {
  int32_t bc, coll, east;
  uint32_t flags;
  double omega;  // T(T(1)/T(tau)), collision.hpp:25-38
  double rho_start, rho_slope, rho_final;
  uint64_t ramp_steps;
  double vec[3];
  double rho_eq;
  double disc[5];
  uint32_t force_mask;  // BGKGuo: bit d set when force component d may be non-zero
  uint32_t pad;
};

// compute_rho_u (macroscopic.hpp:10-35)
// The sums start from their first term instead of 0 + term (equal up to
// the sign of an exact zero); D3Q27 keeps the sums from zero (the shorter
// form raises its register pressure and runs the fused step 2% slower).
template <class L, class T>
__device__ __forceinline__ void rho_u(const T (&f)[L::Q], T& rho, T (&u)[3]) {
  if constexpr (L::Q == 27) {
    rho = T(0);
    u[0] = u[1] = u[2] = T(0);
#pragma unroll
    for (int j = 0; j < L::Q; ++j) {
      rho = rho + f[j];
#pragma unroll
      for (int d = 0; d < L::D; ++d) {
        const int c = cdir<L>(j, d);
        if (c == 1) u[d] = u[d] + f[j];
        if (c == -1) u[d] = u[d] - f[j];
      }
    }
#pragma unroll
    for (int d = 0; d < L::D; ++d) u[d] = u[d] / rho;
    return;
  }
  rho = f[0];
  u[0] = u[1] = u[2] = T(0);
  bool seen[3] = {false, false, false};
#pragma unroll
  for (int j = 0; j < L::Q; ++j) {
    if (j > 0) rho = rho + f[j];
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
      const int c = cdir<L>(j, d);
      if (c == 0) continue;
      if (!seen[d]) u[d] = c == 1 ? f[j] : -f[j];
      else if (c == 1) u[d] = u[d] + f[j];
      else u[d] = u[d] - f[j];
      seen[d] = true;
    }
  }
#pragma unroll
  for (int d = 0; d < L::D; ++d) u[d] = u[d] / rho;
}

// equilibrium (equilibrium.hpp:7-48).  WB: the bracket is evaluated in
// double and narrowed once, as the reference does for FP32 (its literals are
// double); for T = double WB changes nothing.
template <class L, class T, bool WB>
__device__ __forceinline__ T eq_dir(int i, T rho, const T (&u)[3], T vv) {
  T vc = T(0);
#pragma unroll
  for (int d = 0; d < L::D; ++d) {
    const int c = cdir<L>(i, d);
    if (c == 1) vc = vc + u[d];
    if (c == -1) vc = vc - u[d];
  }
  const T a = vc * T(3.0);
  const T wr = T(L::w[i]) * rho;
  const T aa = a * a;
  if (WB) {
    const double br = ((1.0 + double(a)) - 1.5 * double(vv)) + double(aa) * 0.5;
    return T(double(wr) * br);
  } else {
    return wr * (((T(1) + a) - T(1.5) * vv) + aa * T(0.5));
  }
}

template <class L, class T>
__device__ __forceinline__ T vel_sq(const T (&u)[3]) {
  T vv = T(0);
#pragma unroll
  for (int d = 0; d < L::D; ++d) vv = vv + u[d] * u[d];
  return vv;
}

template <class L, class T, bool WB>
__device__ __forceinline__ void equilibrium(T rho, const T (&u)[3], T (&eq)[L::Q]) {
  const T vv = vel_sq<L, T>(u);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) eq[i] = eq_dir<L, T, WB>(i, rho, u, vv);
}

// The collisions below visit the directions in opposite pairs (i, opp(i)).
// For a pair, c·u of the opposite is the exact negation (IEEE rounding is
// symmetric), so a, a², w·ρ and their double conversions are computed once
// and both equilibria come from the same terms — each value is bitwise what
// the per-direction evaluation gives.
template <class L>
__device__ __forceinline__ constexpr bool pair_head(int i) {
  return i == 0 || i < L::opp[i];
}

template <class L, class T>
__device__ __forceinline__ T dot_c(int i, const T (&u)[3]) {
  T vc = T(0);
#pragma unroll
  for (int d = 0; d < L::D; ++d) {
    const int c = cdir<L>(i, d);
    if (c == 1) vc = vc + u[d];
    if (c == -1) vc = vc - u[d];
  }
  return vc;
}

// eq_i and eq_opp(i) from one a = 3 c_i·u
template <class L, class T, bool WB>
__device__ __forceinline__ void eq_pair(int i, T rho, T a, T vv, T& ei, T& eo) {
  const T wr = T(L::w[i]) * rho;
  const T aa = a * a;
  if (WB) {
    const double A = double(a), C = 1.5 * double(vv), Y = double(aa) * 0.5, W = double(wr);
    ei = T(W * (((1.0 + A) - C) + Y));
    eo = T(W * (((1.0 - A) - C) + Y));
  } else {
    const T C = T(1.5) * vv, Y = aa * T(0.5);
    ei = wr * (((T(1) + a) - C) + Y);
    eo = wr * (((T(1) - a) - C) + Y);
  }
}

// BGK relaxation (collision.hpp:91-95)
template <class L, class T, bool WB>
__device__ __forceinline__ void bgk(T (&f)[L::Q], T rho, const T (&u)[3], T om) {
  const T vv = vel_sq<L, T>(u);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    if (!pair_head<L>(i)) continue;
    const int o = L::opp[i];
    const T a = dot_c<L, T>(i, u) * T(3.0);
    T ei, eo;
    eq_pair<L, T, WB>(i, rho, a, vv, ei, eo);
    f[i] = f[i] + om * (ei - f[i]);
    if (o != i) f[o] = f[o] + om * (eo - f[o]);
  }
}

// BGK with Guo source (collision.hpp:280-315), F = cell force + global force.
// Components outside `fmask` are exactly zero, and (t1+t2)·0 adds nothing to
// the projection, so they are skipped.  Per direction the projection term of
// component d is ((c_d − u_d)·3 + c_d·s2)·F_d with s2 = (c·u·3)·3; c_d is
// +1, −1 or 0, so (c_d − u_d)·3 takes one of three values per cell (formed
// once), c_d·s2 is s2, −s2 or a zero, and the pair's opposite uses the
// other sign.  Each term is bitwise the reference's, up to the sign of an
// exact zero (x + (±0) = x).  D3Q27 evaluates each term directly (the
// precomputed form runs its fused step 2% slower there).
template <class L, class T, bool WB>
__device__ __forceinline__ void bgk_guo(T (&f)[L::Q], T rho, const T (&u)[3], T om,
                                        const T (&F)[3], uint32_t fmask) {
  const T vv = vel_sq<L, T>(u);
  const T relax = T(1) - T(0.5f) * om;
  const T ics2 = T(3.0);
  if constexpr (L::Q == 27) {
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
      if (!pair_head<L>(i)) continue;
      const int o = L::opp[i];
      const T a = dot_c<L, T>(i, u) * ics2;
      T ei, eo;
      eq_pair<L, T, WB>(i, rho, a, vv, ei, eo);
      const T s2 = a * ics2;
      T pi = T(0), po = T(0);
#pragma unroll
      for (int d = 0; d < L::D; ++d) {
        if (!(fmask & (1u << d))) continue;
        const T cd = T(cdir<L>(i, d));
        const T t2 = cd * s2;  // the opposite's c_d·s2 is the same product
        pi = pi + ((cd - u[d]) * ics2 + t2) * F[d];
        po = po + (((-cd) - u[d]) * ics2 + t2) * F[d];
      }
      const T wi = T(L::w[i]);
      f[i] = (f[i] + om * (ei - f[i])) + (wi * pi) * relax;
      if (o != i) f[o] = (f[o] + om * (eo - f[o])) + (wi * po) * relax;
    }
    return;
  }
  T up[3], um[3], u0[3];  // (1 − u_d)·3, (−1 − u_d)·3, (0 − u_d)·3
#pragma unroll
  for (int d = 0; d < L::D; ++d) {
    up[d] = (T(1) - u[d]) * ics2;
    um[d] = (T(-1) - u[d]) * ics2;
    u0[d] = (T(0) - u[d]) * ics2;
  }
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    if (!pair_head<L>(i)) continue;
    const int o = L::opp[i];
    const T cu = dot_c<L, T>(i, u);
    const T a = cu * ics2;
    T ei, eo;
    eq_pair<L, T, WB>(i, rho, a, vv, ei, eo);
    const T s2 = a * ics2;  // (c·u·3)·3; the opposite's is −s2
    T pi = T(0), po = T(0);
    bool first = true;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
      if (!(fmask & (1u << d))) continue;
      const int cd = cdir<L>(i, d);
      T ti, to;
      if (cd == 1) {
        ti = (up[d] + s2) * F[d];
        to = (um[d] + s2) * F[d];
      } else if (cd == -1) {
        ti = (um[d] - s2) * F[d];
        to = (up[d] - s2) * F[d];
      } else {
        ti = to = u0[d] * F[d];
      }
      pi = first ? ti : pi + ti;
      po = first ? to : po + to;
      first = false;
    }
    const T wi = T(L::w[i]);
    f[i] = (f[i] + om * (ei - f[i])) + (wi * pi) * relax;
    if (o != i) f[o] = (f[o] + om * (eo - f[o])) + (wi * po) * relax;
  }
}

// PSM solid-fraction collision (psm.hpp:85-88)
template <class L, class T, bool WB>
__device__ __forceinline__ void psm(T (&f)[L::Q], T rho, const T (&u)[3], T om, T B) {
  const T vv = vel_sq<L, T>(u);
  const T flip = T(1.0) - B;
  T cpy[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) cpy[i] = f[i];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    if (!pair_head<L>(i)) continue;
    const int o = L::opp[i];
    T ei, eo;
    eq_pair<L, T, WB>(i, rho, dot_c<L, T>(i, u) * T(3.0), vv, ei, eo);
    f[i] = (cpy[i] + (om * (ei - cpy[i])) * B) + (cpy[o] - cpy[i]) * flip;
    if (o != i) f[o] = (cpy[o] + (om * (eo - cpy[o])) * B) + (cpy[i] - cpy[o]) * flip;
  }
}

// Σ_i (f_i − f_opp(i)) c_i (momentum_gather.hpp:35-44, component/psm.hpp:64-73)
template <class L, class T>
__device__ __forceinline__ void opp_momentum(const T (&f)[L::Q], T (&m)[3]) {
  m[0] = m[1] = m[2] = T(0);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const T diff = f[i] - f[L::opp[i]];
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
      const int c = cdir<L>(i, d);
      if (c == 1) m[d] = m[d] + diff;
      if (c == -1) m[d] = m[d] - diff;
    }
  }
}

// ZouHePressureX<Dir>, D2Q9 (boundary.hpp:186-226)
template <class T>
__device__ __forceinline__ void zouhe_p_d2q9(T* f, bool dir, T rs) {
  T rvx;
  if (dir)
    rvx = rs - (((f[0] + f[3]) + f[4]) + ((f[1] + f[5]) + f[7]) * T(2u));
  else
    rvx = (((f[0] + f[3]) + f[4]) + ((f[2] + f[6]) + f[8]) * T(2u)) - rs;
  const T c23 = T(2.0 / 3.0), c16 = T(1.0 / 6.0), h = T(0.5);
  // both faces' values, then a select per population: the compiler would
  // otherwise merge the two branches' stores into one through a selected
  // address and move the population array to scratch memory
  const T e2 = f[1] + c23 * rvx;
  const T e6 = (f[5] + c16 * rvx) + h * (f[3] - f[4]);
  const T e8 = (f[7] + c16 * rvx) + h * (f[4] - f[3]);
  const T w1 = f[2] - c23 * rvx;
  const T w5 = (f[6] - c16 * rvx) + h * (f[4] - f[3]);
  const T w7 = (f[8] - c16 * rvx) + h * (f[3] - f[4]);
  f[1] = dir ? f[1] : w1;
  f[2] = dir ? e2 : f[2];
  f[5] = dir ? f[5] : w5;
  f[6] = dir ? e6 : f[6];
  f[7] = dir ? f[7] : w7;
  f[8] = dir ? e8 : f[8];
}

// ZouHePressureX<East>, D3Q27 (boundary.hpp:256-754); the body is written for
// the x = max face and the flag is flipped once (AtXMax = !East, :248).
// normal_only (LBM_OPF_ZOUHE_NORMAL_ONLY, no reference counterpart) drops the
// tangential terms of the reconstruction.
template <class T>
__device__ __forceinline__ void zouhe_p_d3q27(T* f, bool east, T rs, bool normal_only = false) {
  // Both faces' sums and reconstructions are formed and the face's values are
  // selected: assignments to face-dependent elements would become stores
  // through a selected address, which keeps the population array in scratch
  // memory in every kernel that contains this operator.  Each face reads only
  // the populations it does not write, so the values are those of the
  // branch-per-face form, operation for operation.
  const bool at_max = !east;
  const T two = T(2.0);
  const T S0 = ((((((((f[0] + f[3]) + f[6]) + f[9]) + f[12]) + f[15]) + f[18]) + f[21]) + f[24]);
  const T y0 = (((((-f[3]) + f[6]) - f[12]) + f[15]) - f[21]) + f[24];
  const T z0 = (((((-f[9]) - f[12]) - f[15]) + f[18]) + f[21]) + f[24];
  const T kmax = ((((((((f[2] + f[5]) + f[8]) + f[11]) + f[14]) + f[17]) + f[20]) + f[23]) + f[26]);
  const T kmin = ((((((((f[1] + f[4]) + f[7]) + f[10]) + f[13]) + f[16]) + f[19]) + f[22]) + f[25]);
  const T ux = at_max ? (S0 + two * kmax) - rs : rs - (S0 + two * kmin);
  T uy = at_max ? y0 + two * (((((-f[5]) + f[8]) - f[14]) + f[17]) - f[23] + f[26])
                : y0 + two * (((((-f[4]) + f[7]) - f[13]) + f[16]) - f[22] + f[25]);
  T uz = at_max ? z0 + two * (((((-f[11]) - f[14]) - f[17]) + f[20]) + f[23] + f[26])
                : z0 + two * (((((-f[10]) - f[13]) - f[16]) + f[19]) + f[22] + f[25]);
  if (normal_only) uy = uz = T(0);
  const T c49 = T(4.0 / 9.0), c19 = T(1.0 / 9.0), c136 = T(1.0 / 36.0);
  // x = max face: the populations entering along -x (odd indices)
  const T m1 = f[2] - c49 * ux;
  const T m4 = f[5] - c19 * (ux + uy);
  const T m7 = f[8] + c19 * ((-ux) + uy);
  const T m10 = f[11] - c19 * (ux + uz);
  const T m13 = f[14] - c136 * ((ux + uy) + uz);
  const T m16 = f[17] + c136 * (((-ux) + uy) + uz);
  const T m19 = f[20] + c19 * ((-ux) + uz);
  const T m22 = f[23] + c136 * (((-ux) - uy) + uz);
  const T m25 = f[26] + c136 * (((-ux) + uy) + uz);
  // x = 0 face: the populations entering along +x (even indices)
  const T p2 = f[1] + c49 * ux;
  const T p5 = f[4] + c19 * (ux - uy);
  const T p8 = f[7] + c19 * (ux + uy);
  const T p11 = f[10] + c19 * (ux - uz);
  const T p14 = f[13] + c136 * ((ux - uy) - uz);
  const T p17 = f[16] + c136 * ((ux + uy) - uz);
  const T p20 = f[19] + c19 * (ux + uz);
  const T p23 = f[22] + c136 * ((ux - uy) + uz);
  const T p26 = f[25] + c136 * ((ux + uy) + uz);
  f[1] = at_max ? m1 : f[1];
  f[4] = at_max ? m4 : f[4];
  f[7] = at_max ? m7 : f[7];
  f[10] = at_max ? m10 : f[10];
  f[13] = at_max ? m13 : f[13];
  f[16] = at_max ? m16 : f[16];
  f[19] = at_max ? m19 : f[19];
  f[22] = at_max ? m22 : f[22];
  f[25] = at_max ? m25 : f[25];
  f[2] = at_max ? f[2] : p2;
  f[5] = at_max ? f[5] : p5;
  f[8] = at_max ? f[8] : p8;
  f[11] = at_max ? f[11] : p11;
  f[14] = at_max ? f[14] : p14;
  f[17] = at_max ? f[17] : p17;
  f[20] = at_max ? f[20] : p20;
  f[23] = at_max ? f[23] : p23;
  f[26] = at_max ? f[26] : p26;
}

// ZouHeVelocityX<East>, D2Q9 index-hard-coded (boundary.hpp:769-803)
template <class T>
__device__ __forceinline__ void zouhe_u_d2q9(T* f, bool east, T u0, T u1) {
  const T ds = east ? ((f[1] + f[5]) + f[7]) : ((f[2] + f[6]) + f[8]);
  const T r = (((f[0] + f[4]) + f[3]) + ds * T(2)) / (u0 + T(1.0));
  const T c23 = T(2.0 / 3.0), c16 = T(1.0 / 6.0), h = T(0.5f);
  // both faces' values, then a select per population (see zouhe_p_d2q9)
  const T e2 = f[1] + (c23 * r) * u0;
  const T e6 = (f[5] + (c16 * r) * u0) + (((f[3] - f[4]) + r * u1) * h);
  const T e8 = (f[7] + (c16 * r) * u0) + (((f[4] - f[3]) + r * u1) * h);
  const T w1 = f[2] + (c23 * r) * u0;
  const T w5 = (f[6] + (c16 * r) * u0) + (((f[4] - f[3]) + r * u1) * h);
  const T w7 = (f[8] + (c16 * r) * u0) + (((f[3] - f[4]) + r * u1) * h);
  f[1] = east ? f[1] : w1;
  f[2] = east ? e2 : f[2];
  f[5] = east ? f[5] : w5;
  f[6] = east ? e6 : f[6];
  f[7] = east ? f[7] : w7;
  f[8] = east ? e8 : f[8];
}

// particle_porosity<..., ParticleSpheroid>::calculate (porosity.hpp:56-89).
// sin on the porosity ramp's range [0, π/2] in double, by series with fixed
// operations (no library call): x + x³ P(x²) below π/4, and cos(π/2 − x) =
// 1 + z Q(z), z = (π/2 − x)², above it, with π/2 split in two doubles so
// π/2 − x keeps full precision.  Both truncations are below 1e-19 on these
// ranges; the result is within about one ulp of sin(x).  Built without FMA
// contraction, it gives the same bits wherever it is evaluated.
__host__ __device__ inline double ramp_sin(double x) {
  if (x > 0.78539816339744831) {
    const double r = (1.5707963267948966 - x) + 6.123233995736766e-17;
    const double z = r * r;
    return 1.0 + z * (-0.5 + z * (4.1666666666666664e-02 + z * (-1.3888888888888889e-03 +
                 z * (2.4801587301587302e-05 + z * (-2.7557319223985888e-07 + z * (2.0876756987868100e-09 +
                 z * (-1.1470745597729725e-11 + z * 4.7794773323873853e-14)))))));
  }
  const double z = x * x;
  return x + x * (z * (-1.6666666666666666e-01 + z * (8.3333333333333332e-03 + z * (-1.9841269841269841e-04 +
             z * (2.7557319223985893e-06 + z * (-2.5052108385441720e-08 + z * (1.6059043836821613e-10 +
             z * (-7.6471637318198164e-13 + z * 2.8114572543455206e-15))))))));
}

// For T = float the ramp argument is formed in double (π is a double) and
// narrowed, as the reference's expression does; sin² is then taken of that
// T value (ramp_sin in double, narrowed to T once).
template <class T, int D>
__device__ __forceinline__ T disc_porosity(const T* rel, T rad, T eps) {
  T s2 = T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) s2 = s2 + rel[d] * rel[d];
  const T s = sqrt(s2);
  const T eps_h = eps / T(2);
  const T lo = rad - eps_h, hi = rad + eps_h;
  if (s <= lo) return T(0);
  if (s >= hi) return T(1);
  const T inner = T((3.14159265358979323846 / double(T(2) * eps)) * double(s - lo));
  const T si = T(ramp_sin(double(inner)));
  return si * si;
}

}  // namespace lbmx
