// Fused stream+collide, AA pattern, one population buffer.
//
// One launch is one reference step: the per-flag boundary and collision of
// the drivers' collide kernel (examples/poiseulle_2d_gpu/poiseulle_2d_gpu.cpp:
// 169-224) followed by the pull stream of every flag > 0 cell
// (stream.hpp:19-58, poiseulle_2d_gpu.cpp:384-399).
//
// Layout: SoA planes f[i][cell], cell = x + nx*(y + ny*z), no ghost ring.
// Phase 0 ("even" kernel) finds f_i(x) in plane i at x, collides and writes
// the post-collision f*_i(x) to plane opp(i) at x — every access is to the
// thread's own cell, fully coalesced.  Phase 1 ("odd" kernel) reads
// f_i(x) = f*_i(x − c_i) from plane opp(i) at x − c_i, collides and writes
// f*_i(x) to plane i at x + c_i, which leaves the natural layout again.
// Every plane element is read and written by exactly one thread per launch.
// Outside the lattice the reference's ghost ring holds zeros: such reads give
// 0, such writes are dropped, and in phase 1 a cell whose upstream neighbour
// is a ghost zeroes its own natural slot (nobody else streams into it).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace lbmx {

// ---- storage traits --------------------------------------------------------
template <class S, class T>
struct Store;
template <>
struct Store<double, double> {
  __device__ static double ld(const double* p, double) { return *p; }
  __device__ static void st(double* p, double v, double) { *p = v; }
  __device__ static double zero(double) { return 0.0; }
};
template <>
struct Store<float, float> {
  __device__ static float ld(const float* p, float) { return *p; }
  __device__ static void st(float* p, float v, float) { *p = v; }
};
// Shifted 16-bit storage: the stored value is f_i − w_i.
template <>
struct Store<__half, float> {
  __device__ static float ld(const __half* p, float w) { return __half2float(*p) + w; }
  __device__ static void st(__half* p, float v, float w) { *p = __float2half(v - w); }
};
template <>
struct Store<__hip_bfloat16, float> {
  __device__ static float ld(const __hip_bfloat16* p, float w) {
    return __bfloat162float(*p) + w;
  }
  __device__ static void st(__hip_bfloat16* p, float v, float w) { *p = __float2bfloat16(v - w); }
};

template <class L, class S, class T, bool WB, bool ODD>
__global__ __launch_bounds__(256) void k_aa_step(StepArgs a) {
  constexpr int Q = L::Q;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  const uint32_t cell = uint32_t(x) + uint32_t(nx) * (uint32_t(y) + uint32_t(ny) * uint32_t(z));
  S* __restrict__ fb = static_cast<S*>(a.f);
  const uint32_t Np = a.Np;

  // neighbour coordinates (wrapped) and validity per axis, for offsets −1, 0, +1
  const bool px = a.periodic & 1u, py = a.periodic & 2u, pz = a.periodic & 4u;
  const uint32_t X[3] = {uint32_t(x == 0 ? nx - 1 : x - 1), uint32_t(x),
                         uint32_t(x == nx - 1 ? 0 : x + 1)};
  const uint32_t Y[3] = {uint32_t(y == 0 ? ny - 1 : y - 1) * uint32_t(nx), uint32_t(y) * uint32_t(nx),
                         uint32_t(y == ny - 1 ? 0 : y + 1) * uint32_t(nx)};
  const uint32_t Z[3] = {uint32_t(z == 0 ? nz - 1 : z - 1) * sxy, uint32_t(z) * sxy,
                         uint32_t(z == nz - 1 ? 0 : z + 1) * sxy};
  const bool VX[3] = {x > 0 || px, true, x < nx - 1 || px};
  const bool VY[3] = {y > 0 || py, true, y < ny - 1 || py};
  const bool VZ[3] = {L::D < 3 || z > 0 || pz, true, L::D < 3 || z < nz - 1 || pz};

  T f[Q];
  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = Store<S, T>::ld(fb + size_t(i) * Np + cell, T(L::w[i]));
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 - L::cx[i], iy = 1 - L::cy[i], iz = 1 - L::cz[i];  // x − c_i
      if (VX[ix] && VY[iy] && VZ[iz]) {
        const uint32_t src = X[ix] + Y[iy] + Z[iz];
        f[i] = Store<S, T>::ld(fb + size_t(L::opp[i]) * Np + src, T(L::w[i]));
      } else {
        f[i] = T(0);
      }
    }
  }

  const uint8_t fl = a.flags[cell];
  if (fl < kMaxOps) {
    const OpDev* op = a.ops + fl;
    const int bc = op->bc;
    const int coll = op->coll;
    switch (bc) {
      case LBM_BC_BOUNCE_BACK: {
        T g[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) g[i] = f[L::opp[i]];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = g[i];
        break;
      }
      case LBM_BC_ANTI_BOUNCE_BACK: {
        if constexpr (Q == 9) {
          const T f3 = f[3], f5 = f[5], f6 = f[6];
          f[3] = f[4];
          f[4] = f3;
          f[5] = f[7];
          f[6] = f[8];
          f[7] = f5;
          f[8] = f6;
        }
        break;
      }
      case LBM_BC_ZOUHE_PRESSURE_X: {
        const double rs = (op->ramp_steps > 0 && a.t < op->ramp_steps)
                              ? op->rho_start + op->rho_slope * double(a.t)
                              : op->rho_final;
        if constexpr (Q == 9) zouhe_p_d2q9<T>(f, op->east != 0, T(rs));
        if constexpr (Q == 27) zouhe_p_d3q27<T>(f, op->east != 0, T(rs));
        break;
      }
      case LBM_BC_ZOUHE_VELOCITY_X: {
        if constexpr (Q == 9) {
          const T* bv = static_cast<const T*>(a.bcvel);
          zouhe_u_d2q9<T>(f, op->east != 0, bv[cell], bv[a.N + cell]);
        }
        break;
      }
      case LBM_BC_EQUILIBRIUM: {
        const T u[3] = {T(op->vec[0]), T(op->vec[1]), T(op->vec[2])};
        equilibrium<L, T, WB>(T(op->rho_eq), u, f);
        break;
      }
      default:
        break;
    }
    if (coll != LBM_COLL_NONE) {
      T rho, u[3];
      T B = T(1);
      T* por = static_cast<T*>(a.por);
      if (coll == LBM_COLL_PSM) {
        if (op->flags & LBM_OPF_PSM_STAMP_DISC) {
          const int xc[3] = {x, y, z};
          T rel[3];
#pragma unroll
          for (int d = 0; d < L::D; ++d) rel[d] = T(xc[d]) - T(op->disc[d]);
          B = disc_porosity<T, L::D>(rel, T(op->disc[3]), T(op->disc[4]));
          por[cell] = B;
        } else {
          B = por[cell];
        }
      }
      rho_u<L, T>(f, rho, u);
      if (a.write_macros) {
        static_cast<T*>(a.rho)[cell] = rho;
        T* um = static_cast<T*>(a.u);
#pragma unroll
        for (int d = 0; d < L::D; ++d) um[size_t(d) * a.N + cell] = u[d];
      }
      const T om = T(op->omega);
      if (coll == LBM_COLL_BGK) {
        bgk<L, T, WB>(f, rho, u, om);
      } else if (coll == LBM_COLL_BGK_GUO) {
        T F[3];
        const T* cf = static_cast<const T*>(a.force);
        const bool use_cf = op->flags & LBM_OPF_GUO_CELL_FORCE;
#pragma unroll
        for (int d = 0; d < L::D; ++d) F[d] = (use_cf ? cf[size_t(d) * a.N + cell] : T(0)) + T(op->vec[d]);
        bgk_guo<L, T, WB>(f, rho, u, om, F);
      } else if (coll == LBM_COLL_PSM) {
        psm<L, T, WB>(f, rho, u, om, B);
        if (op->flags & (LBM_OPF_PSM_CELL_FORCE | LBM_OPF_PSM_FORCE_GATHER)) {
          T m[3];
          opp_momentum<L, T>(f, m);
          const T sc = (op->flags & LBM_OPF_PSM_CELL_FORCE) ? B : T(1.0) - B;
          T* fo = static_cast<T*>(a.force);
#pragma unroll
          for (int d = 0; d < L::D; ++d) fo[size_t(d) * a.N + cell] = m[d] * sc;
        }
      }
    }
  }

  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i)
      Store<S, T>::st(fb + size_t(L::opp[i]) * Np + cell, f[i], T(L::w[i]));
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 + L::cx[i], iy = 1 + L::cy[i], iz = 1 + L::cz[i];  // x + c_i
      if (VX[ix] && VY[iy] && VZ[iz]) {
        const uint32_t dst = X[ix] + Y[iy] + Z[iz];
        Store<S, T>::st(fb + size_t(i) * Np + dst, f[i], T(L::w[i]));
      }
      const int jx = 1 - L::cx[i], jy = 1 - L::cy[i], jz = 1 - L::cz[i];
      if (!(VX[jx] && VY[jy] && VZ[jz])) Store<S, T>::st(fb + size_t(i) * Np + cell, T(0), T(L::w[i]));
    }
  }
}

template <class L, class S, class T, bool WB>
static hipError_t launch_t(bool odd, const StepArgs& a, const LaunchShape& sh, hipStream_t s) {
  dim3 block(sh.bx, sh.by, 1);
  dim3 grid((a.nx + sh.bx - 1) / sh.bx, (a.ny + sh.by - 1) / sh.by, a.nz);
  if (odd)
    hipLaunchKernelGGL((k_aa_step<L, S, T, WB, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_aa_step<L, S, T, WB, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <class L>
static hipError_t launch_l(int storage, bool odd, const StepArgs& a, const LaunchShape& sh,
                           hipStream_t s) {
  switch (storage) {
    case LBM_F64: return launch_t<L, double, double, true>(odd, a, sh, s);
    case LBM_F32: return launch_t<L, float, float, true>(odd, a, sh, s);
    case LBM_F16S: return launch_t<L, __half, float, false>(odd, a, sh, s);
    case LBM_BF16S: return launch_t<L, __hip_bfloat16, float, false>(odd, a, sh, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_aa_step(int lattice, int storage, bool odd, const StepArgs& a,
                          const LaunchShape& sh, hipStream_t s) {
  switch (lattice) {
    case LBM_D2Q9: return launch_l<D2Q9>(storage, odd, a, sh, s);
    case LBM_D3Q19: return launch_l<D3Q19>(storage, odd, a, sh, s);
    case LBM_D3Q27: return launch_l<D3Q27>(storage, odd, a, sh, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace lbmx
