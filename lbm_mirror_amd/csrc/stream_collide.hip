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

// Build-time tuning of the fused kernel: vector width for FP32 (4, 2, or 0 =
// scalar kernel), whether FP64 and the odd (shifted) phase use it.
#ifndef LBMX_VEC_F32
#define LBMX_VEC_F32 0
#endif
#ifndef LBMX_VEC_F64
#define LBMX_VEC_F64 0
#endif
#ifndef LBMX_VEC_ODD
#define LBMX_VEC_ODD 1
#endif
#ifndef LBMX_ROW_MIN_NX
#define LBMX_ROW_MIN_NX (1 << 30)
#endif
constexpr int kVecF32 = LBMX_VEC_F32;
constexpr int kRowKernelMinNx = LBMX_ROW_MIN_NX;
#ifndef LBMX_ROW_BLOCK
#define LBMX_ROW_BLOCK 256
#endif
constexpr int kRowBlock = LBMX_ROW_BLOCK;
constexpr bool kVecF64 = LBMX_VEC_F64 != 0;
constexpr bool kVecOdd = LBMX_VEC_ODD != 0;

// ---- storage traits --------------------------------------------------------
template <class S, class T>
struct Store;
template <>
struct Store<double, double> {
  template <class P>
  __device__ static double ld(P p, double) { return *p; }
  template <class P>
  __device__ static void st(P p, double v, double) { *p = v; }
  __device__ static double zero(double) { return 0.0; }
};
#ifndef LBMX_NT
#define LBMX_NT 0
#endif
template <>
struct Store<float, float> {
  template <class P>
  __device__ static float ld(P p, float) {
    if (LBMX_NT & 1) return __builtin_nontemporal_load(p);
    return *p;
  }
  template <class P>
  __device__ static void st(P p, float v, float) {
    if (LBMX_NT & 2) __builtin_nontemporal_store(v, p);
    else *p = v;
  }
};
// Shifted 16-bit storage: the stored value is f_i − w_i.
template <>
struct Store<__half, float> {
  // as raw 16-bit words: the class type cannot be accessed through a
  // global-typed pointer
  template <class P>
  __device__ static float ld(P p, float w) {
    return __half2float(__ushort_as_half(*(const __attribute__((address_space(1))) unsigned short*)p)) + w;
  }
  template <class P>
  __device__ static void st(P p, float v, float w) {
    *(__attribute__((address_space(1))) unsigned short*)p = __half_as_ushort(__float2half(v - w));
  }
};
template <>
struct Store<__hip_bfloat16, float> {
  template <class P>
  __device__ static float ld(P p, float w) {
    return __bfloat162float(__ushort_as_bfloat16(*(const __attribute__((address_space(1))) unsigned short*)p)) + w;
  }
  template <class P>
  __device__ static void st(P p, float v, float w) {
    *(__attribute__((address_space(1))) unsigned short*)p = __bfloat16_as_ushort(__float2bfloat16(v - w));
  }
};

// Boundary then collision for one cell (the drivers' switch(info)), writing
// the macros when asked.  f holds the pre-collision populations on entry and
// the post-collision ones on exit.
// A field of the lattice, typed as global memory: accesses through it are
// GLOBAL, not FLAT, also where the pointer was loaded from memory (batch
// members).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gfield(const void* p) {
  return (__attribute__((address_space(1))) T*)(p);
}

template <class L, class T, bool WB, class OpP>
__device__ __forceinline__ void apply_op(T (&f)[L::Q], OpP op, const StepArgs& a,
                                         uint32_t cell, int x, int y, int z) {
  constexpr int Q = L::Q;
  {
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
        const uint64_t t = a.t_dev ? *a.t_dev + a.t : a.t;
        const double rs = (op->ramp_steps > 0 && t < op->ramp_steps)
                              ? op->rho_start + op->rho_slope * double(t)
                              : op->rho_final;
        if constexpr (Q == 9) zouhe_p_d2q9<T>(f, op->east != 0, T(rs));
        if constexpr (Q == 27)
          zouhe_p_d3q27<T>(f, op->east != 0, T(rs), (op->flags & LBM_OPF_ZOUHE_NORMAL_ONLY) != 0);
        break;
      }
      case LBM_BC_ZOUHE_VELOCITY_X: {
        if constexpr (Q == 9) {
          const auto bv = gfield<const T>(a.bcvel);
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
      auto por = gfield<T>(a.por);
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
        gfield<T>(a.rho)[cell] = rho;
        auto um = gfield<T>(a.u);
        bool bad = !__builtin_isfinite(rho);
#pragma unroll
        for (int d = 0; d < L::D; ++d) {
          um[size_t(d) * a.N + cell] = u[d];
          bad = bad || !__builtin_isfinite(u[d]);
        }
        // one store per wavefront that meets a non-finite value
        const uint64_t m = __ballot(bad);
        if (m != 0 && a.status && __lane_id() == uint32_t(__ffsll((unsigned long long)m) - 1))
          __hip_atomic_store(a.status + kStatusNonFinite, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      const T om = T(op->omega);
      if (coll == LBM_COLL_BGK) {
        bgk<L, T, WB>(f, rho, u, om);
      } else if (coll == LBM_COLL_BGK_GUO) {
        T F[3];
        const auto cf = gfield<const T>(a.force);
        const bool use_cf = op->flags & LBM_OPF_GUO_CELL_FORCE;
#pragma unroll
        for (int d = 0; d < L::D; ++d) F[d] = (use_cf ? cf[size_t(d) * a.N + cell] : T(0)) + T(op->vec[d]);
        bgk_guo<L, T, WB>(f, rho, u, om, F, op->force_mask);
      } else if (coll == LBM_COLL_FPLBM) {
        // fplbm.hpp:60-133: BGKGuo's arithmetic with the cell force field and
        // u = m/rho; the macros hold rho and the momentum (FpLbmReset)
        T F[3] = {T(0), T(0), T(0)};
        const auto cf = gfield<const T>(a.force);
#pragma unroll
        for (int d = 0; d < L::D; ++d) F[d] = cf[size_t(d) * a.N + cell];
        if (a.write_macros) {
          T m[3] = {T(0), T(0), T(0)};
#pragma unroll
          for (int j = 1; j < Q; ++j)
#pragma unroll
            for (int d = 0; d < L::D; ++d)
              if (cdir<L>(j, d) != 0) m[d] = m[d] + T(cdir<L>(j, d)) * f[j];
          auto um = gfield<T>(a.u);
#pragma unroll
          for (int d = 0; d < L::D; ++d) um[size_t(d) * a.N + cell] = m[d];
        }
        bgk_guo<L, T, WB>(f, rho, u, om, F, 7u);
      } else if (coll == LBM_COLL_HLBM) {
        // hlbm.hpp:62-105: velocity u*B + N(1-B)/D (N when D = 0); the macros
        // receive that velocity, written here after the plain moments
        const T Bp = gfield<const T>(a.por)[cell];
        const T flip = T(1.0) - Bp;
        const T Dw = gfield<const T>(a.pd)[cell];
        const auto pn = gfield<const T>(a.pn);
#pragma unroll
        for (int d = 0; d < L::D; ++d) {
          const T Nd = pn[size_t(d) * a.N + cell];
          const T term = (Dw > T(0)) ? (Nd * flip) / Dw : Nd;
          u[d] = u[d] * Bp + term;
        }
        if (a.write_macros) {
          auto um = gfield<T>(a.u);
#pragma unroll
          for (int d = 0; d < L::D; ++d) um[size_t(d) * a.N + cell] = u[d];
        }
        bgk<L, T, WB>(f, rho, u, om);
      } else if (coll == LBM_COLL_PSM) {
        psm<L, T, WB>(f, rho, u, om, B);
        if (op->flags & (LBM_OPF_PSM_CELL_FORCE | LBM_OPF_PSM_FORCE_GATHER)) {
          T m[3];
          opp_momentum<L, T>(f, m);
          const T sc = (op->flags & LBM_OPF_PSM_CELL_FORCE) ? B : T(1.0) - B;
          auto fo = gfield<T>(a.force);
#pragma unroll
          for (int d = 0; d < L::D; ++d) fo[size_t(d) * a.N + cell] = m[d] * sc;
        }
      }
    }
  }
}

template <class L, class T, bool WB>
__device__ __forceinline__ void collide_cell(T (&f)[L::Q], uint8_t fl, const StepArgs& a,
                                             uint32_t cell, int x, int y, int z) {
  // When every active lane of the wavefront carries the same flag (the bulk
  // of any lattice), the operator is read with scalar loads and its kind is
  // a scalar branch; mixed wavefronts take the per-lane path.
  // The operator table is written only by the host between launches: typed
  // as constant memory, a wavefront-uniform operator is read with scalar
  // loads whatever the kernel stores before its fields are used (as global
  // memory, the fault-word store of the macro-writing branch made the
  // compiler read every operator field with vector loads and wait for them:
  // the D3Q19 phase 1 205 -> 240 us); per-lane reads are vector loads either
  // way.
  using GOp = const __attribute__((address_space(4))) OpDev;
  GOp* ops = (GOp*)a.ops;
  const int fl0 = __builtin_amdgcn_readfirstlane(int(fl));
  if (__all(int(fl) == fl0)) {
    if (fl0 < kMaxOps) apply_op<L, T, WB>(f, ops + fl0, a, cell, x, y, z);
  } else {
    if (fl < kMaxOps) apply_op<L, T, WB>(f, ops + fl, a, cell, x, y, z);
  }
}

#ifndef LBMX_NT_EVEN
#define LBMX_NT_EVEN 3
#endif
// Phase-0 accesses as nontemporal (bit 0: loads, bit 1: stores), FP32 only.
template <class S, class T, class P>
__device__ __forceinline__ T ld_even(P p, T w) {
  if constexpr (sizeof(S) == 4 && sizeof(T) == 4)
    if (LBMX_NT_EVEN & 1) return __builtin_nontemporal_load(p);
  return Store<S, T>::ld(p, w);
}
template <class S, class T, class P>
__device__ __forceinline__ void st_even(P p, T v, T w) {
  if constexpr (sizeof(S) == 4 && sizeof(T) == 4)
    if (LBMX_NT_EVEN & 2) {
      __builtin_nontemporal_store(v, p);
      return;
    }
  Store<S, T>::st(p, v, w);
}

// ---- population planes through a buffer resource ------------------------------
// One descriptor for the population array (scalar registers: base, byte range,
// format word of the gfx9 family), the plane's byte offset as the scalar
// offset and the element's byte offset as the 32-bit vector offset: a cell's
// accesses to all Q planes share one vector register, where the global form
// holds a 64-bit address per plane.  Used where the whole array's bytes fit
// the 32-bit range (Q * Np * sizeof(S) < 2^32); an access outside the
// descriptor's range would read 0 and write nothing, and the kernels make none.
// Per lattice, the phases that take it: bit 0 phase 0, bit 1 phase 1.
// D3Q27 FP32 at 256^3: phase 0 drops from 168 VGPRs (with spills) to 128,
// and the step runs 4% faster in both phases (24.4 -> 25.4 GLUPS); D3Q19
// gains nothing (phase 0 2% slower, phase 1 within noise); D2Q9 FP32 with 8
// waves per SIMD, 1.5% on the C3 particle step.
#ifndef LBMX_BUF_Q9
#define LBMX_BUF_Q9 3
#endif
#ifndef LBMX_BUF_Q19
#define LBMX_BUF_Q19 0
#endif
#ifndef LBMX_BUF_Q27
#define LBMX_BUF_Q27 3
#endif
// FP32 storage only: where it was measured (the D2Q9 FP64 step, C1, is 2%
// slower in the buffer form)
template <class L, class S>
constexpr int buf_phases() {
  if (sizeof(S) != 4) return 0;
  return L::Q == 9 ? LBMX_BUF_Q9 : (L::Q == 19 ? LBMX_BUF_Q19 : LBMX_BUF_Q27);
}
// AUX: cache policy of the access (2: nontemporal)
template <class S, class T>
struct BufIO;
template <>
struct BufIO<float, float> {
  template <int AUX>
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, so, AUX));
  }
  template <int AUX>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float v, float) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, i * 4u, so, AUX);
  }
};
template <>
struct BufIO<double, double> {
  using V2 = __attribute__((__vector_size__(2 * sizeof(unsigned int)))) unsigned int;
  template <int AUX>
  __device__ static double ld(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, so, AUX));
  }
  template <int AUX>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, double v, double) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, i * 8u, so, AUX);
  }
};
template <>
struct BufIO<__half, float> {
  template <int AUX>
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float w) {
    return __half2float(__ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, so, AUX))) + w;
  }
  template <int AUX>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float v, float w) {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half(v - w)), r, i * 2u, so, AUX);
  }
};
template <>
struct BufIO<__hip_bfloat16, float> {
  template <int AUX>
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float w) {
    return __bfloat162float(__ushort_as_bfloat16(__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, so, AUX))) + w;
  }
  template <int AUX>
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t i, uint32_t so, float v, float w) {
    __builtin_amdgcn_raw_buffer_store_b16(__bfloat16_as_ushort(__float2bfloat16(v - w)), r, i * 2u, so, AUX);
  }
};
// the cache policies of the global form: phase 0 nontemporal per
// LBMX_NT_EVEN (FP32), phase 1 per LBMX_NT (FP32)
template <class S, class T>
constexpr int aux_even_ld() { return (sizeof(S) == 4 && sizeof(T) == 4 && (LBMX_NT_EVEN & 1)) ? 2 : 0; }
template <class S, class T>
constexpr int aux_even_st() { return (sizeof(S) == 4 && sizeof(T) == 4 && (LBMX_NT_EVEN & 2)) ? 2 : 0; }
template <class S, class T>
constexpr int aux_odd_ld() { return (sizeof(S) == 4 && sizeof(T) == 4 && (LBMX_NT & 1)) ? 2 : 0; }
template <class S, class T>
constexpr int aux_odd_st() { return (sizeof(S) == 4 && sizeof(T) == 4 && (LBMX_NT & 2)) ? 2 : 0; }

// Minimum waves per SIMD asked of the compiler for the fused step, per
// lattice, compute type, phase and addressing form (1: no request).  D3Q27
// FP32: phase 0 in the buffer form held to 4 waves (128 VGPRs), everything
// else to 3 (phase 1 167 VGPRs, no spills): 256^3 phase 0 671 us against
// 717 us at 3 waves; the global form at 4 waves spills (1250 us).  Asking
// D3Q19 for 4 waves fits it in 128 VGPRs without spills but gains nothing.
#ifndef LBMX_WAVES_Q19
#define LBMX_WAVES_Q19 1
#endif
#ifndef LBMX_WAVES_Q27
#define LBMX_WAVES_Q27 4
#endif
#ifndef LBMX_WAVES_Q27_ODD
#define LBMX_WAVES_Q27_ODD 3
#endif
// D2Q9 FP32: 8 waves (64 VGPRs) with the buffer form (phase 0 then needs 60
// and no spills): the C3 particle step 1.5% faster; 8 waves alone, no change
#ifndef LBMX_WAVES_Q9
#define LBMX_WAVES_Q9 8
#endif
template <class L, class T, bool ODD, bool BUF>
struct StepWaves {
  static constexpr int value = 1;
};
template <bool ODD, bool BUF>
struct StepWaves<D3Q19, float, ODD, BUF> {
  static constexpr int value = LBMX_WAVES_Q19;
};
template <bool ODD, bool BUF>
struct StepWaves<D3Q27, float, ODD, BUF> {
  static constexpr int value = (ODD || !BUF) ? LBMX_WAVES_Q27_ODD : LBMX_WAVES_Q27;
};
template <bool ODD, bool BUF>
struct StepWaves<D2Q9, float, ODD, BUF> {
  static constexpr int value = LBMX_WAVES_Q9;
};
// ROW1: every block is one row segment (blockDim.y == 1), so y and z, the
// offsets of the neighbour rows and their validity are wavefront-uniform
// (scalar registers) and each population access is a uniform row base plus
// a 32-bit per-lane x offset; only x and its two neighbours live per lane.
template <class L, class S, class T, bool WB, bool ODD, bool BUF, bool ROW1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(StepWaves<L, T, ODD, BUF>::value))) void k_aa_step(
    StepArgs a) {
  constexpr int Q = L::Q;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = ROW1 ? int(blockIdx.y) : int(blockIdx.y * blockDim.y + threadIdx.y);
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  const uint32_t row = uint32_t(nx) * (uint32_t(y) + uint32_t(ny) * uint32_t(z));
  const uint32_t cell = row + uint32_t(x);
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

  // plane p, row offset r, x offset e: load and store in this phase's cache
  // policy (r joins the plane into the scalar part of the address when ROW1)
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(a.f, 0, int(uint32_t(Q) * Np * uint32_t(sizeof(S))), 0x00020000);
  // (a buffer access's scalar offset must be uniform: without ROW1 the row
  // offset joins the per-lane part)
  auto ld = [&](int p, uint32_t r, uint32_t e, T w) -> T {
    if constexpr (BUF && ROW1)
      return BufIO<S, T>::template ld<ODD ? aux_odd_ld<S, T>() : aux_even_ld<S, T>()>(
          rs, e, (uint32_t(p) * Np + r) * uint32_t(sizeof(S)), w);
    else if constexpr (BUF)
      return BufIO<S, T>::template ld<ODD ? aux_odd_ld<S, T>() : aux_even_ld<S, T>()>(
          rs, r + e, uint32_t(p) * Np * uint32_t(sizeof(S)), w);
    else if constexpr (ODD)
      return Store<S, T>::ld(fb + (size_t(p) * Np + r) + e, w);
    else
      return ld_even<S, T>(fb + (size_t(p) * Np + r) + e, w);
  };
  auto st = [&](int p, uint32_t r, uint32_t e, T v, T w) {
    if constexpr (BUF && ROW1)
      BufIO<S, T>::template st<ODD ? aux_odd_st<S, T>() : aux_even_st<S, T>()>(
          rs, e, (uint32_t(p) * Np + r) * uint32_t(sizeof(S)), v, w);
    else if constexpr (BUF)
      BufIO<S, T>::template st<ODD ? aux_odd_st<S, T>() : aux_even_st<S, T>()>(
          rs, r + e, uint32_t(p) * Np * uint32_t(sizeof(S)), v, w);
    else if constexpr (ODD)
      Store<S, T>::st(fb + (size_t(p) * Np + r) + e, v, w);
    else
      st_even<S, T>(fb + (size_t(p) * Np + r) + e, v, w);
  };

  T f[Q];
  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = ld(i, row, uint32_t(x), T(L::w[i]));
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 - L::cx[i], iy = 1 - L::cy[i], iz = 1 - L::cz[i];  // x − c_i
      const bool inside = VX[ix] && VY[iy] && VZ[iz];
      if constexpr (sizeof(S) == 2) {
        // a 16-bit value is converted after its load: a load in a per-lane
        // branch whose converted result leaves the branch holds the
        // wavefront there until it returns (one memory latency per
        // direction), so the load is unconditional (the wrapped neighbour
        // always lies inside the lattice) and the ghost's zero is selected
        // after it
        const T v = ld(L::opp[i], Y[iy] + Z[iz], X[ix], T(L::w[i]));
        f[i] = inside ? v : T(0);
      } else if (inside) {
        f[i] = ld(L::opp[i], Y[iy] + Z[iz], X[ix], T(L::w[i]));
      } else {
        f[i] = T(0);
      }
    }
  }

  collide_cell<L, T, WB>(f, ((const __attribute__((address_space(1))) uint8_t*)a.flags + row)[x], a, cell, x,
                         y, z);

  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) st(L::opp[i], row, uint32_t(x), f[i], T(L::w[i]));
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 + L::cx[i], iy = 1 + L::cy[i], iz = 1 + L::cz[i];  // x + c_i
      if (VX[ix] && VY[iy] && VZ[iz]) st(i, Y[iy] + Z[iz], X[ix], f[i], T(L::w[i]));
      const int jx = 1 - L::cx[i], jy = 1 - L::cy[i], jz = 1 - L::cz[i];
      if (!(VX[jx] && VY[jy] && VZ[jz])) st(i, row, uint32_t(x), T(0), T(L::w[i]));
    }
  }
}


// Row-block variant of the scalar kernel: a block covers a contiguous run of
// one row (blockIdx.y = y, blockIdx.z = z), so the row offsets of the ±1
// neighbour rows and their validity are wavefront-uniform, and the partial
// cache lines that the ±1 shifts along x create stay inside one block.
// One cell of one row: y and z are uniform across the calling wavefront,
// x varies per lane (x < nx for every lane that calls).
template <class L, class S, class T, bool WB, bool ODD>
__device__ __forceinline__ void aa_row_cell(const StepArgs& a, int x, int y, int z) {
  constexpr int Q = L::Q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  const bool px = a.periodic & 1u, py = a.periodic & 2u, pz = a.periodic & 4u;
  // wavefront-uniform: the base of every (plane, neighbour row) pair, and
  // whether that row exists
  const uint32_t Y[3] = {uint32_t(y == 0 ? ny - 1 : y - 1) * uint32_t(nx), uint32_t(y) * uint32_t(nx),
                         uint32_t(y == ny - 1 ? 0 : y + 1) * uint32_t(nx)};
  const uint32_t Z[3] = {uint32_t(z == 0 ? nz - 1 : z - 1) * sxy, uint32_t(z) * sxy,
                         uint32_t(z == nz - 1 ? 0 : z + 1) * sxy};
  const bool VY[3] = {y > 0 || py, true, y < ny - 1 || py};
  const bool VZ[3] = {L::D < 3 || z > 0 || pz, true, L::D < 3 || z < nz - 1 || pz};
  // per lane: x neighbours, as 32-bit offsets within a row
  const uint32_t X[3] = {uint32_t(x == 0 ? nx - 1 : x - 1), uint32_t(x),
                         uint32_t(x == nx - 1 ? 0 : x + 1)};
  const bool VX[3] = {x > 0 || px, true, x < nx - 1 || px};
  const uint32_t row = Y[1] + Z[1];
  // a batch member's pointers are loaded from memory, where the compiler
  // cannot tell they address global memory: typing them global keeps the
  // population accesses GLOBAL rather than FLAT (which wait on both the LDS
  // and the vector-memory counters)
  using GS = __attribute__((address_space(1))) S;
  GS* __restrict__ fb = (GS*)static_cast<S*>(a.f);
  const size_t Np = a.Np;
  auto rowp = [&](int plane, int iy, int iz) -> GS* {
    return fb + (size_t(plane) * Np + (Y[iy] + Z[iz]));
  };
  // every neighbour of every lane in the lattice (or wrapped): no predication
  const bool inner_rows = VY[0] && VY[2] && VZ[0] && VZ[2];
  const bool inner = inner_rows && __all(VX[0] && VX[2]);

  T f[Q];
  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = ld_even<S, T>(rowp(i, 1, 1) + X[1], T(L::w[i]));
  } else {
    // one unconditional (wrapped, in-lattice) load per direction, then the
    // ghost's zero: no loads in per-lane branches.  (An all-inner arm beside a
    // predicated arm made the compiler wait for every load in flight before
    // the second arm reused the first arm's destination registers: one memory
    // latency more per cell.)
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 - L::cx[i], iy = 1 - L::cy[i], iz = 1 - L::cz[i];  // x − c_i
      const bool inside = VX[ix] && VY[iy] && VZ[iz];
      const T v = Store<S, T>::ld(rowp(L::opp[i], iy, iz) + X[ix], T(L::w[i]));
      f[i] = inside ? v : T(0);
    }
  }

  collide_cell<L, T, WB>(f, ((const __attribute__((address_space(1))) uint8_t*)a.flags)[row + X[1]], a,
                        row + X[1], x, y, z);

  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) st_even<S, T>(rowp(L::opp[i], 1, 1) + X[1], f[i], T(L::w[i]));
  } else if (inner) {
#pragma unroll
    for (int i = 0; i < Q; ++i)
      Store<S, T>::st(rowp(i, 1 + L::cy[i], 1 + L::cz[i]) + X[1 + L::cx[i]], f[i], T(L::w[i]));
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int ix = 1 + L::cx[i], iy = 1 + L::cy[i], iz = 1 + L::cz[i];  // x + c_i
      if (VX[ix] && VY[iy] && VZ[iz]) Store<S, T>::st(rowp(i, iy, iz) + X[ix], f[i], T(L::w[i]));
      const int jx = 1 - L::cx[i], jy = 1 - L::cy[i], jz = 1 - L::cz[i];
      if (!(VX[jx] && VY[jy] && VZ[jz])) Store<S, T>::st(rowp(i, 1, 1) + X[1], T(0), T(L::w[i]));
    }
  }
}

template <class L, class S, class T, bool WB, bool ODD>
__global__ __launch_bounds__(1024) void k_aa_row(StepArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < a.nx) aa_row_cell<L, S, T, WB, ODD>(a, x, blockIdx.y, blockIdx.z);
}

// ---- 16-bit storage: two cells per thread ------------------------------------
// A thread takes cells x0 = 2k and x0 + 1 of one row (one-row blocks) and
// moves each plane's two 16-bit values as one 32-bit word: the FP32 kernel's
// bytes per lane and per memory instruction at half the bytes per cell.  In
// phase 1 the x-shifted words start at an odd element, an unaligned access
// that gfx950's memory path serves directly (the compiler's unaligned-access
// mode).  Phase-1 loads are all issued before any value is used (no load in
// a per-lane branch); where a pair would cross the row's end (periodic wrap,
// or the zero ghost ring) the word is clamped into the row and the missing
// value loaded on its own, and such pairs are stored one value at a time.
// Even rows only (nx even); every element is still read and written by
// exactly one thread.
template <class S>
struct Bits16;
template <>
struct Bits16<__half> {
  __device__ static float f(uint32_t b) { return __half2float(__ushort_as_half((unsigned short)b)); }
  __device__ static uint32_t b(float v) { return __half_as_ushort(__float2half(v)); }
};
template <>
struct Bits16<__hip_bfloat16> {
  __device__ static float f(uint32_t b) { return __bfloat162float(__ushort_as_bfloat16((unsigned short)b)); }
  __device__ static uint32_t b(float v) { return __bfloat16_as_ushort(__float2bfloat16(v)); }
};

#ifndef LBMX_NT_PAIR
#define LBMX_NT_PAIR 3
#endif
template <class L, class S, bool ODD>
__global__ __launch_bounds__(256) void k_aa_pair(StepArgs a) {
  using T = float;
  constexpr int Q = L::Q;
  using G16 = __attribute__((address_space(1))) uint16_t;
  using G32 = __attribute__((address_space(1))) uint32_t;
  typedef uint32_t __attribute__((aligned(2))) u32_a2;  // a word at an even byte address
  using G32u = __attribute__((address_space(1))) u32_a2;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = 2 * int(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = blockIdx.y, z = blockIdx.z;  // one-row blocks: uniform
  if (x0 >= nx) return;                     // nx even: x0 + 1 < nx too
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  const uint32_t row = uint32_t(nx) * (uint32_t(y) + uint32_t(ny) * uint32_t(z));
  G16* fb = (G16*)a.f;
  const size_t Np = a.Np;
  const bool px = a.periodic & 1u, py = a.periodic & 2u, pz = a.periodic & 4u;
  const uint32_t Y[3] = {uint32_t(y == 0 ? ny - 1 : y - 1) * uint32_t(nx), uint32_t(y) * uint32_t(nx),
                         uint32_t(y == ny - 1 ? 0 : y + 1) * uint32_t(nx)};
  const uint32_t Z[3] = {uint32_t(z == 0 ? nz - 1 : z - 1) * sxy, uint32_t(z) * sxy,
                         uint32_t(z == nz - 1 ? 0 : z + 1) * sxy};
  const bool VY[3] = {y > 0 || py, true, y < ny - 1 || py};
  const bool VZ[3] = {L::D < 3 || z > 0 || pz, true, L::D < 3 || z < nz - 1 || pz};
  // the x position of x + d inside the row, or -1 in the ghost ring
  auto wrap = [&](int x) -> int { return x < 0 ? (px ? x + nx : -1) : (x >= nx ? (px ? x - nx : -1) : x); };

  T f[2][Q];
  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const G32* p = (const G32*)(fb + (size_t(i) * Np + row + uint32_t(x0)));
      const uint32_t v = (LBMX_NT_PAIR & 1) ? __builtin_nontemporal_load(p) : *p;
      f[0][i] = Bits16<S>::f(v & 0xffffu) + T(L::w[i]);
      f[1][i] = Bits16<S>::f(v >> 16) + T(L::w[i]);
    }
  } else {
    // Every load is issued before any value is used: a load whose converted
    // value left a per-lane branch would hold the wavefront there until it
    // returned.  Direction i's sources are the word at x0 − c_x (cells x0 and
    // x0 + 1), clamped into the row; a source row in the ghost ring reads this
    // cell's own row instead (the value is then replaced by the ghost's zero).
    // The row's first pair (c_x = +1) and last pair (c_x = −1) take one of
    // their two values from the row's other end (periodic wrap) or the ghost.
    uint32_t raw[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int iy = 1 - L::cy[i], iz = 1 - L::cz[i];
      const uint32_t rb = (VY[iy] && VZ[iz]) ? Y[iy] + Z[iz] : row;
      const int xs = x0 - L::cx[i];
      const int xa = L::cx[i] > 0 ? (xs < 0 ? 0 : xs) : (L::cx[i] < 0 ? (xs > nx - 2 ? nx - 2 : xs) : xs);
      raw[i] = *(const G32u*)(fb + (size_t(L::opp[i]) * Np + rb + uint32_t(xa)));
    }
    const bool first = x0 == 0, last = x0 == nx - 2;
    uint32_t edge[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) edge[i] = 0u;
    if (__any(first || last)) {  // the wavefronts that hold a row end
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        if (L::cx[i] == 0) continue;
        const int iy = 1 - L::cy[i], iz = 1 - L::cz[i];
        const uint32_t rb = (VY[iy] && VZ[iz]) ? Y[iy] + Z[iz] : row;
        edge[i] = fb[size_t(L::opp[i]) * Np + rb + uint32_t(L::cx[i] > 0 ? nx - 1 : 0)];
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int iy = 1 - L::cy[i], iz = 1 - L::cz[i];
      const bool rv = VY[iy] && VZ[iz];
      uint32_t lo = raw[i] & 0xffffu, hi = raw[i] >> 16;
      bool g0 = !rv, g1 = !rv;
      if (L::cx[i] > 0) {
        hi = first ? lo : hi;
        lo = first ? edge[i] : lo;
        g0 = g0 || (first && !px);
      } else if (L::cx[i] < 0) {
        lo = last ? hi : lo;
        hi = last ? edge[i] : hi;
        g1 = g1 || (last && !px);
      }
      const T w = T(L::w[i]);
      f[0][i] = g0 ? T(0) : Bits16<S>::f(lo) + w;
      f[1][i] = g1 ? T(0) : Bits16<S>::f(hi) + w;
    }
  }

  const uint32_t fl2 = ((const G16*)((const __attribute__((address_space(1))) uint8_t*)a.flags + row))[x0 >> 1];
  collide_cell<L, T, false>(f[0], uint8_t(fl2 & 0xffu), a, row + uint32_t(x0), x0, y, z);
  collide_cell<L, T, false>(f[1], uint8_t(fl2 >> 8), a, row + uint32_t(x0) + 1u, x0 + 1, y, z);

  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const T w = T(L::w[i]);
      const uint32_t v = Bits16<S>::b(f[0][i] - w) | (Bits16<S>::b(f[1][i] - w) << 16);
      G32* p = (G32*)(fb + (size_t(L::opp[i]) * Np + row + uint32_t(x0)));
      if (LBMX_NT_PAIR & 2) __builtin_nontemporal_store(v, p);
      else *p = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const T w = T(L::w[i]);
      const uint32_t b0 = Bits16<S>::b(f[0][i] - w), b1 = Bits16<S>::b(f[1][i] - w);
      const int jy = 1 + L::cy[i], jz = 1 + L::cz[i];
      const int xd = x0 + L::cx[i];  // the destination of cell x0 (x0 + 1's is xd + 1)
      if (VY[jy] && VZ[jz]) {
        G16* base = fb + (size_t(i) * Np + Y[jy] + Z[jz]);
        if (xd >= 0 && xd + 1 < nx) {
          *(G32u*)(base + uint32_t(xd)) = b0 | (b1 << 16);
        } else {
          const int d0 = wrap(xd), d1 = wrap(xd + 1);
          if (d0 >= 0) base[d0] = (uint16_t)b0;
          if (d1 >= 0) base[d1] = (uint16_t)b1;
        }
      }
      // a cell whose upstream neighbour is in the ghost ring: its own slot
      // takes the ghost's zero (nobody else streams into it)
      const int uy = 1 - L::cy[i], uz = 1 - L::cz[i];
      const bool rows_ok = VY[uy] && VZ[uz];
      const bool up0 = rows_ok && wrap(x0 - L::cx[i]) >= 0, up1 = rows_ok && wrap(x0 + 1 - L::cx[i]) >= 0;
      if (!up0 || !up1) {
        G16* own = fb + (size_t(i) * Np + row + uint32_t(x0));
        const uint32_t zb = Bits16<S>::b(T(0) - w);
        if (!up0) own[0] = (uint16_t)zb;
        if (!up1) own[1] = (uint16_t)zb;
      }
    }
  }
}

// ---- vectorised variant ------------------------------------------------------
// Each thread owns V consecutive cells of one row and moves every plane with
// one aligned V-wide access.  In phase 1 the ±1 shifts along x are resolved
// inside the 64-wide wavefront: a lane takes the element it needs from its
// neighbour lane's aligned load (__shfl_up/__shfl_down) and hands its edge
// value to the neighbour for the shifted stores; only the first and last lane
// of a row segment (row start/end, or wavefront boundary) issue one extra
// scalar access.  Every plane element is still written exactly once.
template <class S, int V>
struct VecT;
template <>
struct VecT<float, 4> {
  using type = float4;
};
template <>
struct VecT<float, 2> {
  using type = float2;
};
template <>
struct VecT<double, 2> {
  using type = double2;
};

template <class S, int V>
__device__ __forceinline__ void vload(const S* p, S (&e)[V]) {
  const typename VecT<S, V>::type t = *reinterpret_cast<const typename VecT<S, V>::type*>(p);
  const S* q = reinterpret_cast<const S*>(&t);
#pragma unroll
  for (int v = 0; v < V; ++v) e[v] = q[v];
}
template <class S, int V>
__device__ __forceinline__ void vstore(S* p, const S (&e)[V]) {
  typename VecT<S, V>::type t;
  S* q = reinterpret_cast<S*>(&t);
#pragma unroll
  for (int v = 0; v < V; ++v) q[v] = e[v];
  *reinterpret_cast<typename VecT<S, V>::type*>(p) = t;
}

template <class L, class T, bool WB, bool ODD, int V>
__global__ __launch_bounds__(256) void k_aa_vec(StepArgs a) {
  constexpr int Q = L::Q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = (blockIdx.x * blockDim.x + threadIdx.x) * V;
  const int yt = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const bool active = x0 < nx && yt < ny;
  const int y = active ? yt : 0;
  const int lane = __lane_id();
  const bool seg_start = x0 == 0 || lane == 0;
  const bool seg_end = x0 + V >= nx || lane == 63;
  const uint32_t sxy = uint32_t(nx) * uint32_t(ny);
  const bool px = a.periodic & 1u, py = a.periodic & 2u, pz = a.periodic & 4u;
  const uint32_t Y[3] = {uint32_t(y == 0 ? ny - 1 : y - 1) * uint32_t(nx), uint32_t(y) * uint32_t(nx),
                         uint32_t(y == ny - 1 ? 0 : y + 1) * uint32_t(nx)};
  const uint32_t Z[3] = {uint32_t(z == 0 ? nz - 1 : z - 1) * sxy, uint32_t(z) * sxy,
                         uint32_t(z == nz - 1 ? 0 : z + 1) * sxy};
  const bool VY[3] = {y > 0 || py, true, y < ny - 1 || py};
  const bool VZ[3] = {L::D < 3 || z > 0 || pz, true, L::D < 3 || z < nz - 1 || pz};
  const uint32_t row = Y[1] + Z[1];
  T* __restrict__ fb = static_cast<T*>(a.f);
  const size_t Np = a.Np;

  T f[V][Q];
  if (!ODD) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      T e[V];
      if (active) vload<T, V>(fb + size_t(i) * Np + row + x0, e);
      else
#pragma unroll
        for (int v = 0; v < V; ++v) e[v] = T(0);
#pragma unroll
      for (int v = 0; v < V; ++v) f[v][i] = e[v];
    }
  } else {
    // all aligned loads first, then the edge elements, then the shuffles, so
    // that every load of the cell block is in flight at once
    T e[Q][V];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cy = L::cy[i], cz = L::cz[i];
      const bool src_ok = active && VY[1 - cy] && VZ[1 - cz];
      const T* pl = fb + size_t(L::opp[i]) * Np + (Y[1 - cy] + Z[1 - cz]);
      if (src_ok) vload<T, V>(pl + x0, e[i]);
      else
#pragma unroll
        for (int v = 0; v < V; ++v) e[i][v] = T(0);
    }
    T edge[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::cx[i], cy = L::cy[i], cz = L::cz[i];
      const bool src_ok = active && VY[1 - cy] && VZ[1 - cz];
      const T* pl = fb + size_t(L::opp[i]) * Np + (Y[1 - cy] + Z[1 - cz]);
      edge[i] = T(0);
      if (cx == 1) {  // element x0-1 for the first lane of a segment
        const bool inrow = x0 > 0;
        if (seg_start && src_ok && (inrow || px)) edge[i] = pl[inrow ? x0 - 1 : nx - 1];
      } else if (cx == -1) {  // element x0+V for the last lane of a segment
        const bool inrow = x0 + V < nx;
        if (seg_end && src_ok && (inrow || px)) edge[i] = pl[inrow ? x0 + V : 0];
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::cx[i];
      if (cx == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) f[v][i] = e[i][v];
      } else if (cx == 1) {  // needs x0-1 .. x0+V-2
        const T up = __shfl_up(e[i][V - 1], 1, 64);
        f[0][i] = seg_start ? edge[i] : up;
#pragma unroll
        for (int v = 1; v < V; ++v) f[v][i] = e[i][v - 1];
      } else {  // needs x0+1 .. x0+V
        const T dn = __shfl_down(e[i][0], 1, 64);
#pragma unroll
        for (int v = 0; v < V - 1; ++v) f[v][i] = e[i][v + 1];
        f[V - 1][i] = seg_end ? edge[i] : dn;
      }
    }
  }

  if (active) {
    uint8_t fl[V];
    if (V == 4) {
      const uchar4 q = *reinterpret_cast<const uchar4*>(a.flags + row + x0);
      fl[0] = q.x;
      fl[1 % V] = q.y;
      fl[2 % V] = q.z;
      fl[3 % V] = q.w;
    } else {
      const uchar2 q = *reinterpret_cast<const uchar2*>(a.flags + row + x0);
      fl[0] = q.x;
      fl[1 % V] = q.y;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) collide_cell<L, T, WB>(f[v], fl[v], a, row + x0 + v, x0 + v, y, z);
  }

  if (!ODD) {
    if (active) {
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        T e[V];
#pragma unroll
        for (int v = 0; v < V; ++v) e[v] = f[v][i];
        vstore<T, V>(fb + size_t(L::opp[i]) * Np + row + x0, e);
      }
    }
  } else {
    T nb[Q];  // neighbour lane's edge value per x-shifted plane
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::cx[i];
      nb[i] = T(0);
      if (cx == 1) nb[i] = __shfl_up(f[V - 1][i], 1, 64);
      if (cx == -1) nb[i] = __shfl_down(f[0][i], 1, 64);
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::cx[i], cy = L::cy[i], cz = L::cz[i];
      T* pl = fb + size_t(i) * Np;
      // own row: its upstream row is a ghost row, nobody streams into it
      if (active && !(VY[1 - cy] && VZ[1 - cz])) {
        T zr[V];
#pragma unroll
        for (int v = 0; v < V; ++v) zr[v] = T(0);
        vstore<T, V>(pl + row + x0, zr);
      }
      const bool dst_ok = active && VY[1 + cy] && VZ[1 + cz];
      T* pd = pl + (Y[1 + cy] + Z[1 + cz]);
      if (!dst_ok) continue;
      if (cx == 0) {
        T ev[V];
#pragma unroll
        for (int v = 0; v < V; ++v) ev[v] = f[v][i];
        vstore<T, V>(pd + x0, ev);
      } else if (cx == 1) {  // cell x0+v lands on x0+v+1
        if (!seg_start) {
          T ev[V];
          ev[0] = nb[i];
#pragma unroll
          for (int v = 1; v < V; ++v) ev[v] = f[v - 1][i];
          vstore<T, V>(pd + x0, ev);
        } else {
#pragma unroll
          for (int v = 1; v < V; ++v) pd[x0 + v] = f[v - 1][i];
          if (x0 == 0 && !px) pd[0] = T(0);  // upstream of x = 0 is the ghost column
        }
        if (seg_end) {
          if (x0 + V < nx) pd[x0 + V] = f[V - 1][i];
          else if (px) pd[0] = f[V - 1][i];
        }
      } else {  // cx == -1: cell x0+v lands on x0+v-1
        if (!seg_end) {
          T ev[V];
#pragma unroll
          for (int v = 0; v < V - 1; ++v) ev[v] = f[v + 1][i];
          ev[V - 1] = nb[i];
          vstore<T, V>(pd + x0, ev);
        } else {
#pragma unroll
          for (int v = 0; v < V - 1; ++v) pd[x0 + v] = f[v + 1][i];
          if (x0 + V >= nx && !px) pd[x0 + V - 1] = T(0);  // upstream is the ghost column
        }
        if (seg_start) {
          if (x0 > 0) pd[x0 - 1] = f[0][i];
          else if (px) pd[nx - 1] = f[0][i];
        }
      }
    }
  }
}

template <class L, class T, bool WB, int V>
static hipError_t launch_vec(bool odd, const StepArgs& a, hipStream_t s) {
  const int tx = (a.nx + V - 1) / V;  // threads per row
  int bx = 1;
  while (bx < tx && bx < 256) bx *= 2;
  const int by = 256 / bx;
  dim3 block(bx, by, 1);
  dim3 grid((tx + bx - 1) / bx, (a.ny + by - 1) / by, a.nz);
  if (odd)
    hipLaunchKernelGGL((k_aa_vec<L, T, WB, true, V>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_aa_vec<L, T, WB, false, V>), grid, block, 0, s, a);
  return hipGetLastError();
}

// ---- batched ensembles --------------------------------------------------------
// Many independent lattices of uneven size, one launch per step: persistent
// blocks stride over a work list of tiles ordered member by member; a tile is
// rows x blockDim consecutive cells of one member in natural order (x fastest,
// wrapping across rows), so every lane has a cell whatever the member's width
// except in the member's last tile; each cell runs the single-lattice body.
// The step index is read from device memory (t0 + k) so a chain of steps can
// be captured once in a hipGraph and replayed.
// D2Q9 batch: 6 waves per SIMD (≤ 80 VGPRs; it takes 90 unconstrained, 5
// waves): 67.4 -> 69.7 GLUPS on C4; 5 waves 67.1, 7 waves 63.3, 8 spill (50.3)
#ifndef LBMX_BATCH_WAVES
#define LBMX_BATCH_WAVES 6
#endif
template <class L, class S, class T, bool WB, bool ODD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L::Q == 9 ? LBMX_BATCH_WAVES : 1))) void k_aa_batch(
    BatchArgs b) {
  const uint64_t t = *b.t0 + b.k;
  for (uint32_t tile = blockIdx.x; tile < b.ntiles; tile += gridDim.x) {
    const Tile tl = b.tiles[tile];
    StepArgs a = b.members[tl.member];
    a.t = t;
    a.t_dev = nullptr;
    // batch members report divergence through their residual (NaN); without
    // the fault-word store the kernel keeps its registers (it spilled with it)
    a.status = nullptr;
    a.write_macros = b.write_macros != 0;
    if (b.write_macros == 2) a.u = b.u_prev[tl.member];

    const uint32_t nx = uint32_t(a.nx), sxy = nx * uint32_t(a.ny);
    for (uint32_t r = 0; r < b.rows; ++r) {
      const uint32_t c = tl.c0 + r * blockDim.x + threadIdx.x;
      if (c >= a.N) break;
      const uint32_t z = c / sxy, rem = c - z * sxy, y = rem / nx;
      aa_row_cell<L, S, T, WB, ODD>(a, int(rem - y * nx), int(y), int(z));
    }
  }
}

template <class T, int D>
__global__ __launch_bounds__(256) void k_batch_residual(BatchArgs b, double* partial) {
  __shared__ double red[4][2];
  for (uint32_t tile = blockIdx.x; tile < b.ntiles; tile += gridDim.x) {
    const Tile tl = b.tiles[tile];
    const StepArgs& a = b.members[tl.member];
    const T* u = static_cast<const T*>(a.u);
    const T* up = static_cast<const T*>(b.u_prev[tl.member]);
    double dsum = 0.0, usum = 0.0;
    for (uint32_t r = 0; r < b.rows; ++r) {
      const uint32_t c = tl.c0 + r * blockDim.x + threadIdx.x;
      if (c >= a.N) break;
      for (int d = 0; d < D; ++d) {
        const double v = double(u[size_t(d) * a.N + c]);
        const double w = double(up[size_t(d) * a.N + c]);
        dsum += (v - w) * (v - w);
        usum += v * v;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      dsum += __shfl_xor(dsum, o, 64);
      usum += __shfl_xor(usum, o, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
      red[wv][0] = dsum;
      red[wv][1] = usum;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      double s = 0.0;
      for (int w2 = 0; w2 < int((blockDim.x + 63) >> 6); ++w2) s += red[w2][threadIdx.x];
      partial[size_t(tile) * 2 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

template <class L, class S, class T, bool WB>
static hipError_t launch_batch_t(bool odd, const BatchArgs& b, const BatchShape& sh, hipStream_t s) {
  const uint32_t grid = sh.grid > 0 ? uint32_t(sh.grid) : b.ntiles;
  if (b.ntiles == 0) return hipSuccess;
  if (odd)
    hipLaunchKernelGGL((k_aa_batch<L, S, T, WB, true>), dim3(grid), dim3(sh.block), 0, s, b);
  else
    hipLaunchKernelGGL((k_aa_batch<L, S, T, WB, false>), dim3(grid), dim3(sh.block), 0, s, b);
  return hipGetLastError();
}

template <class L>
static hipError_t launch_batch_l(int storage, bool odd, const BatchArgs& b, const BatchShape& sh,
                                 hipStream_t s) {
  switch (storage) {
    case LBM_F64: return launch_batch_t<L, double, double, true>(odd, b, sh, s);
    case LBM_F32: return launch_batch_t<L, float, float, true>(odd, b, sh, s);
    case LBM_F16S: return launch_batch_t<L, __half, float, false>(odd, b, sh, s);
    case LBM_BF16S: return launch_batch_t<L, __hip_bfloat16, float, false>(odd, b, sh, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_aa_batch(int lattice, int storage, bool odd, const BatchArgs& b,
                           const BatchShape& sh, hipStream_t s) {
  switch (lattice) {
    case LBM_D2Q9: return launch_batch_l<D2Q9>(storage, odd, b, sh, s);
    case LBM_D3Q19: return launch_batch_l<D3Q19>(storage, odd, b, sh, s);
    case LBM_D3Q27: return launch_batch_l<D3Q27>(storage, odd, b, sh, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_residual(int dims, int storage, const BatchArgs& b, const BatchShape& sh,
                                 double* partial, const uint32_t* tile_begin, uint32_t members,
                                 double* out, hipStream_t s) {
  if (b.ntiles == 0) return hipSuccess;
  const uint32_t grid = sh.grid > 0 ? uint32_t(sh.grid) : b.ntiles;
  const bool dbl = storage == LBM_F64;
  if (dims == 2) {
    if (dbl) hipLaunchKernelGGL((k_batch_residual<double, 2>), dim3(grid), dim3(sh.block), 0, s, b, partial);
    else hipLaunchKernelGGL((k_batch_residual<float, 2>), dim3(grid), dim3(sh.block), 0, s, b, partial);
  } else {
    if (dbl) hipLaunchKernelGGL((k_batch_residual<double, 3>), dim3(grid), dim3(sh.block), 0, s, b, partial);
    else hipLaunchKernelGGL((k_batch_residual<float, 3>), dim3(grid), dim3(sh.block), 0, s, b, partial);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // per member, its tiles' partial sums in a fixed order (one block each)
  return launch_segment_sums(partial, tile_begin, members, out, s);
}

template <class L, class S, class T, bool WB>
static hipError_t launch_t(bool odd, const StepArgs& a, const LaunchShape& sh, hipStream_t s) {
  // 16-bit storage, even rows: two cells per thread (form 3; the default for
  // D2Q9 and D3Q19, D3Q27's 54 values per thread spill); one-row blocks of
  // sh.bx threads over the row's pairs
  if constexpr (sizeof(S) == 2) {
    const int form = sh.form[odd ? 1 : 0];
    if ((form == 3 || (form == 0 && L::Q <= 19)) && a.nx % 2 == 0) {
      const int pairs = a.nx / 2;
      int bx = 64;
      while (bx < pairs && bx < sh.bx) bx *= 2;
      dim3 block(bx, 1, 1);
      dim3 grid((pairs + bx - 1) / bx, a.ny, a.nz);
      if (odd)
        hipLaunchKernelGGL((k_aa_pair<L, S, true>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((k_aa_pair<L, S, false>), grid, block, 0, s, a);
      return hipGetLastError();
    }
  }
  if (a.nx >= kRowKernelMinNx) {
    // whole rows per block (up to 1024 cells), rows along grid y and z
    int bx = 64;
    while (bx < a.nx && bx < kRowBlock) bx *= 2;
    dim3 block(bx, 1, 1);
    dim3 grid((a.nx + bx - 1) / bx, a.ny, a.nz);
    if (odd)
      hipLaunchKernelGGL((k_aa_row<L, S, T, WB, true>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_aa_row<L, S, T, WB, false>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  dim3 block(sh.bx, sh.by, 1);
  dim3 grid((a.nx + sh.bx - 1) / sh.bx, (a.ny + sh.by - 1) / sh.by, a.nz);
  // the buffer-resource form where the population array's bytes fit the 32-bit range
  const int form = sh.form[odd ? 1 : 0];
  const bool want_buf = form == 0 ? (buf_phases<L, S>() & (odd ? 2 : 1)) != 0 : form == 2;
  const bool buf = want_buf && uint64_t(L::Q) * a.Np * sizeof(S) <= 0xffffffffull;
  // one-row blocks take the row-uniform addressing in the global form
  // (D3Q19 FP32 1%, D2Q9 FP64 10% faster); the buffer form keeps per-lane
  // rows (row-uniform scalar offsets are 4% slower in D3Q27 phase 1, 3% in
  // D2Q9 FP32)
  if (buf) {
    if (odd) hipLaunchKernelGGL((k_aa_step<L, S, T, WB, true, true, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_aa_step<L, S, T, WB, false, true, false>), grid, block, 0, s, a);
  } else if (sh.by == 1) {
    if (odd) hipLaunchKernelGGL((k_aa_step<L, S, T, WB, true, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_aa_step<L, S, T, WB, false, false, true>), grid, block, 0, s, a);
  } else {
    if (odd) hipLaunchKernelGGL((k_aa_step<L, S, T, WB, true, false, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_aa_step<L, S, T, WB, false, false, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

template <class L>
static hipError_t launch_l(int storage, bool odd, const StepArgs& a, const LaunchShape& sh,
                           hipStream_t s) {
  switch (storage) {
    case LBM_F64:
      if constexpr (kVecF64)
        if (sh.vec && a.nx % 2 == 0) return launch_vec<L, double, true, 2>(odd, a, s);
      return launch_t<L, double, double, true>(odd, a, sh, s);
    case LBM_F32:
      if constexpr (kVecF32 > 0)
        if (sh.vec && (!odd || kVecOdd) && a.nx % kVecF32 == 0)
          return launch_vec<L, float, true, (kVecF32 > 0 ? kVecF32 : 4)>(odd, a, s);
      return launch_t<L, float, float, true>(odd, a, sh, s);
    case LBM_F16S: return launch_t<L, __half, float, false>(odd, a, sh, s);
    case LBM_BF16S: return launch_t<L, __hip_bfloat16, float, false>(odd, a, sh, s);
  }
  return hipErrorInvalidValue;
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) {
  if (threadIdx.x == 0) *p = v;
}

hipError_t launch_set_u64(uint64_t* p, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

__global__ void k_add_u64(uint64_t* p, uint64_t v) {
  if (threadIdx.x == 0) *p = *p + v;
}

hipError_t launch_add_u64(uint64_t* p, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
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
