// Field transfer kernels: populations between the caller's natural SoA
// doubles and the device layout (storage type, AA phase), and T <-> double
// conversion of per-cell fields.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "fields.hpp"

namespace lbmx {

template <class S>
__device__ inline double to_d(S v);
template <>
__device__ inline double to_d<double>(double v) { return v; }
template <>
__device__ inline double to_d<float>(float v) { return double(v); }
template <>
__device__ inline double to_d<__half>(__half v) { return double(__half2float(v)); }
template <>
__device__ inline double to_d<__hip_bfloat16>(__hip_bfloat16 v) {
  return double(__bfloat162float(v));
}
template <class S>
__device__ inline S from_d(double v);
template <>
__device__ inline double from_d<double>(double v) { return v; }
template <>
__device__ inline float from_d<float>(double v) { return float(v); }
template <>
__device__ inline __half from_d<__half>(double v) { return __float2half(float(v)); }
template <>
__device__ inline __hip_bfloat16 from_d<__hip_bfloat16>(double v) {
  return __float2bfloat16(float(v));
}

template <class S>
struct Shift {  // shifted storage keeps f − w
  static constexpr bool on = false;
};
template <>
struct Shift<__half> {
  static constexpr bool on = true;
};
template <>
struct Shift<__hip_bfloat16> {
  static constexpr bool on = true;
};

// Index of population i of cell (x, y, z) in a host-side population array:
// SoA doubles [i*N + cell], or the reference's chunk layout (chunk.hpp:
// 11-24, flatten.hpp:11-39): a ghost ring of one cell on every side, cells
// x-fastest over (nx+2)(ny+2)[(nz+2)], the Q populations of a cell
// consecutive (FixedArray<T,Q> per cell).
template <int D, bool AOS>
__device__ inline size_t pop_index(const Geometry& g, int x, int y, int z, uint32_t cell, int i, int Q) {
  if (!AOS) return size_t(i) * g.N + cell;
  const size_t gx = size_t(g.nx) + 2, gy = size_t(g.ny) + 2;
  const size_t gz = D == 3 ? size_t(z) + 1 : 0;
  return ((size_t(x) + 1) + gx * ((size_t(y) + 1) + gy * gz)) * size_t(Q) + size_t(i);
}

// logical f_i(x) of the current state, natural SoA doubles out[i*N + cell]
// (or the chunk layout; its ghost ring is left as the caller cleared it)
template <class L, class S, bool AOS>
__global__ void k_unpack(const S* __restrict__ f, double* __restrict__ out, Geometry g, int phase) {
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.N) return;
  const int x = cell % g.nx, y = (cell / g.nx) % g.ny, z = cell / (uint32_t(g.nx) * g.ny);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const float w = float(L::w[i]);
    double v;
    if (phase == 0) {
      v = to_d<S>(f[size_t(i) * g.Np + cell]);
      if (Shift<S>::on) v = double(float(v) + w);
    } else {
      int xs = x - L::cx[i], ys = y - L::cy[i], zs = z - L::cz[i];
      bool ok = true;
      if (xs < 0 || xs >= g.nx) { if (g.periodic & 1u) xs = (xs + g.nx) % g.nx; else ok = false; }
      if (ys < 0 || ys >= g.ny) { if (g.periodic & 2u) ys = (ys + g.ny) % g.ny; else ok = false; }
      if (zs < 0 || zs >= g.nz) { if (g.periodic & 4u) zs = (zs + g.nz) % g.nz; else ok = false; }
      if (ok) {
        const uint32_t src = uint32_t(xs) + uint32_t(g.nx) * (uint32_t(ys) + uint32_t(g.ny) * uint32_t(zs));
        v = to_d<S>(f[size_t(L::opp[i]) * g.Np + src]);
        if (Shift<S>::on) v = double(float(v) + w);
      } else {
        v = 0.0;
      }
    }
    out[pop_index<L::D, AOS>(g, x, y, z, cell, i, L::Q)] = v;
  }
}

// natural layout from doubles (phase 0); the chunk layout's ghost ring is not read
template <class L, class S, bool AOS>
__global__ void k_pack(S* __restrict__ f, const double* __restrict__ in, Geometry g) {
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.N) return;
  const int x = cell % g.nx, y = (cell / g.nx) % g.ny, z = cell / (uint32_t(g.nx) * g.ny);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    double v = in[pop_index<L::D, AOS>(g, x, y, z, cell, i, L::Q)];
    if (Shift<S>::on) v = double(float(v) - float(L::w[i]));
    f[size_t(i) * g.Np + cell] = from_d<S>(v);
  }
}

template <class T>
__global__ void k_to_double(const T* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = double(a[i]);
}
template <class T>
__global__ void k_from_double(T* __restrict__ a, const double* __restrict__ b, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = T(b[i]);
}

template <class L, bool AOS>
static hipError_t unpack_l(int storage, const void* f, double* out, const Geometry& g, int phase,
                           hipStream_t s) {
  dim3 b(256), gr((g.N + 255) / 256);
  switch (storage) {
    case LBM_F64: hipLaunchKernelGGL((k_unpack<L, double, AOS>), gr, b, 0, s, (const double*)f, out, g, phase); break;
    case LBM_F32: hipLaunchKernelGGL((k_unpack<L, float, AOS>), gr, b, 0, s, (const float*)f, out, g, phase); break;
    case LBM_F16S: hipLaunchKernelGGL((k_unpack<L, __half, AOS>), gr, b, 0, s, (const __half*)f, out, g, phase); break;
    case LBM_BF16S: hipLaunchKernelGGL((k_unpack<L, __hip_bfloat16, AOS>), gr, b, 0, s, (const __hip_bfloat16*)f, out, g, phase); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <class L, bool AOS>
static hipError_t pack_l(int storage, void* f, const double* in, const Geometry& g, hipStream_t s) {
  dim3 b(256), gr((g.N + 255) / 256);
  switch (storage) {
    case LBM_F64: hipLaunchKernelGGL((k_pack<L, double, AOS>), gr, b, 0, s, (double*)f, in, g); break;
    case LBM_F32: hipLaunchKernelGGL((k_pack<L, float, AOS>), gr, b, 0, s, (float*)f, in, g); break;
    case LBM_F16S: hipLaunchKernelGGL((k_pack<L, __half, AOS>), gr, b, 0, s, (__half*)f, in, g); break;
    case LBM_BF16S: hipLaunchKernelGGL((k_pack<L, __hip_bfloat16, AOS>), gr, b, 0, s, (__hip_bfloat16*)f, in, g); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t population_count(int lattice, const Geometry& g, int layout) {
  const int Q = lattice == LBM_D2Q9 ? 9 : (lattice == LBM_D3Q19 ? 19 : 27);
  if (layout != LBM_LAYOUT_REF_AOS_GHOST) return size_t(Q) * g.N;
  size_t cells = (size_t(g.nx) + 2) * (size_t(g.ny) + 2);
  if (lattice != LBM_D2Q9) cells *= size_t(g.nz) + 2;
  return cells * size_t(Q);
}

hipError_t unpack_populations(int lattice, int storage, const void* f, double* out,
                              const Geometry& g, int phase, hipStream_t s, int layout) {
  const bool aos = layout == LBM_LAYOUT_REF_AOS_GHOST;
  switch (lattice) {
    case LBM_D2Q9: return aos ? unpack_l<D2Q9, true>(storage, f, out, g, phase, s) : unpack_l<D2Q9, false>(storage, f, out, g, phase, s);
    case LBM_D3Q19: return aos ? unpack_l<D3Q19, true>(storage, f, out, g, phase, s) : unpack_l<D3Q19, false>(storage, f, out, g, phase, s);
    case LBM_D3Q27: return aos ? unpack_l<D3Q27, true>(storage, f, out, g, phase, s) : unpack_l<D3Q27, false>(storage, f, out, g, phase, s);
  }
  return hipErrorInvalidValue;
}
hipError_t pack_populations(int lattice, int storage, void* f, const double* in, const Geometry& g,
                            hipStream_t s, int layout) {
  const bool aos = layout == LBM_LAYOUT_REF_AOS_GHOST;
  switch (lattice) {
    case LBM_D2Q9: return aos ? pack_l<D2Q9, true>(storage, f, in, g, s) : pack_l<D2Q9, false>(storage, f, in, g, s);
    case LBM_D3Q19: return aos ? pack_l<D3Q19, true>(storage, f, in, g, s) : pack_l<D3Q19, false>(storage, f, in, g, s);
    case LBM_D3Q27: return aos ? pack_l<D3Q27, true>(storage, f, in, g, s) : pack_l<D3Q27, false>(storage, f, in, g, s);
  }
  return hipErrorInvalidValue;
}

hipError_t field_to_double(bool is_double, const void* a, double* b, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  dim3 bl(256), gr(unsigned((n + 255) / 256));
  if (is_double)
    hipLaunchKernelGGL((k_to_double<double>), gr, bl, 0, s, (const double*)a, b, n);
  else
    hipLaunchKernelGGL((k_to_double<float>), gr, bl, 0, s, (const float*)a, b, n);
  return hipGetLastError();
}
hipError_t field_from_double(bool is_double, void* a, const double* b, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  dim3 bl(256), gr(unsigned((n + 255) / 256));
  if (is_double)
    hipLaunchKernelGGL((k_from_double<double>), gr, bl, 0, s, (double*)a, b, n);
  else
    hipLaunchKernelGGL((k_from_double<float>), gr, bl, 0, s, (float*)a, b, n);
  return hipGetLastError();
}

}  // namespace lbmx
