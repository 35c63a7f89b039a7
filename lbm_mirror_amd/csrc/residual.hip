// FP64 steady-state residual sums (no reference counterpart).
//
// A lattice's residual is r = |u_t - u_{t-1}|_2 / |u_t|_2 over every cell and
// velocity component.  The sums are formed in FP64 in a fixed order, so the
// same fields give the same bits on every call: block b adds the cells
// [2048 b, 2048 (b+1)) (thread t the cells 2048 b + t + 256 k, k = 0..7, in k
// order; then the 64 lanes of a wavefront by a shuffle butterfly, then the
// four wavefronts in order), and one block per segment adds the block sums
// (thread t the sums t, t + 256, ... in order; then the same tree).
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace lbmx {

namespace {

// the block's 256 values summed in a fixed order; every thread gets the total
__device__ __forceinline__ double block_sum(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) lds[wv] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = int((blockDim.x + 63) >> 6);
  for (int w = 0; w < nw; ++w) s += lds[w];
  __syncthreads();
  return s;
}

template <class T, int D>
__global__ __launch_bounds__(256) void k_residual_blocks(const T* __restrict__ u, const T* __restrict__ up,
                                                         uint32_t N, double* __restrict__ partial) {
  __shared__ double lds[4];
  const uint32_t base = blockIdx.x * kResidualCellsPerBlock;
  double ds = 0.0, us = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kResidualCellsPerBlock / 256; ++k) {
    const uint32_t c = base + threadIdx.x + 256u * k;
    if (c < N) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double v = double(u[size_t(d) * N + c]);
        const double w = double(up[size_t(d) * N + c]);
        ds += (v - w) * (v - w);
        us += v * v;
      }
    }
  }
  ds = block_sum(ds, lds);
  us = block_sum(us, lds);
  if (threadIdx.x == 0) {
    partial[2 * size_t(blockIdx.x)] = ds;
    partial[2 * size_t(blockIdx.x) + 1] = us;
  }
}

// One block per segment [begin[m], begin[m+1]) of partial pairs (begin null:
// one segment [0, count)).  RATIO: out[3m..3m+2] = (r, Σ|Δu|², Σ|u|²), else
// out[2m..2m+1] = (Σ|Δu|², Σ|u|²).
template <bool RATIO>
__global__ __launch_bounds__(256) void k_segment_sums(const double* __restrict__ partial,
                                                      const uint32_t* __restrict__ begin, uint32_t count,
                                                      double* __restrict__ out) {
  __shared__ double lds[4];
  const uint32_t m = blockIdx.x;
  const uint32_t b0 = begin ? begin[m] : 0u, b1 = begin ? begin[m + 1] : count;
  double ds = 0.0, us = 0.0;
  for (uint32_t k = b0 + threadIdx.x; k < b1; k += blockDim.x) {
    ds += partial[2 * size_t(k)];
    us += partial[2 * size_t(k) + 1];
  }
  ds = block_sum(ds, lds);
  us = block_sum(us, lds);
  if (threadIdx.x == 0) {
    if (RATIO) {
      const bool finite = isfinite(ds) && isfinite(us);
      out[3 * size_t(m)] = !finite ? __builtin_nan("") : (us > 0.0 ? sqrt(ds / us) : 0.0);
      out[3 * size_t(m) + 1] = ds;
      out[3 * size_t(m) + 2] = us;
    } else {
      out[2 * size_t(m)] = ds;
      out[2 * size_t(m) + 1] = us;
    }
  }
}

}  // namespace

hipError_t launch_residual(int dims, bool dbl, const void* u, const void* up, uint32_t N, double* partial,
                           double* out, hipStream_t s) {
  const uint32_t nb = residual_blocks(N);
  if (nb == 0) return hipErrorInvalidValue;
  if (dims == 2) {
    if (dbl)
      hipLaunchKernelGGL((k_residual_blocks<double, 2>), dim3(nb), dim3(256), 0, s, (const double*)u,
                         (const double*)up, N, partial);
    else
      hipLaunchKernelGGL((k_residual_blocks<float, 2>), dim3(nb), dim3(256), 0, s, (const float*)u,
                         (const float*)up, N, partial);
  } else {
    if (dbl)
      hipLaunchKernelGGL((k_residual_blocks<double, 3>), dim3(nb), dim3(256), 0, s, (const double*)u,
                         (const double*)up, N, partial);
    else
      hipLaunchKernelGGL((k_residual_blocks<float, 3>), dim3(nb), dim3(256), 0, s, (const float*)u,
                         (const float*)up, N, partial);
  }
  hipLaunchKernelGGL((k_segment_sums<true>), dim3(1), dim3(256), 0, s, (const double*)partial,
                     (const uint32_t*)nullptr, nb, out);
  return hipGetLastError();
}

hipError_t launch_segment_sums(const double* partial, const uint32_t* begin, uint32_t segments, double* out,
                               hipStream_t s) {
  if (segments == 0) return hipSuccess;
  hipLaunchKernelGGL((k_segment_sums<false>), dim3(segments), dim3(256), 0, s, partial, begin, 0u, out);
  return hipGetLastError();
}

}  // namespace lbmx
