// Measured HBM ceilings of the device the engine runs on (lbm_hbm_ceiling).
//
// The roofline of every population kernel is HBM bandwidth.  Against the
// MI355X's published 8 TB/s, two plain data movements give the bandwidth a
// kernel can actually reach on the same device, under the same conditions:
//  - LBM_CEILING_COPY: b = a over two buffers, 16 bytes per lane;
//  - LBM_CEILING_PLANES: one in-place pass over 19 planes, one 4-byte
//    element of every plane per lane, read then written back nontemporally —
//    the access pattern of the D3Q19 AA even phase without its arithmetic.
// Both count the bytes read plus the bytes written per pass.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/lbm_mirror_amd/lbm.h"

lbm_status lbmx_report(lbm_status s, const char* msg);

namespace {

__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ a, float4* __restrict__ b,
                                                size_t n4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

constexpr int kPlanes = 19;

// src and dst are the same buffer; as two pointers the compiler cannot prove
// it, so it keeps every load and store (a value stored back to the address it
// was loaded from through one pointer would be a removable no-op)
__global__ __launch_bounds__(256) void k_planes_nt(const float* src, float* dst, size_t cells) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  float v[kPlanes];
#pragma unroll
  for (int i = 0; i < kPlanes; ++i) v[i] = __builtin_nontemporal_load(src + i * cells + c);
#pragma unroll
  for (int i = 0; i < kPlanes; ++i) __builtin_nontemporal_store(v[i], dst + i * cells + c);
}

lbm_status hip_fail(hipError_t e, const char* what) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return lbmx_report(e == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_CRITICAL, buf);
}

}  // namespace

extern "C" lbm_status lbm_hbm_ceiling(int32_t device, int32_t kind, uint64_t bytes, uint32_t reps,
                                      double* gbps) {
  if (!gbps) return lbmx_report(LBM_ERR_INVALID_ARG, "null argument");
  if (kind != LBM_CEILING_COPY && kind != LBM_CEILING_PLANES)
    return lbmx_report(LBM_ERR_INVALID_ARG, "unknown ceiling kind");
  if (reps == 0 || bytes < (uint64_t(1) << 20) || bytes > (uint64_t(64) << 30))
    return lbmx_report(LBM_ERR_INVALID_ARG, "bytes must be in [1 MiB, 64 GiB] and reps > 0");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  // whole 16-byte vectors (copy) or whole cells of 19 planes (in place); the
  // plane length a multiple of 256 cells, so every plane starts 1 KiB-aligned
  // (as the lattice's padded planes do)
  const size_t unit = kind == LBM_CEILING_COPY ? 16 : 4 * kPlanes;
  const size_t n = kind == LBM_CEILING_COPY ? bytes / unit : bytes / unit / 256 * 256;
  const size_t used = n * unit;
  float *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  lbm_status st = LBM_OK;
  auto cleanup = [&]() {
    if (s) (void)hipStreamSynchronize(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  };
#define CEIL_TRY(expr, what)        \
  do {                              \
    hipError_t e_ = (expr);         \
    if (e_ != hipSuccess) {         \
      st = hip_fail(e_, what);      \
      cleanup();                    \
      return st;                    \
    }                               \
  } while (0)
  CEIL_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  CEIL_TRY(hipEventCreate(&e0), "hipEventCreate");
  CEIL_TRY(hipEventCreate(&e1), "hipEventCreate");
  CEIL_TRY(hipMalloc(&a, used), "hipMalloc");
  CEIL_TRY(hipMemsetAsync(a, 0, used, s), "hipMemsetAsync");
  if (kind == LBM_CEILING_COPY) {
    CEIL_TRY(hipMalloc(&b, used), "hipMalloc");
    CEIL_TRY(hipMemsetAsync(b, 0, used, s), "hipMemsetAsync");
  }
  const unsigned grid = unsigned((n + 255) / 256);
  auto launch = [&]() {
    if (kind == LBM_CEILING_COPY)
      hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (const float4*)a, (float4*)b, n);
    else
      hipLaunchKernelGGL(k_planes_nt, dim3(grid), dim3(256), 0, s, a, a, n);
    return hipGetLastError();
  };
  CEIL_TRY(launch(), "launch");  // untimed first pass
  CEIL_TRY(hipEventRecord(e0, s), "hipEventRecord");
  for (uint32_t r = 0; r < reps; ++r) CEIL_TRY(launch(), "launch");
  CEIL_TRY(hipEventRecord(e1, s), "hipEventRecord");
  CEIL_TRY(hipEventSynchronize(e1), "hipEventSynchronize");
  float ms = 0.f;
  CEIL_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
#undef CEIL_TRY
  *gbps = 2.0 * double(used) * reps / (double(ms) * 1e-3) / 1e9;
  cleanup();
  return LBM_OK;
}
