// copy_probe — measured HBM ceiling on this GPU and the byte-counter
// calibration workload: one plain float copy of a known size, repeated.
// Usage: copy_probe [MiB per buffer] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

__global__ void k_copy1(const float* __restrict__ a, float* __restrict__ b, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// The LBM-shaped ceiling: Q planes read and Q planes written per cell, one
// cell per lane (the even AA phase without arithmetic).
template <int Q, bool NT>
__global__ void k_planes(const float* __restrict__ a, float* __restrict__ b, size_t n, size_t np) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float v[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) v[i] = NT ? __builtin_nontemporal_load(a + i * np + c) : a[i * np + c];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (NT) __builtin_nontemporal_store(v[i], b + i * np + c);
    else b[i * np + c] = v[i];
  }
}

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e = (x);                                                         \
    if (e != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));               \
      return 1;                                                                 \
    }                                                                           \
  } while (0)

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 640;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(float);
  float *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (int variant = 0; variant < 2; ++variant) {
    auto launch = [&]() {
      if (variant == 0)
        hipLaunchKernelGGL(k_copy4, dim3(unsigned((n / 4 + 255) / 256)), dim3(256), 0, 0,
                           (const float4*)a, (float4*)b, n / 4);
      else
        hipLaunchKernelGGL(k_copy1, dim3(unsigned((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
    };
    launch();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double gbs = 2.0 * double(bytes) * reps / (ms * 1e-3) / 1e9;
    std::printf("{\"copy\": \"%s\", \"bytes_per_launch\": %zu, \"ms_per_launch\": %.5f, \"GBps\": %.1f}\n",
                variant == 0 ? "float4" : "float", 2 * bytes, ms / reps, gbs);
  }
  {  // 19 planes of 8,388,608 cells (the D3Q19 512x128x128 lattice), in place
    const size_t cells = size_t(512) * 128 * 128;
    const size_t np = cells;
    const size_t need = 19 * np * sizeof(float);
    float* p = nullptr;
    CK(hipMalloc(&p, need));
    CK(hipMemset(p, 0, need));
    for (int nt = 0; nt < 2; ++nt) {
      auto launch = [&]() {
        if (nt) hipLaunchKernelGGL((k_planes<19, true>), dim3(unsigned((cells + 255) / 256)), dim3(256), 0, 0, p, p, cells, np);
        else hipLaunchKernelGGL((k_planes<19, false>), dim3(unsigned((cells + 255) / 256)), dim3(256), 0, 0, p, p, cells, np);
      };
      launch();
      CK(hipDeviceSynchronize());
      CK(hipEventRecord(e0, 0));
      for (int r = 0; r < reps; ++r) launch();
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      const double bytes2 = 2.0 * double(need);
      std::printf("{\"copy\": \"19-plane in place%s\", \"bytes_per_launch\": %.0f, \"ms_per_launch\": %.5f, \"GBps\": %.1f}\n",
                  nt ? " nontemporal" : "", bytes2, ms / reps, bytes2 * reps / (ms * 1e-3) / 1e9);
    }
    CK(hipFree(p));
  }
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
