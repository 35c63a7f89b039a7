// Field transfer kernels (fields.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "lattice.hpp"
#include "../../include/lbm_mirror_amd/lbm.h"

namespace lbmx {

struct Geometry {
  int32_t nx, ny, nz;
  uint32_t N, Np;
  uint32_t periodic;
};

hipError_t unpack_populations(int lattice, int storage, const void* f, double* out,
                              const Geometry& g, int phase, hipStream_t s);
hipError_t pack_populations(int lattice, int storage, void* f, const double* in,
                            const Geometry& g, hipStream_t s);
hipError_t field_to_double(bool is_double, const void* a, double* b, size_t n, hipStream_t s);
hipError_t field_from_double(bool is_double, void* a, const double* b, size_t n, hipStream_t s);

}  // namespace lbmx
