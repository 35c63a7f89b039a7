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

// host-side population arrays in `layout` (lbm_layout): LBM_LAYOUT_SOA or the
// reference's chunk layout with its ghost ring (LBM_LAYOUT_REF_AOS_GHOST)
size_t population_count(int lattice, const Geometry& g, int layout);
hipError_t unpack_populations(int lattice, int storage, const void* f, double* out,
                              const Geometry& g, int phase, hipStream_t s, int layout = LBM_LAYOUT_SOA);
hipError_t pack_populations(int lattice, int storage, void* f, const double* in,
                            const Geometry& g, hipStream_t s, int layout = LBM_LAYOUT_SOA);
hipError_t field_to_double(bool is_double, const void* a, double* b, size_t n, hipStream_t s);
hipError_t field_from_double(bool is_double, void* a, const double* b, size_t n, hipStream_t s);

}  // namespace lbmx
