// Launch interface between the C ABI (lbm_abi.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cell_ops.hpp"

namespace lbmx {

constexpr int kMaxOps = 16;  // flags 0..15 may carry an operator

// Everything one stream+collide launch needs.  Field pointers are in the
// compute type T of the storage mode (double for F64, float otherwise).
struct StepArgs {
  void* f;               // Q planes of Np populations, storage type
  const uint8_t* flags;  // N cell flags
  const OpDev* ops;      // kMaxOps device operators
  void* rho;             // N
  void* u;               // D*N
  void* force;           // D*N
  void* por;             // N
  const void* bcvel;     // D*N
  int32_t nx, ny, nz;
  uint32_t N, Np;
  uint32_t periodic;
  uint64_t t;
  int32_t write_macros;
};

struct LaunchShape {
  int bx = 256, by = 1;  // threads per block along x and y (scalar kernel)
  bool vec = true;       // vectorised kernel where the row length allows it
};

hipError_t launch_aa_step(int lattice, int storage, bool odd, const StepArgs& a,
                          const LaunchShape& shape, hipStream_t s);

// Particle coupling (particles.hip)
struct ParticleArgs {
  // lattice and fields
  const void* f;
  void* por;
  void* force;
  int32_t nx, ny, nz;
  uint32_t N, Np;
  uint32_t periodic;
  int32_t phase;  // AA phase of f: 0 natural, 1 swapped
  // particles (T-typed SoA records of kPartFields values each)
  void* parts;
  uint32_t n;
  double* f64;      // n*3 force sums (FP64)
  double* t64;      // n*3 torque sums (FP64)
  double* partial;  // per-block partial sums
  uint32_t blocks_per_particle;
  // parameters
  double eps;
  uint64_t sub_steps;
  int32_t porosity_mode;
  int32_t psm_torque;
  double k_n, zeta_n, gamma_t, mu;
  int32_t walls;
  double g[3];
  // cell list
  uint64_t* cl_count;
  uint64_t* cl_slot;
  uint64_t* cl_overflow;
  uint32_t* owner;  // per-cell owner (scatter mode)
  double bin;       // bin size (T-rounded)
  uint64_t grid[3];
  uint32_t cap;
};

// Per-particle record layout (T values)
enum PartField {
  PF_POS = 0, PF_POS_OLD = 3, PF_ROT = 6, PF_ROT_OLD = 9, PF_ACC = 12, PF_ANG_ACC = 15,
  PF_FORCE = 18, PF_TORQUE = 21, PF_MASS = 24, PF_INERTIA = 25, PF_DENSITY = 26,
  PF_RADIUS = 27, kPartFields = 28
};

hipError_t launch_psm_couple(int lattice, int storage, const ParticleArgs& a, hipStream_t s);
hipError_t launch_dem_contact(int dims, int storage, const ParticleArgs& a, hipStream_t s);
hipError_t launch_particle_integrate(int dims, int storage, const ParticleArgs& a, hipStream_t s);

}  // namespace lbmx
