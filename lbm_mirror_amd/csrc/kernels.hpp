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
  const void* pn;        // D*N HLBM particle velocity sums N
  const void* pd;        // N   HLBM particle weights D
  int32_t nx, ny, nz;
  uint32_t N, Np;
  uint32_t periodic;
  uint64_t t;
  const uint64_t* t_dev;  // null: t is the step index; else the index is *t_dev + t
  int32_t write_macros;
  // fault words the host reads without a copy (pinned, mapped host memory):
  // [kStatusNonFinite] set to 1 when a macro-writing pass meets a non-finite
  // density or velocity; [kStatusOverflow] when a cell-list bin was full
  uint32_t* status;
};

constexpr int kStatusNonFinite = 0;
constexpr int kStatusOverflow = 1;

struct LaunchShape {
  int bx = 256, by = 1;  // threads per block along x and y (scalar kernel)
  bool vec = true;       // vectorised kernel where the row length allows it
  // population addressing per AA phase: 0 the lattice's default, 1 global
  // addresses, 2 one buffer resource (used only where the array's bytes fit
  // its 32-bit range), 3 two cells per thread as 32-bit words (16-bit
  // storage, even rows)
  int32_t form[2] = {0, 0};
};

hipError_t launch_aa_step(int lattice, int storage, bool odd, const StepArgs& a,
                          const LaunchShape& shape, hipStream_t s);

// *p = v and *p += v, in stream order (a step index that captured launches read)
hipError_t launch_set_u64(uint64_t* p, uint64_t v, hipStream_t s);
hipError_t launch_add_u64(uint64_t* p, uint64_t v, hipStream_t s);

// ---- batched ensembles -------------------------------------------------------
// A tile is `rows` consecutive rows (from y) of one member, blockDim cells wide
// from x0.
struct Tile {
  uint32_t member, c0;  // first cell (natural order) of the tile within the member
};

struct BatchArgs {
  const StepArgs* members;  // device array; t and write_macros are set per launch
  const Tile* tiles;
  uint32_t ntiles;
  uint32_t rows;            // rows per tile
  const uint64_t* t0;       // device: step index of the first step of a chain
  uint32_t k;               // this launch's offset in the chain
  int32_t write_macros;     // 0 none, 1 into the members' velocity, 2 into u_prev
  void* const* u_prev;      // per member: second velocity field for the residual
};

struct BatchShape {
  int block = 256;          // threads per block (cells per tile row)
  int rows = 1;             // rows per tile
  int grid = 0;             // persistent blocks (0: one per tile)
};

hipError_t launch_aa_batch(int lattice, int storage, bool odd, const BatchArgs& b,
                           const BatchShape& shape, hipStream_t s);

// Per member: Σ|u − u_prev|² and Σ|u|² in FP64 over every cell, tile partials
// reduced in a fixed order.  partial: 2 doubles per tile; out: 2 per member;
// tile_begin: ntiles-of-member prefix (members+1 entries).
hipError_t launch_batch_residual(int dims, int storage, const BatchArgs& b, const BatchShape& shape,
                                 double* partial, const uint32_t* tile_begin, uint32_t members,
                                 double* out, hipStream_t s);

// ---- residual sums (residual.hip) ---------------------------------------------
// FP64 Σ|u − u_prev|² and Σ|u|² of one lattice in a fixed order: one partial
// pair per block of kResidualCellsPerBlock cells, then out[0..2] = (r, Σ|Δu|²,
// Σ|u|²) with r = sqrt(Σ|Δu|²/Σ|u|²) (0 when Σ|u|² = 0, NaN when a sum is not
// finite).  partial holds 2 * residual_blocks(N) doubles.
constexpr uint32_t kResidualCellsPerBlock = 2048;
inline uint32_t residual_blocks(uint32_t N) {
  return (N + kResidualCellsPerBlock - 1) / kResidualCellsPerBlock;
}
hipError_t launch_residual(int dims, bool dbl, const void* u, const void* up, uint32_t N, double* partial,
                           double* out, hipStream_t s);
// out[2m..2m+1] = the partial pairs [begin[m], begin[m+1]) summed in a fixed
// order, one block per segment
hipError_t launch_segment_sums(const double* partial, const uint32_t* begin, uint32_t segments, double* out,
                               hipStream_t s);

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
  uint32_t* owner;  // per-cell owners (more than kLocalOwnerMax particles), else null
  double bin;       // bin size (T-rounded)
  uint64_t grid[3];
  uint32_t cap;
  int32_t state64;  // particle records and arithmetic in FP64 whatever the storage
  int32_t hlbm;     // HLBM coupling: raw momentum as cell force, torque, stamp of N and D
  int32_t fplbm;    // FPLBM coupling: penalty force (u_s - m)(1 - B), no torque
  const uint8_t* cflags;  // cell flags (FPLBM: flag-2 cells take m from the populations)
  const void* umac;       // D*N macro velocity field (FPLBM: m of the other cells)
  void* pn;         // D*N HLBM N sums (fluid field type)
  void* pd;         // N   HLBM D sums
  int32_t fuse_finalize;  // the contact kernel finalizes the gather sums (full steps)
  int32_t fuse_integrate; // the contact kernel integrates its particle (full steps):
                          // partners are read from snap, written by the gather
  void* snap;             // n records: the particles as the gather saw them
  int32_t fixed_axis;  // ParticleAxisIntegrate instead of ParticleIntegrate
  double axis[3];      // its normalised axis, T-rounded (2D: the sign in [0])
  uint32_t* status;    // the context's fault words (StepArgs::status)
  // PSM scatter porosity without the per-cell reset: n*12 ints, particle p's
  // box of this gather at [12p, 12p+6) (lo, hi) and of the previous one at
  // [12p+6, 12p+12).  The gather sweeps the union of the two and sets the
  // cells of the previous box that no current box holds back to 1 (null:
  // every cell is reset before the gather)
  int32_t* boxes;
};

// Up to this many particles the gather decides cell ownership per block (no
// per-cell owner array, no owner pass); above it the owner pass runs.
constexpr uint32_t kLocalOwnerMax = 1024;

// Per-particle record layout (T values)
enum PartField {
  PF_POS = 0, PF_POS_OLD = 3, PF_ROT = 6, PF_ROT_OLD = 9, PF_ACC = 12, PF_ANG_ACC = 15,
  PF_FORCE = 18, PF_TORQUE = 21, PF_MASS = 24, PF_INERTIA = 25, PF_DENSITY = 26,
  PF_RADIUS = 27, kPartFields = 28
};

hipError_t launch_psm_couple(int lattice, int storage, const ParticleArgs& a, hipStream_t s);
// Porosity 1 in every cell and every particle's previous box empty: the
// state from which the gathers keep the porosity without a per-cell reset
hipError_t launch_sweep_reset(bool dbl, void* por, uint32_t N, int32_t* boxes, uint32_t n, hipStream_t s);
// HLBM stamp after the move: N += u_p (1 - B), D += 1 - B over each post-move AABB
hipError_t launch_hlbm_stamp(int dims, int storage, const ParticleArgs& a, hipStream_t s);
hipError_t launch_dem_contact(int dims, int storage, const ParticleArgs& a, hipStream_t s);
hipError_t launch_particle_integrate(int dims, int storage, const ParticleArgs& a, hipStream_t s);

}  // namespace lbmx
