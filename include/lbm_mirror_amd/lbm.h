/*
 * lbm_mirror_amd — C ABI of the MI355X (gfx950) lattice Boltzmann engine.
 *
 * Drop-in boundary for the reference's component API (keldu/lbm_mirror,
 * modules/core + modules/sycl).  The reference builds
 * `component<T, Desc, cmpt::Kind, encode::Sycl>` objects on the host, captures
 * them by value in a `parallel_for` lambda and calls `apply(field, macros,
 * index, t)` once per cell (examples/poiseulle_2d_gpu/poiseulle_2d_gpu.cpp:
 * 158-235).  A per-cell `apply` cannot cross a C ABI, so the same component
 * kinds and parameters are handed over at field granularity: a table maps
 * each cell flag to (boundary component, collision component), exactly the
 * `switch(info)` of the reference drivers, and one call advances whole steps.
 *
 * Conventions shared by every entry point
 *  - Every function returns an lbm_status (0 = ok).  Codes map 1:1 onto the
 *    reference error categories (modules/core/c++/abstract/error.hpp:209-259).
 *    No C++ exception crosses this boundary.  lbm_last_error() returns text.
 *  - Host buffers are owned by the caller, copied during the call, never kept.
 *  - Cells are x-fastest: cell = x + nx*(y + ny*z), the reference's
 *    flatten_index order (modules/core/c++/flatten.hpp:11-39) without its
 *    ghost ring.  The ghost ring (flag 0, populations 0) is implicit.
 *  - Population arrays at the boundary are SoA doubles f[i*N + cell], i in
 *    the reference's direction order (descriptor/d2q9.hpp:19-29,
 *    d3q19.hpp:9-40, d3q27.hpp:17-48).
 *  - One context owns one HIP stream; calls are stream-ordered and become
 *    host-visible at the get_* calls or lbm_sync.  A context is not thread
 *    safe: one host thread per context.
 */
#ifndef LBM_MIRROR_AMD_LBM_H
#define LBM_MIRROR_AMD_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: abstract/error.hpp:209-259 categories ---------------- */
typedef enum lbm_status {
  LBM_OK = 0,
  LBM_ERR_CRITICAL = 1,       /* err::critical — HIP runtime failure        */
  LBM_ERR_NOT_FOUND = 2,      /* err::not_found                             */
  LBM_ERR_OUT_OF_MEMORY = 3,  /* err::out_of_memory — hipMalloc failed      */
  LBM_ERR_INVALID_STATE = 4,  /* err::invalid_state — call order / NaN      */
  LBM_ERR_NOT_SUPPORTED = 5,  /* err::not_supported                         */
  LBM_ERR_INVALID_ARG = 6,    /* err::invalid_argument                      */
  LBM_ERR_OVERFLOW = 7        /* err::buffer_exhausted: a cell-list bin was
                                 full (cell_list.hpp:96-110)                */
} lbm_status;

/* ---- lattices: descriptor/{d2q9,d3q19,d3q27}.hpp ------------------------ */
typedef enum lbm_lattice { LBM_D2Q9 = 0, LBM_D3Q19 = 1, LBM_D3Q27 = 2 } lbm_lattice;

/* ---- storage / arithmetic -------------------------------------------------
 * F64, F32: populations stored and computed in T, with the reference's
 *   per-operation rounding (abstract/data.hpp:170-188) and its FP32
 *   equilibrium bracket evaluated in double (equilibrium.hpp:36-44).
 * F16S, BF16S: populations stored as the deviation f_i - w_i in 16 bits,
 *   arithmetic in FP32 (no reference counterpart; the reference's
 *   MixedPrecision schema, abstract/schema.hpp:95-112, is the precedent).   */
typedef enum lbm_storage { LBM_F64 = 0, LBM_F32 = 1, LBM_F16S = 2, LBM_BF16S = 3 } lbm_storage;

/* ---- per-flag operators: the driver switch(info) --------------------------
 * boundary runs first, then collision, on the pre-collision populations of
 * the step; every cell whose flag is > 0 is streamed.                      */
typedef enum lbm_boundary_kind {
  LBM_BC_NONE = 0,
  LBM_BC_BOUNCE_BACK = 1,       /* cmpt::BounceBack        boundary.hpp:57-75   */
  LBM_BC_ANTI_BOUNCE_BACK = 2,  /* cmpt::AntiBounceBack<0> boundary.hpp:106-132 (D2Q9) */
  LBM_BC_ZOUHE_PRESSURE_X = 3,  /* cmpt::ZouHePressureX<east> boundary.hpp:175-227 (D2Q9), 229-756 (D3Q27) */
  LBM_BC_ZOUHE_VELOCITY_X = 4,  /* cmpt::ZouHeVelocityX<east> boundary.hpp:759-805 (D2Q9) */
  LBM_BC_EQUILIBRIUM = 5        /* cmpt::Equilibrium       boundary.hpp:135-161 */
} lbm_boundary_kind;

typedef enum lbm_collision_kind {
  LBM_COLL_NONE = 0,
  LBM_COLL_BGK = 1,      /* cmpt::BGK     collision.hpp:76-96            */
  LBM_COLL_BGK_GUO = 2,  /* cmpt::BGKGuo  collision.hpp:135-204          */
  LBM_COLL_PSM = 3,      /* cmpt::Psm     psm.hpp:58-89                  */
  LBM_COLL_HLBM = 4,     /* cmpt::Hlbm    hlbm.hpp:62-105                */
  LBM_COLL_FPLBM = 5     /* cmpt::FpLbm   fplbm.hpp:60-133: Guo forcing with the cell
                            force field, u = m/rho; writes rho and the momentum m */
} lbm_collision_kind;

typedef enum lbm_particle_coupling {
  LBM_COUPLING_PSM = 0,  /* PsmReset + PsmOneParticleGather (psm.hpp:20-32,101-445)        */
  LBM_COUPLING_HLBM = 1, /* HlbmReset + HlbmOneParticleGather/Stamp (hlbm.hpp:28-49,111-325) */
  LBM_COUPLING_FPLBM = 2 /* FpLbmReset + FpLbmOneParticle (fplbm.hpp:21-58,135-247)          */
} lbm_particle_coupling;

/* lbm_op_desc.flags */
#define LBM_OPF_GUO_CELL_FORCE  0x1u  /* BGKGuo also reads the per-cell force field (collision.hpp:260-263) */
#define LBM_OPF_PSM_STAMP_DISC  0x2u  /* cmpt::OneParticleAt before Psm (particle/particle_opa.hpp:14-46) */
#define LBM_OPF_PSM_FORCE_GATHER 0x4u /* cmpt::ForceGather after Psm (momentum_gather.hpp:15-52)         */
#define LBM_OPF_PSM_CELL_FORCE  0x8u  /* cmpt::PSM variant writing Σ(f_i-f_opp)c_i·B (component/psm.hpp:16-80) */
/* D3Q27 Zou/He pressure without the tangential terms: each unknown is its
 * mirror across the face plus 6 w_i ρu_x c_ix only.  No reference
 * counterpart.  The reference's reconstruction (boundary.hpp:539-754) adds
 * the tangential terms to mirror pairs, which leaves the cell with 4/3 of
 * its tangential momentum every step; this variant keeps mass (Σf = ρ_set)
 * and tangential momentum, so 3D channels stay stable.  Default off. */
#define LBM_OPF_ZOUHE_NORMAL_ONLY 0x10u

typedef struct lbm_op_desc {
  int32_t boundary;      /* lbm_boundary_kind                                   */
  int32_t collision;     /* lbm_collision_kind                                  */
  int32_t east;          /* Zou/He template flag, reference meaning: 1 = x=0 face */
  uint32_t flags;        /* LBM_OPF_*                                           */
  double tau;            /* relaxation time; ω = T(1)/T(τ) (collision.hpp:25-38) */
  /* Zou/He pressure set-point, as the drivers compute it per step:
   *   ρ_set(t) = t < ramp_steps ? rho_start + rho_slope·t : rho_final
   * (poiseulle_2d_gpu.cpp:182-190).  ramp_steps = 0 means ρ_set = rho_final. */
  double rho_start, rho_slope, rho_final;
  uint64_t ramp_steps;
  double vec[3];         /* BGKGuo global force / Equilibrium velocity          */
  double rho_eq;         /* Equilibrium density                                 */
  double disc[5];        /* OneParticleAt: centre x,y,z, radius, eps            */
} lbm_op_desc;

/* ---- particles: particle/schema.hpp ParticleRigidBody + Spheroid ----------
 * One record per particle, doubles; 2D uses component 0 of the rotation
 * fields (the scalar out-of-plane rotation, schema.hpp:22-33).             */
typedef struct lbm_particle {
  double position[3], position_old[3];
  double rotation[3], rotation_old[3];
  double acceleration[3], angular_acceleration[3];
  double force[3], torque[3];
  double mass, inertia, density, radius;
} lbm_particle;

typedef enum lbm_porosity_mode {
  LBM_POROSITY_SCATTER = 0,  /* PsmOneParticleGather stamp: each AABB cell takes
                                the porosity of the highest-index particle whose
                                AABB holds it (psm.hpp:258-271, serial order)   */
  LBM_POROSITY_PRODUCT = 1   /* cell_porosity: product over cell-list candidates
                                (particle/couple.hpp:26-60)                     */
} lbm_porosity_mode;

typedef struct lbm_particle_params {
  double k_n, zeta_n, gamma_t, mu;  /* dem_params (particle/contact.hpp:36-61)      */
  int32_t walls;                    /* ParticleContact with extent (contact.hpp:533-546) */
  int32_t porosity_mode;            /* lbm_porosity_mode                             */
  int32_t psm_torque;               /* 0: reference PSM (no torque, psm.hpp:404-443);
                                       1: add -Σ r×F_cell (hlbm.hpp:234,249-252)    */
  int32_t cell_capacity;            /* ParticleFluidCouple<Cap> bin capacity         */
  uint64_t sub_steps;               /* particle sub-steps; 0 means 1 (integrate.hpp:16-19) */
  double body_acceleration[3];      /* ParticleIntegrate body acceleration           */
  double eps;                       /* porosity ramp width (psm.hpp:213: 1.5)        */
  double max_radius;                /* cell-list sizing (cell_list.hpp:185-206)      */
  int32_t state_fp64;               /* 0: particle state and arithmetic in the lattice's
                                       compute type, as the reference (default);
                                       1: FP64 particle state and arithmetic whatever
                                       the storage (no reference counterpart)       */
  int32_t coupling;                 /* lbm_particle_coupling (HLBM needs scatter porosity) */
  int32_t fixed_axis;               /* 0: ParticleIntegrate; 1: ParticleAxisIntegrate
                                       (integrate.hpp:134-189,298-327): acceleration 0,
                                       position pinned, spin about spin_axis only      */
  double spin_axis[3];              /* fixed axis (normalised in the particle type as the
                                       component does; 2D uses the sign of [0])        */
} lbm_particle_params;

typedef struct lbm_ctx lbm_ctx;

/* ---- context ------------------------------------------------------------- */
const char* lbm_last_error(void);
const char* lbm_version(void);
/* Number of visible GPUs (0 when none); never allocates. */
int lbm_device_count(void);

lbm_status lbm_ctx_create(lbm_ctx** out, int32_t lattice, int32_t storage,
                          const uint64_t extent[3], int32_t device);
lbm_status lbm_ctx_destroy(lbm_ctx* ctx);

/* Cell flags (reference meaning: 0 ghost, 1 wall, 2 fluid, 3+ driver-defined,
 * geometry.hpp:16-19).  Flag 0 inside the lattice is rejected with
 * LBM_ERR_NOT_SUPPORTED: the single-buffer AA layout cannot hold the
 * reference's stale second buffer for such cells. */
lbm_status lbm_set_flags(lbm_ctx* ctx, const uint8_t* flags);
lbm_status lbm_get_flags(lbm_ctx* ctx, uint8_t* flags);
/* Lattice, storage and extent the context was created with (the
 * compile-time Desc, T and chunk Sides of the reference's field types). */
lbm_status lbm_ctx_info(lbm_ctx* ctx, int32_t* lattice, int32_t* storage, uint64_t extent[3]);
/* Periodic wrap per axis (bit 0 = x, 1 = y, 2 = z).  No reference
 * counterpart: the reference only has zero ghosts (stream.hpp:47). */
lbm_status lbm_set_periodic(lbm_ctx* ctx, uint32_t axis_mask);
lbm_status lbm_set_operator(lbm_ctx* ctx, uint8_t flag, const lbm_op_desc* op);
/* Per-cell velocity for LBM_BC_ZOUHE_VELOCITY_X cells, D doubles per cell,
 * u[d*N + cell] (the per-cell vel built in stdkr_2d_bgk/step.hpp:44-58). */
lbm_status lbm_set_boundary_velocity(lbm_ctx* ctx, const double* u);
/* Per-cell force field ("force" macro), D doubles per cell. */
lbm_status lbm_set_force_field(lbm_ctx* ctx, const double* force);
lbm_status lbm_get_force_field(lbm_ctx* ctx, double* force);
lbm_status lbm_set_porosity(lbm_ctx* ctx, const double* porosity);
lbm_status lbm_get_porosity(lbm_ctx* ctx, double* porosity);
/* HLBM particle sums as the last step left them: N, D doubles per cell
 * (n[d*N + cell]), and D, one per cell — the reference's particle_N and
 * particle_D fields (HlbmReset / HlbmOneParticleStamp, hlbm.hpp:28-49,264-308).
 * Zero where no HLBM coupling has run. */
lbm_status lbm_get_hlbm_sums(lbm_ctx* ctx, double* n, double* d);

/* Populations: f[i*N + cell], the pre-collision (post-stream) populations of
 * the next step, i.e. what the reference holds in dfs_old (t even) or dfs
 * (t odd) — whatever the physical AA phase is. */
lbm_status lbm_set_populations(lbm_ctx* ctx, const double* f);
lbm_status lbm_get_populations(lbm_ctx* ctx, double* f);
/* The same populations in a chosen host layout:
 *  LBM_LAYOUT_SOA            f[i*N + cell] (as above), Q*N doubles;
 *  LBM_LAYOUT_REF_AOS_GHOST  the reference's dfs field, a
 *    Chunk<FixedArray<T,Q>, 1, nx, ny[, nz]> (chunk.hpp:11-24, flatten.hpp:
 *    11-39): cells of the ghost-widened extent x-fastest, cell (x, y, z) of
 *    the lattice at ((x+1) + (nx+2)*((y+1) + (ny+2)*(z+1))), the Q values of
 *    a cell consecutive; Q*(nx+2)*(ny+2)[*(nz+2)] doubles.  Get writes the
 *    ghost ring as zeros (the reference's ghost values); set ignores it.
 * lbm_population_count gives the array length a layout takes. */
typedef enum lbm_layout { LBM_LAYOUT_SOA = 0, LBM_LAYOUT_REF_AOS_GHOST = 1 } lbm_layout;
lbm_status lbm_population_count(lbm_ctx* ctx, int32_t layout, uint64_t* count);
lbm_status lbm_set_populations_layout(lbm_ctx* ctx, const double* f, int32_t layout);
lbm_status lbm_get_populations_layout(lbm_ctx* ctx, double* f, int32_t layout);
/* Fill every cell with eq(ρ, u) in storage precision (init of the drivers,
 * poiseulle_2d_gpu.cpp:104-134) and reset macros to (ρ, u). */
lbm_status lbm_init_equilibrium(lbm_ctx* ctx, double rho, const double u[3]);
/* Per-cell variant: cell x gets eq(rho[x], u[d*N + x]) and macros (rho, u).
 * No reference counterpart; it starts a lattice from a predicted field (the
 * learned warm start below). */
lbm_status lbm_init_equilibrium_field(lbm_ctx* ctx, const double* rho, const double* u);

/* Macros with the reference's meaning: density/velocity written by the
 * collision from the pre-collision populations of the last step; wall cells
 * keep what they held (collision.hpp:83-89, psm.hpp:67-73). */
lbm_status lbm_get_macros(lbm_ctx* ctx, double* rho, double* u);
lbm_status lbm_set_macros(lbm_ctx* ctx, const double* rho, const double* u);

/* Advance nsteps reference steps t0 .. t0+nsteps-1.  Without particles one
 * step is collide+stream (poiseulle_2d_gpu.cpp:342-399); with particles it is
 * porosity reset, particle reset/gather/contact/integrate, collide, stream
 * (moving_poiseulle_particles_2d_psm_gpu/sim.cpp:176-280). */
lbm_status lbm_step(lbm_ctx* ctx, uint64_t t0, uint64_t nsteps);
/* Steps per hipGraph replay inside lbm_step (0 launches every step; even,
 * at most 4096; default 32 up to 2^20 cells, 0 above, where a step's kernels
 * are long enough that replays gain nothing).  No reference counterpart.  Once every Zou/He
 * density ramp has ended the step index no longer enters any kernel, so whole
 * runs of steps are replayed from one captured graph (recaptured whenever
 * an argument it captured changes); results are bit-identical either way. */
lbm_status lbm_set_step_graph(lbm_ctx* ctx, uint32_t steps);
/* Waits for the context's queued work.  Kernel-side faults are kept in a
 * device status word (the reference's kernels have no error channel and
 * return early, e.g. particle/aabb.hpp:25-40) and reported here, then
 * cleared:
 *   LBM_ERR_INVALID_STATE  a density or velocity the collision wrote was not
 *                          finite (checked on every macro-writing step: the
 *                          last step of each lbm_step call, or every step
 *                          where the macros are kept each step);
 *   LBM_ERR_OVERFLOW       a cell-list bin was full (product porosity mode;
 *                          the count is in lbm_cell_list_overflow).
 * lbm_step reports a fault of an earlier call's finished work the same way
 * (without waiting for unfinished work), before it queues anything. */
lbm_status lbm_sync(lbm_ctx* ctx);
/* FP64 steady-state residual (no reference counterpart; the batch calls
 * use the same definition): with tracking on, each lbm_step call also keeps
 * the velocity of the step before its last, and lbm_residual gives
 *   out[0] = r = |u_t - u_{t-1}|_2 / |u_t|_2  (0 when both are 0, NaN when
 *            a sum is not finite),
 *   out[1] = sum over cells and components of (u_t - u_{t-1})^2,
 *   out[2] = sum of u_t^2,
 * summed in FP64 on the GPU in a fixed order (blocks of 2048 cells, then the
 * block sums), so equal inputs give equal bits.  u_t is the velocity field
 * the last step wrote; u_{t-1} the one the step before wrote (for a
 * one-step call, the previous call's last step).  lbm_set_residual
 * allocates (enable = 1) or frees (0) the second velocity field. */
lbm_status lbm_set_residual(lbm_ctx* ctx, int32_t enable);
lbm_status lbm_residual(lbm_ctx* ctx, double out[3]);
/* lbm_residual without allocating anything.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The same kernels and summation order as lbm_residual (equal bits), with
 * the block sums in a static device array of the library (8192 pairs of
 * doubles, 128 KiB per GPU, held until the process ends) instead of the
 * context's buffer; `out_device` is caller-owned device memory for the 3
 * doubles, written asynchronously on the context's stream (visible after
 * lbm_sync or any wait on that stream; no copy, no wait here).  Lattices
 * over 16,777,216 cells are rejected with LBM_ERR_NOT_SUPPORTED.  All calls
 * on one GPU share that array and every call rewrites all it reads, so
 * consecutive calls need nothing cleared; calls that overlap on one GPU
 * (contexts on other streams, host threads) race. */
lbm_status lbm_residual_noalloc(lbm_ctx* ctx, double* out_device);
/* Device time of the last lbm_step call, from HIP events recorded on the
 * context's stream around its launches. */
lbm_status lbm_last_step_ms(lbm_ctx* ctx, double* ms);

/* ---- launch shape of the fused step (no reference counterpart) -----------
 * The fused stream+collide kernel runs one row segment of block_x cells per
 * block, and addresses the population planes either with per-lane global
 * addresses (form 1) or through one buffer resource with the plane as a
 * scalar offset (form 2, where the array's bytes fit its 32-bit range); with
 * 16-bit storage and an even nx, form 3 gives each thread two cells whose
 * values move as 32-bit words (block_x threads then cover 2*block_x cells);
 * form 0 takes the lattice's default.  Results are bitwise the same in every
 * shape and form.  lbm_autotune times every width (64, 128, 256, no wider
 * than the row) in every form that applies, per AA phase, on the context's own lattice over
 * `rounds` passes (a pass runs four even/odd step pairs back to back per
 * candidate, each launch between HIP events on the context's stream; the
 * mean per phase of a pass, the least over passes, is kept), restores
 * the populations and the cell force and porosity fields it advanced, and
 * keeps the fastest shape; `chosen->us` is the chosen shape's time per phase.
 * Call it from an even step count, after flags and operators are set. */
typedef struct lbm_step_tuning {
  int32_t block_x;  /* threads per block along x: 64, 128 or 256     */
  int32_t form[2];  /* per AA phase (even, odd): 0 default, 1 global, 2 buffer, 3 pairs */
  double us[2];     /* lbm_autotune: measured time per launch of each phase */
} lbm_step_tuning;
lbm_status lbm_autotune(lbm_ctx* ctx, uint32_t rounds, lbm_step_tuning* chosen);
lbm_status lbm_set_step_tuning(lbm_ctx* ctx, const lbm_step_tuning* tuning);
lbm_status lbm_get_step_tuning(lbm_ctx* ctx, lbm_step_tuning* tuning);

/* ---- measured HBM ceiling (no reference counterpart) ----------------------
 * The bandwidth plain data movement reaches on `device`, as a measured
 * ceiling beside the published peak for the population kernels' roofline:
 *   LBM_CEILING_COPY    b = a between two buffers of `bytes`, 16 B per lane;
 *   LBM_CEILING_PLANES  one in-place nontemporal pass over 19 planes of
 *                       `bytes` in total, 4 B of every plane per lane (the
 *                       D3Q19 even AA phase without its arithmetic).
 * One untimed pass, then `reps` passes timed with HIP events on a stream of
 * its own; *gbps = (bytes read + bytes written) / mean pass time.  Allocates
 * and frees its buffers; bytes in [1 MiB, 64 GiB]. */
typedef enum lbm_ceiling_kind { LBM_CEILING_COPY = 0, LBM_CEILING_PLANES = 1 } lbm_ceiling_kind;
lbm_status lbm_hbm_ceiling(int32_t device, int32_t kind, uint64_t bytes, uint32_t reps, double* gbps);

/* ---- result output: write_vtk.hpp, write_csv.hpp, abstract/forst.hpp ------
 * `members` lists the fields of the struct being written, comma-separated, in
 * file order; each is one of the reference drivers' member keys
 *   velocity, force   Vector<T,D> chunks (lbm_get_macros / lbm_get_force_field)
 *   density, porosity Scalar<T> chunks
 *   info              the UInt8 flag chunk
 *   dfs               the FixedArray<T,Q> population chunk (ghost ring 1), holding
 *                     the populations of lbm_get_populations
 * (e.g. the MacroStruct of examples/stokes_drag_particle_2d_hlbm_gpu/sim.cpp:
 * 47-52 is "velocity,density,force,porosity").  T is double for F64 lattices
 * and float otherwise.  Values are printed as the reference's ostream does
 * (%g, 6 significant digits), so the files are byte-identical to the
 * reference writers' for the same field values. */

/* Host copies of a lattice's fields, for writing results that are already on
 * the host (batch members, post-processing) and for reading snapshots.  SoA
 * doubles, cell x-fastest, like the get/set calls; a member that is listed
 * needs its buffer, the others may be NULL. */
typedef struct lbm_field_set {
  int32_t lattice;   /* lbm_lattice: D and Q                                 */
  int32_t storage;   /* lbm_storage: T = double for F64, float otherwise       */
  uint64_t extent[3];
  double* velocity;  /* D*N */
  double* density;   /* N   */
  double* force;     /* D*N */
  double* porosity;  /* N   */
  uint8_t* info;     /* N   */
  double* dfs;       /* Q*N */
} lbm_field_set;

/* write_vtk_file (write_vtk.hpp:218-252): <out_dir>/vtk/<file_name>_<t>.vtk,
 * legacy ASCII STRUCTURED_POINTS; directories are created. */
lbm_status lbm_write_vtk(lbm_ctx* ctx, const char* out_dir, const char* file_name, uint64_t t,
                         const char* members);
/* write_csv_file (write_csv.hpp:168-181): one file per member,
 * <out_dir>/csv/<file_name>_<member>_<t>.csv, one cell per line (x fastest),
 * components comma-separated. */
lbm_status lbm_write_csv(lbm_ctx* ctx, const char* out_dir, const char* file_name, uint64_t t,
                         const char* members);
/* codec<Struct<Member<Chunk...>...>, enc::Forst>::encode (abstract/forst.hpp:
 * 580-600, forst_chunk.hpp:29-45) written to `path`: the members' chunk bytes
 * back to back, little endian, no framing; a chunk covers its ghost ring
 * (zeros here).  Populations need F64 or F32 storage. */
lbm_status lbm_write_forst(lbm_ctx* ctx, const char* path, const char* members);
/* codec<...>::decode (abstract/forst.hpp:602-609) from `path` into the
 * context: info -> lbm_set_flags, dfs -> lbm_set_populations (ghost ring
 * ignored), velocity/density -> lbm_set_macros, force, porosity.  A file
 * shorter than the struct is rejected (err::forst_truncated) with
 * LBM_ERR_INVALID_ARG; bytes past it are ignored, as the reference does.
 * With "info,dfs,velocity,density" (plus force, porosity when used) a
 * context restored from lbm_write_forst continues bit for bit. */
lbm_status lbm_read_forst(lbm_ctx* ctx, const char* path, const char* members);
/* The same four on host buffers; lbm_read_forst_fields fills the listed
 * members' buffers of `set` (extent, lattice and storage say what to expect). */
lbm_status lbm_write_vtk_fields(const lbm_field_set* set, const char* out_dir, const char* file_name,
                                uint64_t t, const char* members);
lbm_status lbm_write_csv_fields(const lbm_field_set* set, const char* out_dir, const char* file_name,
                                uint64_t t, const char* members);
lbm_status lbm_write_forst_fields(const lbm_field_set* set, const char* path, const char* members);
lbm_status lbm_read_forst_fields(lbm_field_set* set, const char* path, const char* members);

/* ---- particles ------------------------------------------------------------ */
lbm_status lbm_particles_set(lbm_ctx* ctx, const lbm_particle* p, uint32_t n,
                             const lbm_particle_params* params);
lbm_status lbm_particles_get(lbm_ctx* ctx, lbm_particle* p, uint32_t n);
/* Disc/sphere mass and inertia exactly as create_spheroid_particle_group
 * computes them in storage precision (particle/particle.hpp:79-101). */
lbm_status lbm_spheroid_mass(int32_t dims, int32_t storage, double radius,
                             double density, double* mass, double* inertia);
/* One PSM coupling pass at step t on the current populations: porosity field,
 * per-cell force field and the per-particle FP64 force/torque reduction
 * (psm.hpp:95-446).  The FP64 sums are read with lbm_particle_reduce. */
lbm_status lbm_psm_couple(lbm_ctx* ctx, uint64_t t);
lbm_status lbm_particle_reduce(lbm_ctx* ctx, double* force, double* torque);
/* DEM contact (contact.hpp:522-567) and integration (integrate.hpp:263-296). */
lbm_status lbm_dem_contact(lbm_ctx* ctx);
lbm_status lbm_particle_integrate(lbm_ctx* ctx);
/* Cell-list overflow count of the last build (cell_list.hpp:96-101). */
lbm_status lbm_cell_list_overflow(lbm_ctx* ctx, uint64_t* overflow);

/* ---- batched ensembles (no reference counterpart) ------------------------
 * `count` independent lattices of uneven size stepped by one persistent launch
 * per step, chains of steps replayed from a hipGraph.  Each member is an
 * ordinary lbm_ctx for set-up and read-back (flags, operators, populations,
 * macros); step members only through the lbm_batch_* calls.               */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_tuning {
  int32_t block;       /* threads per block = cells per tile row (64..256)  */
  int32_t rows;        /* rows per tile                                     */
  int32_t grid;        /* persistent blocks (0 = one per tile)              */
  double ms_per_step;  /* measured by lbm_batch_autotune                    */
} lbm_batch_tuning;

/* extents: count*3 values (nz ignored for D2Q9) */
lbm_status lbm_batch_create(lbm_batch** out, int32_t lattice, int32_t storage, uint32_t count,
                            const uint64_t* extents, int32_t device);
lbm_status lbm_batch_destroy(lbm_batch* b);
/* The i-th member's context (owned by the batch; NULL if out of range). */
lbm_ctx* lbm_batch_member(lbm_batch* b, uint32_t i);
/* Advance every active member nsteps reference steps t0 .. t0+nsteps-1. */
lbm_status lbm_batch_step(lbm_batch* b, uint64_t t0, uint64_t nsteps);
/* Run to steady state: every check_every steps (even, >= 2) each active
 * member's residual r = |u_t - u_{t-1}|_2 / |u_t|_2 over all cells is formed
 * in FP64 on the GPU; members with r < tol stop.  Returns when all stopped or
 * after max_steps.  steps_out[i]: steps member i ran; res_out[i]: its last r
 * (NaN for a member whose fields stopped being finite; it stops too). */
lbm_status lbm_batch_run(lbm_batch* b, uint64_t t0, double tol, uint64_t check_every,
                         uint64_t max_steps, uint64_t* steps_out, double* res_out);
/* Times candidate tile shapes on the batch (state restored afterwards) and
 * keeps the fastest. */
lbm_status lbm_batch_autotune(lbm_batch* b, lbm_batch_tuning* chosen);
lbm_status lbm_batch_set_tuning(lbm_batch* b, const lbm_batch_tuning* tuning);
/* Device time and member cell updates of the last step/run call. */
lbm_status lbm_batch_stats(lbm_batch* b, double* ms, uint64_t* cell_updates);

/* ---- learned warm start (no reference counterpart) -------------------------
 * A convolutional encoder-decoder on 2D lattices of height x width cells
 * (both multiples of 8) that maps (obstacle mask, inlet u_max, tau) to the
 * steady (ux, uy, rho).  Activations are NHWC bf16, convolutions are 3x3
 * same-padded bf16 MFMA GEMMs accumulating in FP32; master weights and Adam
 * state are FP32.  Layers, in parameter order, each W[3][3][Cin][Cout] then
 * b[Cout] (w = widths):
 *   e1a 8->w0, e1b w0->w0, pool, e2a w0->w1, e2b w1->w1, pool,
 *   e3a w1->w2, e3b w2->w2, pool, e4a w2->w3, e4b w3->w3,
 *   up+cat(e3b) d3 (w3+w2)->w2, up+cat(e2b) d2 (w2+w1)->w1,
 *   up+cat(e1b) d1 (w1+w0)->w0, out w0->3 (no ReLU; every other layer ReLU).
 * Input features per cell (lbm_net_features): fluid indicator (flag >= 2),
 * u_max/0.08, (tau-0.5)/0.7, the inlet profile shape 4(y-1/2)(ny-3/2-y)/(ny-2)^2,
 * x/(nx-1), y/(ny-1), log1p(P(x)) with P(x) = 12 nu (Q/u_max) sum_{x'>x} h(x')^-3
 * / u_max (a 1D hydraulic pressure estimate, h = open cells per column, Q the
 * inlet flux, nu = (tau-1/2)/3), and log((ny-2)/h(x)).  Targets on fluid
 * cells: ux/u_max, uy/u_max and (rho-1)/(3 u_max^2).                         */
typedef struct lbm_net lbm_net;
typedef struct lbm_net_config {
  int32_t height, width; /* lattice rows (ny) and columns (nx)              */
  int32_t widths[4];     /* channels per level, multiples of 8              */
  int32_t batch;         /* samples per training / inference launch         */
  int32_t capacity;      /* samples the on-device training set may hold     */
  uint64_t seed;         /* weight initialisation (He normal)               */
} lbm_net_config;

lbm_status lbm_net_create(lbm_net** out, const lbm_net_config* cfg, int32_t device);
lbm_status lbm_net_destroy(lbm_net* net);
lbm_status lbm_net_param_count(lbm_net* net, uint64_t* count);
/* Flat FP32 parameters in the layer order above. */
lbm_status lbm_net_get_params(lbm_net* net, float* params);
lbm_status lbm_net_set_params(lbm_net* net, const float* params);
/* Features of one lattice: flags (ny*nx, reference flag values) -> 8 floats
 * per cell, cell-major (the values the network sees, before bf16 rounding). */
lbm_status lbm_net_features(lbm_net* net, const uint8_t* flags, double u_max, double tau, float* features);
/* Append one steady state (flags, u_max, tau, rho[N], u[2N] in the
 * lbm_get_macros layout) to the on-device training set. */
lbm_status lbm_net_add_sample(lbm_net* net, const uint8_t* flags, double u_max, double tau,
                              const double* rho, const double* u);
lbm_status lbm_net_clear_samples(lbm_net* net);
lbm_status lbm_net_sample_count(lbm_net* net, uint32_t* count);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches drawn without
 * replacement per epoch (seeded); losses[k] = masked MSE of step k (may be NULL). */
lbm_status lbm_net_train(lbm_net* net, uint32_t steps, float lr, uint64_t seed, float* losses);
/* Predict count lattices: rho[count*N], u[count*2N] (lbm_get_macros layout,
 * solid cells rho 1, u 0). */
lbm_status lbm_net_predict(lbm_net* net, uint32_t count, const uint8_t* flags, const double* u_max,
                           const double* tau, double* rho, double* u);
/* Mass-flux correction of predicted velocities u[count*2N] (in place): in
 * steady flow every column x carries the inlet's flux, u_max times the sum
 * of the inlet parabola; each column whose predicted flux sum_y u_x exceeds
 * a fifth of it is rescaled (u_x and u_y) to carry exactly that flux.  The
 * network's column-wise flux errors are the slowest-decaying part of a warm
 * start's error (no reference counterpart). */
lbm_status lbm_net_correct_flux(lbm_net* net, uint32_t count, const uint8_t* flags, const double* u_max,
                                double* u);
/* Raw forward: x[count*N*8] features -> y[count*N*3] normalised outputs. */
lbm_status lbm_net_forward_raw(lbm_net* net, uint32_t count, const float* x, float* y);
/* Raw loss and gradient for the current parameters: features x, targets
 * t[count*N*3], mask m[count*N] (1 = counted); grad has param_count floats. */
lbm_status lbm_net_loss_grad_raw(lbm_net* net, uint32_t count, const float* x, const float* t,
                                 const float* m, double* loss, float* grad);
/* Output standardisation: the network's outputs are (target - mean) / stdev
 * per channel (ux, uy, rho targets above).  The first lbm_net_train call takes
 * mean and stdev from the training set's fluid cells and freezes them; a
 * network restored with lbm_net_set_params needs them restored too. */
lbm_status lbm_net_get_output_scale(lbm_net* net, float mean[3], float stdev[3]);
lbm_status lbm_net_set_output_scale(lbm_net* net, const float mean[3], const float stdev[3]);
/* Device time of the last train / predict call. */
lbm_status lbm_net_last_ms(lbm_net* net, double* ms);

#ifdef __cplusplus
}
#endif
#endif
