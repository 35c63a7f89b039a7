// C ABI of the engine (include/lbm_mirror_amd/lbm.h).
//
// One lbm_ctx owns one lattice on one GPU: its population buffer (AA layout,
// single copy), flags, operator table, macro fields and particle state, and
// one HIP stream on which every launch of the context is ordered.  No C++
// exception crosses the boundary; every entry point returns an lbm_status and
// leaves a message for lbm_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fields.hpp"
#include "kernels.hpp"

using namespace lbmx;

namespace {

thread_local std::string g_err;

lbm_status fail(lbm_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_CRITICAL, \
                  "%s: %s", #expr, hipGetErrorString(e_));                             \
  } while (0)

int dims_of(int lat) { return lat == LBM_D2Q9 ? 2 : 3; }
int q_of(int lat) { return lat == LBM_D2Q9 ? 9 : (lat == LBM_D3Q19 ? 19 : 27); }
size_t storage_bytes(int st) { return st == LBM_F64 ? 8 : (st == LBM_F32 ? 4 : 2); }
bool compute_double(int st) { return st == LBM_F64; }

// Rounds a double parameter to the storage's compute type T, as the
// reference's component constructors do (native_data_type<T>).
double to_T(int st, double v) { return compute_double(st) ? v : double(float(v)); }

const int kMaxParticles = 4096;

// ---- lbm_residual_noalloc's scratch ------------------------------------------
// The block sums of the residual (2 doubles per block of
// kResidualCellsPerBlock cells) for callers that cannot allocate: a file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in): each GPU's address is looked up once, with that GPU
// current, when the first context on it is created.
constexpr uint32_t kResidualScratchBlocks = 8192;  // lattices up to 16,777,216 cells
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[2 * kResidualScratchBlocks];
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`; call with `device` current
hipError_t scratch_lookup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return hipSuccess;
}

}  // namespace

// Error channel shared with the other translation units of the library.
lbm_status lbmx_report(lbm_status s, const char* msg) {
  g_err = msg;
  return s;
}

struct lbm_ctx {
  int lattice = 0, storage = 0, D = 2, Q = 9, device = 0;
  uint64_t n[3] = {1, 1, 1};
  uint32_t N = 0, Np = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  void* f = nullptr;
  uint8_t* flags = nullptr;
  OpDev* ops = nullptr;
  OpDev ops_h[kMaxOps];
  bool has_op[kMaxOps] = {};
  void* rho = nullptr;
  void* u = nullptr;
  void* force = nullptr;
  void* por = nullptr;
  void* bcvel = nullptr;
  void* pn = nullptr;  // HLBM N sums (D*N)
  void* pd = nullptr;  // HLBM D sums (N)
  uint32_t periodic = 0;
  int phase = 0;
  bool flags_set = false;
  LaunchShape shape;
  // lbm_step replays: steps_per_graph steps captured once, keyed by the bytes
  // of the launch arguments they captured (0 steps: every step launched)
  uint32_t steps_per_graph = 0;
  hipGraphExec_t step_graph = nullptr;
  uint64_t* t_dev = nullptr;  // the replayed graph's first step index
  uint64_t t_dev_value = ~uint64_t(0);  // what *t_dev holds once queued work ran (~0: unknown)
  std::vector<uint8_t> graph_key;
  // particles
  uint32_t np = 0;
  bool pdbl = false;  // particle records are doubles
  void* parts = nullptr;
  void* snap = nullptr;  // the records as the gather saw them (partners of the fused contact)
  double* f64 = nullptr;
  double* t64 = nullptr;
  double* partial = nullptr;
  uint32_t bpp = 1;
  lbm_particle_params pp{};
  uint64_t* cl_count = nullptr;
  uint64_t* cl_slot = nullptr;
  uint64_t* cl_overflow = nullptr;
  uint32_t* owner = nullptr;
  // PSM scatter porosity kept by the gathers' sweeps (ParticleArgs::boxes):
  // 12 ints per particle; por_fresh once the porosity is 1 outside the
  // previous boxes (after a full reset, kept by every sweep since)
  int32_t* boxes = nullptr;
  bool por_fresh = false;
  double bin = 1.0;
  uint64_t grid[3] = {1, 1, 1};
  uint64_t bins = 1;
  // fault words written by the kernels (kStatusNonFinite, kStatusOverflow):
  // pinned host memory, read by the host without a copy
  uint32_t* status_h = nullptr;
  uint32_t* status_d = nullptr;
  // residual tracking: the velocity of the step before the last, the block
  // partial sums and (r, Σ|Δu|², Σ|u|²)
  void* u_prev = nullptr;
  double* res_partial = nullptr;
  double* res_out = nullptr;
  size_t tsz() const { return compute_double(storage) ? 8 : 4; }
};

namespace {

hipError_t alloc_zero(void** p, size_t bytes, hipStream_t s) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return e;
  return hipMemsetAsync(*p, 0, bytes, s);
}

lbm_status ensure_field(lbm_ctx* c, void** field, size_t count, double fill) {
  if (*field) return LBM_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(alloc_zero(field, count * c->tsz(), c->stream));
  if (fill != 0.0) {
    std::vector<double> h(count, fill);
    double* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, count * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(tmp, h.data(), count * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(field_from_double(compute_double(c->storage), *field, tmp, count, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(tmp));
  }
  return LBM_OK;
}

// Porosity (1), cell force (0) and the HLBM sums N, D (0): what the HLBM
// collision and the HLBM particle coupling read and write.
lbm_status ensure_hlbm_fields(lbm_ctx* c) {
  lbm_status s = ensure_field(c, &c->por, c->N, 1.0);
  if (s) return s;
  s = ensure_field(c, &c->force, size_t(c->D) * c->N, 0.0);
  if (s) return s;
  s = ensure_field(c, &c->pn, size_t(c->D) * c->N, 0.0);
  if (s) return s;
  return ensure_field(c, &c->pd, c->N, 0.0);
}

Geometry geom(const lbm_ctx* c) {
  Geometry g;
  g.nx = int32_t(c->n[0]);
  g.ny = int32_t(c->n[1]);
  g.nz = int32_t(c->n[2]);
  g.N = c->N;
  g.Np = c->Np;
  g.periodic = c->periodic;
  return g;
}

lbm_status put_field(lbm_ctx* c, void* field, const double* host, size_t count, int as = -1) {
  double* tmp = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(&tmp, count * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(tmp, host, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(field_from_double(compute_double(as < 0 ? c->storage : as), field, tmp, count, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(tmp));
  return LBM_OK;
}

lbm_status get_field(lbm_ctx* c, const void* field, double* host, size_t count, int as = -1) {
  double* tmp = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(&tmp, count * sizeof(double)));
  HIP_TRY(field_to_double(compute_double(as < 0 ? c->storage : as), field, tmp, count, c->stream));
  HIP_TRY(hipMemcpyAsync(host, tmp, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(tmp));
  return LBM_OK;
}

// Storage whose compute type the particle records and arithmetic use.
int particle_storage(int storage, const lbm_particle_params& pp) {
  return (storage == LBM_F64 || pp.state_fp64) ? LBM_F64 : LBM_F32;
}

// normalise_spin_axis (integrate.hpp:115-131): the sign in 2D; in 3D the
// axis over its length, in T (math::normalize, abstract/math.hpp:70-87: a
// zero vector stays zero).  Results are exact in double.
template <class T>
void spin_axis_normalised(int D, const double* in, double* out) {
  out[0] = out[1] = out[2] = 0.0;
  if (D == 2) {
    out[0] = T(in[0]) < T(0) ? -1.0 : 1.0;
    return;
  }
  const T a[3] = {T(in[0]), T(in[1]), T(in[2])};
  T dot = T(0);
  for (int i = 0; i < 3; ++i) dot = T(dot + T(a[i] * a[i]));
  if (dot <= T(0)) {
    for (int i = 0; i < 3; ++i) out[i] = double(a[i]);
    return;
  }
  const T len = std::sqrt(dot);
  for (int i = 0; i < 3; ++i) out[i] = double(T(a[i] / len));
}

// AABB cells per gather block (one per thread): against two per thread, C3's
// 200 discs run 7% faster and a single r-19 sphere (69k cells, 256 blocks)
// 4% faster
#ifndef LBMX_GATHER_CELLS
#define LBMX_GATHER_CELLS 256
#endif
constexpr uint64_t kGatherCells = LBMX_GATHER_CELLS;

// PSM coupling with scatter porosity and per-block ownership: the gathers
// keep the porosity without a per-cell reset
bool sweep_eligible(const lbm_ctx* c) {
  return c->np > 0 && c->boxes && c->pp.coupling == LBM_COUPLING_PSM &&
         c->pp.porosity_mode == LBM_POROSITY_SCATTER && !c->owner;
}
bool sweep_ready(const lbm_ctx* c) { return sweep_eligible(c) && c->por_fresh; }

// Before the couplings of a call: one full reset (porosity 1, no previous
// boxes) where the porosity was set from outside or the particles changed
lbm_status prepare_sweep(lbm_ctx* c) {
  if (!sweep_eligible(c) || c->por_fresh) return LBM_OK;
  HIP_TRY(launch_sweep_reset(c->storage == LBM_F64, c->por, c->N, c->boxes, c->np, c->stream));
  c->por_fresh = true;
  return LBM_OK;
}

ParticleArgs particle_args(lbm_ctx* c) {
  ParticleArgs a;
  std::memset(&a, 0, sizeof a);  // padding too: the bytes key the step graphs
  const int pst = particle_storage(c->storage, c->pp);
  a.state64 = pst == LBM_F64 ? 1 : 0;
  a.f = c->f;
  a.por = c->por;
  a.force = c->force;
  a.nx = int32_t(c->n[0]);
  a.ny = int32_t(c->n[1]);
  a.nz = int32_t(c->n[2]);
  a.N = c->N;
  a.Np = c->Np;
  a.periodic = c->periodic;
  a.phase = c->phase;
  a.parts = c->parts;
  a.snap = c->snap;
  a.n = c->np;
  a.boxes = sweep_ready(c) ? c->boxes : nullptr;
  a.f64 = c->f64;
  a.t64 = c->t64;
  a.partial = c->partial;
  a.blocks_per_particle = c->bpp;
  a.eps = to_T(pst, c->pp.eps);
  a.sub_steps = c->pp.sub_steps;
  a.porosity_mode = c->pp.porosity_mode;
  a.psm_torque = c->pp.psm_torque;
  a.k_n = c->pp.k_n;
  a.zeta_n = c->pp.zeta_n;
  a.gamma_t = c->pp.gamma_t;
  a.mu = c->pp.mu;
  a.walls = c->pp.walls;
  for (int d = 0; d < 3; ++d) a.g[d] = to_T(pst, c->pp.body_acceleration[d]);
  a.cl_count = c->cl_count;
  a.cl_slot = c->cl_slot;
  a.cl_overflow = c->cl_overflow;
  a.owner = c->np > kLocalOwnerMax ? c->owner : nullptr;
  a.bin = c->bin;
  for (int d = 0; d < 3; ++d) a.grid[d] = c->grid[d];
  a.cap = uint32_t(c->pp.cell_capacity);
  a.hlbm = c->pp.coupling == LBM_COUPLING_HLBM ? 1 : 0;
  a.fplbm = c->pp.coupling == LBM_COUPLING_FPLBM ? 1 : 0;
  a.pn = c->pn;
  a.pd = c->pd;
  a.cflags = c->flags;
  a.umac = c->u;
  // fixed ramp width of the HLBM gather and the FPLBM particle (hlbm.hpp:173, fplbm.hpp:193)
  if (a.hlbm || a.fplbm) a.eps = to_T(pst, 1.5);
  a.status = c->status_d;
  a.fixed_axis = c->pp.fixed_axis ? 1 : 0;
  if (pst == LBM_F64)
    spin_axis_normalised<double>(c->D, c->pp.spin_axis, a.axis);
  else
    spin_axis_normalised<float>(c->D, c->pp.spin_axis, a.axis);
  return a;
}

lbm_status check_ready(lbm_ctx* c) {
  if (!c->flags_set) return fail(LBM_ERR_INVALID_STATE, "lbm_set_flags has not been called");
  for (int k = 0; k < kMaxOps; ++k) {
    if (!c->has_op[k]) continue;
    const OpDev& o = c->ops_h[k];
    if (o.coll == LBM_COLL_PSM && !c->por) return fail(LBM_ERR_INVALID_STATE, "PSM without porosity field");
    if (o.bc == LBM_BC_ZOUHE_VELOCITY_X && !c->bcvel)
      return fail(LBM_ERR_INVALID_STATE, "Zou/He velocity without lbm_set_boundary_velocity");
  }
  return LBM_OK;
}

lbm_status particle_pass(lbm_ctx* c, uint64_t t) {
  (void)t;
  ParticleArgs a = particle_args(c);
  a.fuse_finalize = 1;
  a.fuse_integrate = 1;
  HIP_TRY(launch_psm_couple(c->lattice, c->storage, a, c->stream));
  HIP_TRY(launch_dem_contact(c->D, c->storage, a, c->stream));
  if (a.hlbm) HIP_TRY(launch_hlbm_stamp(c->D, c->storage, a, c->stream));
  return LBM_OK;
}

// t_dev: where captured launches find the step index of the first step they
// replay (t is then the offset from it), else null
StepArgs step_args(const lbm_ctx* c, uint64_t t, bool write_macros, const uint64_t* t_dev = nullptr) {
  StepArgs a;
  std::memset(&a, 0, sizeof a);  // padding too: the bytes key the step graphs
  a.f = c->f;
  a.flags = c->flags;
  a.ops = c->ops;
  a.rho = c->rho;
  a.u = c->u;
  a.force = c->force;
  a.por = c->por;
  a.bcvel = c->bcvel;
  a.pn = c->pn;
  a.pd = c->pd;
  a.nx = int32_t(c->n[0]);
  a.ny = int32_t(c->n[1]);
  a.nz = int32_t(c->n[2]);
  a.N = c->N;
  a.Np = c->Np;
  a.periodic = c->periodic;
  a.t = t;
  a.t_dev = t_dev;
  a.write_macros = write_macros ? 1 : 0;
  a.status = c->status_d;
  return a;
}

// u_into: where this pass writes the velocity when it writes macros (null:
// the macro field)
lbm_status fluid_pass(lbm_ctx* c, uint64_t t, bool write_macros, const uint64_t* t_dev = nullptr,
                      void* u_into = nullptr) {
  StepArgs a = step_args(c, t, write_macros, t_dev);
  if (u_into) a.u = u_into;
  HIP_TRY(launch_aa_step(c->lattice, c->storage, c->phase == 1, a, c->shape, c->stream));
  c->phase ^= 1;
  return LBM_OK;
}

// Faults the kernels recorded since the last report, reported once.
lbm_status report_faults(lbm_ctx* c) {
  if (!c->status_h) return LBM_OK;
  volatile uint32_t* w = c->status_h;
  const uint32_t nonfinite = w[kStatusNonFinite], overflow = w[kStatusOverflow];
  if (!nonfinite && !overflow) return LBM_OK;
  w[kStatusNonFinite] = 0;
  w[kStatusOverflow] = 0;
  if (nonfinite)
    return fail(LBM_ERR_INVALID_STATE, "a macro-writing step met a non-finite density or velocity%s",
                overflow ? " (and a cell-list bin was full)" : "");
  return fail(LBM_ERR_OVERFLOW, "a cell-list bin was full (lbm_cell_list_overflow has the count)");
}

template <class X>
void append_bytes(std::vector<uint8_t>& k, const X& x) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(&x);
  k.insert(k.end(), p, p + sizeof(X));
}

// steps_per_graph full steps from step t (particle pass and fluid pass, from
// phase 0) launched as one graph, captured again whenever anything its
// launches captured has changed.  The step index enters the kernels only
// through the Zou/He density ramps: the captured fluid launches read it as
// *t_dev + their offset in the graph, and the graph's last node advances
// *t_dev by its length, so one graph serves every start step, inside the
// ramps or past them; *t_dev is set only where a replay does not continue
// the previous one.
lbm_status replay_steps(lbm_ctx* c, uint64_t t, bool every) {
  if (!c->t_dev) HIP_TRY(hipMalloc((void**)&c->t_dev, sizeof(uint64_t)));
  std::vector<uint8_t> key;
  append_bytes(key, step_args(c, 0, every, c->t_dev));
  append_bytes(key, c->shape);
  append_bytes(key, c->steps_per_graph);
  append_bytes(key, c->np);
  if (c->np > 0) {
    ParticleArgs pa = particle_args(c);
    pa.fuse_finalize = 1;
    append_bytes(key, pa);
  }
  if (!c->step_graph || key != c->graph_key) {
    if (c->step_graph) {  // (after its queued replays)
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipGraphExecDestroy(c->step_graph));
    }
    c->step_graph = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    lbm_status s = LBM_OK;
    for (uint32_t i = 0; i < c->steps_per_graph && !s; ++i) {
      if (c->np > 0) s = particle_pass(c, i);
      if (!s) s = fluid_pass(c, i, every, c->t_dev);
    }
    if (!s && launch_add_u64(c->t_dev, c->steps_per_graph, c->stream) != hipSuccess)
      s = fail(LBM_ERR_CRITICAL, "step index advance");
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (s) return s;
    HIP_TRY(e);
    HIP_TRY(hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(g));
    c->graph_key = key;  // capture flipped the phase an even number of times
  }
  if (c->t_dev_value != t) HIP_TRY(launch_set_u64(c->t_dev, t, c->stream));
  HIP_TRY(hipGraphLaunch(c->step_graph, c->stream));
  c->t_dev_value = t + c->steps_per_graph;
  return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_last_error(void) { return g_err.c_str(); }
const char* lbm_version(void) { return "lbm_mirror_amd 0.1 (gfx950)"; }

int lbm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

lbm_status lbm_ctx_create(lbm_ctx** out, int32_t lattice, int32_t storage,
                          const uint64_t extent[3], int32_t device) {
  if (!out || !extent) return fail(LBM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (lattice < LBM_D2Q9 || lattice > LBM_D3Q27) return fail(LBM_ERR_INVALID_ARG, "unknown lattice %d", lattice);
  if (storage < LBM_F64 || storage > LBM_BF16S) return fail(LBM_ERR_INVALID_ARG, "unknown storage %d", storage);
  const int D = dims_of(lattice);
  uint64_t n[3] = {extent[0], extent[1], D == 3 ? extent[2] : 1};
  uint64_t N = 1;
  for (int d = 0; d < 3; ++d) {
    if (n[d] == 0) return fail(LBM_ERR_INVALID_ARG, "empty extent on axis %d", d);
    N *= n[d];
  }
  const int Q = q_of(lattice);
  const uint64_t Np = (N + 255) / 256 * 256;
  if (N >= (1ull << 31) || n[0] >= (1u << 30) || n[1] >= (1u << 30) || n[2] >= 65536)
    return fail(LBM_ERR_NOT_SUPPORTED, "lattice too large for 32-bit cell indices");
  int ndev = lbm_device_count();
  if (ndev < 1) return fail(LBM_ERR_NOT_FOUND, "no GPU visible");
  if (device < 0 || device >= ndev) return fail(LBM_ERR_INVALID_ARG, "device %d of %d", device, ndev);
  lbm_ctx* c = new lbm_ctx();
  c->lattice = lattice;
  c->storage = storage;
  c->D = D;
  c->Q = Q;
  c->device = device;
  for (int d = 0; d < 3; ++d) c->n[d] = n[d];
  c->N = uint32_t(N);
  c->Np = uint32_t(Np);
  std::memset(c->ops_h, 0, sizeof(c->ops_h));
  auto bail = [&](lbm_status s) {
    lbm_ctx_destroy(c);
    return s;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(LBM_ERR_CRITICAL, "hipSetDevice"));
  {
    const hipError_t e = scratch_lookup(device);
    if (e != hipSuccess)
      return bail(fail(LBM_ERR_CRITICAL, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)));
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(LBM_ERR_CRITICAL, "hipStreamCreate"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(LBM_ERR_CRITICAL, "hipEventCreate"));
  if (hipHostMalloc((void**)&c->status_h, 64, hipHostMallocMapped) != hipSuccess)
    return bail(fail(LBM_ERR_OUT_OF_MEMORY, "fault words (pinned host memory)"));
  std::memset(c->status_h, 0, 64);
  if (hipHostGetDevicePointer((void**)&c->status_d, c->status_h, 0) != hipSuccess)
    return bail(fail(LBM_ERR_CRITICAL, "hipHostGetDevicePointer"));
  hipError_t e = alloc_zero(&c->f, size_t(Q) * Np * storage_bytes(storage), c->stream);
  if (e == hipSuccess) e = alloc_zero((void**)&c->flags, N, c->stream);
  if (e == hipSuccess) e = alloc_zero((void**)&c->ops, sizeof(OpDev) * kMaxOps, c->stream);
  if (e == hipSuccess) e = alloc_zero(&c->rho, N * c->tsz(), c->stream);
  if (e == hipSuccess) e = alloc_zero(&c->u, size_t(D) * N * c->tsz(), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess)
    return bail(fail(e == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_CRITICAL,
                     "allocation: %s", hipGetErrorString(e)));
  // block shape: one row segment of up to 256 cells, rows stacked to fill 256 threads
  int bx = 64;
  while (bx < 256 && uint64_t(bx) < n[0]) bx *= 2;
  c->shape.bx = bx;
  c->shape.by = 256 / bx;
  // step graphs pay off where launches are a visible share of a step; past
  // 2^20 cells (a 200 µs step at 8 M cells) they change nothing but cost a
  // capture
  c->steps_per_graph = N <= (uint64_t(1) << 20) ? 32u : 0u;
  *out = c;
  return LBM_OK;
}

lbm_status lbm_ctx_destroy(lbm_ctx* c) {
  if (!c) return LBM_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->f, c->flags, c->ops, c->rho, c->u, c->force, c->por, c->bcvel, c->pn, c->pd,
                  c->parts, c->snap, c->f64, c->t64, c->partial, c->cl_count, c->cl_slot,
                  c->cl_overflow, c->owner, c->t_dev, c->u_prev, c->res_partial, c->res_out, c->boxes};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->status_h) (void)hipHostFree(c->status_h);
  if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LBM_OK;
}

lbm_status lbm_set_flags(lbm_ctx* c, const uint8_t* flags) {
  if (!c || !flags) return fail(LBM_ERR_INVALID_ARG, "null argument");
  for (uint32_t i = 0; i < c->N; ++i)
    if (flags[i] == 0)
      return fail(LBM_ERR_NOT_SUPPORTED,
                  "cell %u has flag 0 (ghost) inside the lattice; the AA single-buffer layout "
                  "keeps no stale second buffer for it",
                  i);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(c->flags, flags, c->N, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->flags_set = true;
  return LBM_OK;
}

lbm_status lbm_get_flags(lbm_ctx* c, uint8_t* flags) {
  if (!c || !flags) return fail(LBM_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(flags, c->flags, c->N, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

lbm_status lbm_ctx_info(lbm_ctx* c, int32_t* lattice, int32_t* storage, uint64_t extent[3]) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (lattice) *lattice = c->lattice;
  if (storage) *storage = c->storage;
  if (extent)
    for (int d = 0; d < 3; ++d) extent[d] = c->n[d];
  return LBM_OK;
}

lbm_status lbm_set_step_graph(lbm_ctx* c, uint32_t steps) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (steps % 2 || steps > 4096) return fail(LBM_ERR_INVALID_ARG, "steps per graph must be even and <= 4096");
  c->steps_per_graph = steps;
  return LBM_OK;
}

lbm_status lbm_set_periodic(lbm_ctx* c, uint32_t mask) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (mask > 7) return fail(LBM_ERR_INVALID_ARG, "axis mask %u", mask);
  if (c->D == 2) mask &= 3u;
  if (c->phase != 0) {
    // the swapped layout's ghost handling depends on the wrap; change it only
    // in the natural layout
    return fail(LBM_ERR_INVALID_STATE, "set periodicity after an even number of steps");
  }
  c->periodic = mask;
  return LBM_OK;
}

lbm_status lbm_set_operator(lbm_ctx* c, uint8_t flag, const lbm_op_desc* op) {
  if (!c || !op) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (flag == 0) return fail(LBM_ERR_INVALID_ARG, "flag 0 is the ghost flag and carries no operator");
  if (flag >= kMaxOps) return fail(LBM_ERR_NOT_SUPPORTED, "operators on flags >= %d", kMaxOps);
  if (op->boundary < LBM_BC_NONE || op->boundary > LBM_BC_EQUILIBRIUM)
    return fail(LBM_ERR_INVALID_ARG, "unknown boundary kind %d", op->boundary);
  if (op->collision < LBM_COLL_NONE || op->collision > LBM_COLL_FPLBM)
    return fail(LBM_ERR_INVALID_ARG, "unknown collision kind %d", op->collision);
  if (op->boundary == LBM_BC_ZOUHE_PRESSURE_X && c->lattice == LBM_D3Q19)
    return fail(LBM_ERR_NOT_SUPPORTED, "the reference has Zou/He pressure for D2Q9 and D3Q27 only");
  if ((op->boundary == LBM_BC_ZOUHE_VELOCITY_X || op->boundary == LBM_BC_ANTI_BOUNCE_BACK) &&
      c->lattice != LBM_D2Q9)
    return fail(LBM_ERR_NOT_SUPPORTED, "this boundary is D2Q9-index-hard-coded in the reference");
  if (op->collision != LBM_COLL_NONE && !(op->tau > 0.0))
    return fail(LBM_ERR_INVALID_ARG, "tau must be positive");
  const int st = c->storage;
  OpDev o{};
  o.bc = op->boundary;
  o.coll = op->collision;
  o.east = op->east;
  o.flags = op->flags;
  o.omega = compute_double(st) ? 1.0 / op->tau : double(1.0f / float(op->tau));
  o.rho_start = op->rho_start;
  o.rho_slope = op->rho_slope;
  o.rho_final = op->rho_final;
  o.ramp_steps = op->ramp_steps;
  for (int d = 0; d < 3; ++d) o.vec[d] = to_T(st, op->vec[d]);
  o.rho_eq = to_T(st, op->rho_eq);
  for (int k = 0; k < 5; ++k) o.disc[k] = to_T(st, op->disc[k]);
  // components of F = cell force + global force that can be non-zero
  o.force_mask = 0;
  for (int d = 0; d < c->D; ++d)
    if ((o.flags & LBM_OPF_GUO_CELL_FORCE) || o.vec[d] != 0.0) o.force_mask |= 1u << d;
  if (o.coll == LBM_COLL_HLBM) {
    lbm_status s = ensure_hlbm_fields(c);
    if (s) return s;
  }
  if (o.coll == LBM_COLL_PSM) {
    lbm_status s = ensure_field(c, &c->por, c->N, 1.0);
    if (s) return s;
  }
  if (o.coll == LBM_COLL_FPLBM ||
      (o.coll == LBM_COLL_BGK_GUO && (o.flags & LBM_OPF_GUO_CELL_FORCE)) ||
      (o.coll == LBM_COLL_PSM && (o.flags & (LBM_OPF_PSM_CELL_FORCE | LBM_OPF_PSM_FORCE_GATHER)))) {
    lbm_status s = ensure_field(c, &c->force, size_t(c->D) * c->N, 0.0);
    if (s) return s;
  }
  c->ops_h[flag] = o;
  c->has_op[flag] = true;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(c->ops + flag, &o, sizeof(OpDev), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

lbm_status lbm_set_boundary_velocity(lbm_ctx* c, const double* u) {
  if (!c || !u) return fail(LBM_ERR_INVALID_ARG, "null argument");
  lbm_status s = ensure_field(c, &c->bcvel, size_t(c->D) * c->N, 0.0);
  if (s) return s;
  return put_field(c, c->bcvel, u, size_t(c->D) * c->N);
}

lbm_status lbm_set_force_field(lbm_ctx* c, const double* f) {
  if (!c || !f) return fail(LBM_ERR_INVALID_ARG, "null argument");
  lbm_status s = ensure_field(c, &c->force, size_t(c->D) * c->N, 0.0);
  if (s) return s;
  return put_field(c, c->force, f, size_t(c->D) * c->N);
}

lbm_status lbm_get_force_field(lbm_ctx* c, double* f) {
  if (!c || !f) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (!c->force) {
    std::fill(f, f + size_t(c->D) * c->N, 0.0);
    return LBM_OK;
  }
  return get_field(c, c->force, f, size_t(c->D) * c->N);
}

lbm_status lbm_set_porosity(lbm_ctx* c, const double* p) {
  if (!c || !p) return fail(LBM_ERR_INVALID_ARG, "null argument");
  lbm_status s = ensure_field(c, &c->por, c->N, 1.0);
  if (s) return s;
  c->por_fresh = false;  // the next coupling resets every cell, as PsmReset does
  return put_field(c, c->por, p, c->N);
}

lbm_status lbm_get_hlbm_sums(lbm_ctx* c, double* n, double* d) {
  if (!c || !n || !d) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (!c->pn || !c->pd) {
    std::fill(n, n + size_t(c->D) * c->N, 0.0);
    std::fill(d, d + c->N, 0.0);
    return LBM_OK;
  }
  lbm_status s = get_field(c, c->pn, n, size_t(c->D) * c->N);
  if (s) return s;
  return get_field(c, c->pd, d, c->N);
}

lbm_status lbm_get_porosity(lbm_ctx* c, double* p) {
  if (!c || !p) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (!c->por) {
    std::fill(p, p + c->N, 1.0);
    return LBM_OK;
  }
  return get_field(c, c->por, p, c->N);
}

lbm_status lbm_population_count(lbm_ctx* c, int32_t layout, uint64_t* count) {
  if (!c || !count) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (layout != LBM_LAYOUT_SOA && layout != LBM_LAYOUT_REF_AOS_GHOST)
    return fail(LBM_ERR_INVALID_ARG, "unknown population layout %d", layout);
  *count = population_count(c->lattice, geom(c), layout);
  return LBM_OK;
}

lbm_status lbm_set_populations_layout(lbm_ctx* c, const double* f, int32_t layout) {
  if (!c || !f) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (layout != LBM_LAYOUT_SOA && layout != LBM_LAYOUT_REF_AOS_GHOST)
    return fail(LBM_ERR_INVALID_ARG, "unknown population layout %d", layout);
  const size_t cnt = population_count(c->lattice, geom(c), layout);
  double* tmp = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(&tmp, cnt * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(tmp, f, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(pack_populations(c->lattice, c->storage, c->f, tmp, geom(c), c->stream, layout));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(tmp));
  c->phase = 0;
  return LBM_OK;
}

lbm_status lbm_get_populations_layout(lbm_ctx* c, double* f, int32_t layout) {
  if (!c || !f) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (layout != LBM_LAYOUT_SOA && layout != LBM_LAYOUT_REF_AOS_GHOST)
    return fail(LBM_ERR_INVALID_ARG, "unknown population layout %d", layout);
  const size_t cnt = population_count(c->lattice, geom(c), layout);
  double* tmp = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(&tmp, cnt * sizeof(double)));
  if (layout == LBM_LAYOUT_REF_AOS_GHOST)  // the ghost ring holds zeros
    HIP_TRY(hipMemsetAsync(tmp, 0, cnt * sizeof(double), c->stream));
  HIP_TRY(unpack_populations(c->lattice, c->storage, c->f, tmp, geom(c), c->phase, c->stream, layout));
  HIP_TRY(hipMemcpyAsync(f, tmp, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(tmp));
  return LBM_OK;
}

lbm_status lbm_set_populations(lbm_ctx* c, const double* f) {
  return lbm_set_populations_layout(c, f, LBM_LAYOUT_SOA);
}

lbm_status lbm_get_populations(lbm_ctx* c, double* f) {
  return lbm_get_populations_layout(c, f, LBM_LAYOUT_SOA);
}

}  // extern "C"

namespace {

// eq(ρ, u) in T with the reference's arithmetic (equilibrium.hpp:7-48),
// evaluated on the host; out[i] holds the value rounded to T.
void host_equilibrium(int lattice, int storage, double rho, const double* u, double* out) {
  auto fill = [&](auto L, auto zero) {
    using LL = decltype(L);
    using T = decltype(zero);
    T uu[3] = {T(0), T(0), T(0)};
    for (int d = 0; d < LL::D; ++d) uu[d] = T(u[d]);
    T r = T(rho);
    T vv = T(0);
    for (int d = 0; d < LL::D; ++d) vv = T(vv + T(uu[d] * uu[d]));
    for (int i = 0; i < LL::Q; ++i) {
      T vc = T(0);
      for (int d = 0; d < LL::D; ++d) vc = T(vc + T(uu[d] * T(cdir<LL>(i, d))));
      T a = T(vc * T(3.0));
      T wr = T(T(LL::w[i]) * r);
      T aa = T(a * a);
      out[i] = double(T(double(wr) * (((1.0 + double(a)) - 1.5 * double(vv)) + double(aa) * 0.5)));
    }
  };
  const bool dbl = compute_double(storage);
  if (lattice == LBM_D2Q9) dbl ? fill(D2Q9{}, 0.0) : fill(D2Q9{}, 0.0f);
  if (lattice == LBM_D3Q19) dbl ? fill(D3Q19{}, 0.0) : fill(D3Q19{}, 0.0f);
  if (lattice == LBM_D3Q27) dbl ? fill(D3Q27{}, 0.0) : fill(D3Q27{}, 0.0f);
}

}  // namespace

extern "C" {

lbm_status lbm_init_equilibrium(lbm_ctx* c, double rho, const double u[3]) {
  if (!c || !u) return fail(LBM_ERR_INVALID_ARG, "null argument");
  double eq[27];
  const int st = c->storage;
  host_equilibrium(c->lattice, st, rho, u, eq);
  std::vector<double> h(size_t(c->Q) * c->N);
  for (int i = 0; i < c->Q; ++i) std::fill(h.begin() + size_t(i) * c->N, h.begin() + size_t(i + 1) * c->N, eq[i]);
  lbm_status s = lbm_set_populations(c, h.data());
  if (s) return s;
  std::vector<double> r(c->N, to_T(st, rho)), uv(size_t(c->D) * c->N);
  for (int d = 0; d < c->D; ++d)
    std::fill(uv.begin() + size_t(d) * c->N, uv.begin() + size_t(d + 1) * c->N, to_T(st, u[d]));
  return lbm_set_macros(c, r.data(), uv.data());
}

lbm_status lbm_init_equilibrium_field(lbm_ctx* c, const double* rho, const double* u) {
  if (!c || !rho || !u) return fail(LBM_ERR_INVALID_ARG, "null argument");
  const int st = c->storage;
  const size_t N = c->N;
  std::vector<double> h(size_t(c->Q) * N), r(N), uv(size_t(c->D) * N);
  double eq[27], uc[3] = {0.0, 0.0, 0.0};
  for (size_t x = 0; x < N; ++x) {
    for (int d = 0; d < c->D; ++d) uc[d] = u[size_t(d) * N + x];
    host_equilibrium(c->lattice, st, rho[x], uc, eq);
    for (int i = 0; i < c->Q; ++i) h[size_t(i) * N + x] = eq[i];
    r[x] = to_T(st, rho[x]);
    for (int d = 0; d < c->D; ++d) uv[size_t(d) * N + x] = to_T(st, uc[d]);
  }
  lbm_status s = lbm_set_populations(c, h.data());
  if (s) return s;
  return lbm_set_macros(c, r.data(), uv.data());
}

lbm_status lbm_get_macros(lbm_ctx* c, double* rho, double* u) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (rho) {
    lbm_status s = get_field(c, c->rho, rho, c->N);
    if (s) return s;
  }
  if (u) return get_field(c, c->u, u, size_t(c->D) * c->N);
  return LBM_OK;
}

lbm_status lbm_set_macros(lbm_ctx* c, const double* rho, const double* u) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (rho) {
    lbm_status s = put_field(c, c->rho, rho, c->N);
    if (s) return s;
  }
  if (u) return put_field(c, c->u, u, size_t(c->D) * c->N);
  return LBM_OK;
}

lbm_status lbm_step(lbm_ctx* c, uint64_t t0, uint64_t nsteps) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  lbm_status s = check_ready(c);
  if (s) return s;
  s = report_faults(c);  // of earlier calls' work that has run
  if (s) return s;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  s = prepare_sweep(c);  // (outside any step graph)
  if (s) return s;
  // FPLBM reads the stored macro velocity of non-fluid cells every step
  const bool every = c->np > 0 && c->pp.coupling == LBM_COUPLING_FPLBM;
  // residual tracking: the step before the last keeps its velocity in u_prev
  const bool track = c->u_prev != nullptr;
  const size_t ubytes = size_t(c->D) * c->N * c->tsz();
  if (track && nsteps == 1)  // the step before is the previous call's last
    HIP_TRY(hipMemcpyAsync(c->u_prev, c->u, ubytes, hipMemcpyDeviceToDevice, c->stream));
  const uint64_t G = c->steps_per_graph;
  uint64_t k = 0;
  while (k < nsteps) {
    const uint64_t t = t0 + k;
    // a whole graph of steps: from phase 0, and never the last step of the
    // call (it writes the macros), nor the one before it when tracking
    if (G > 0 && c->phase == 0 && nsteps - k > G + (track ? 1u : 0u)) {
      s = replay_steps(c, t, every);
      if (s) return s;
      k += G;
      continue;
    }
    if (c->np > 0) {
      s = particle_pass(c, t);
      if (s) return s;
    }
    const bool before_last = track && k + 2 == nsteps;
    if (before_last && !every) {
      s = fluid_pass(c, t, true, nullptr, c->u_prev);
    } else {
      s = fluid_pass(c, t, k + 1 == nsteps || every);
      if (!s && before_last)
        HIP_TRY(hipMemcpyAsync(c->u_prev, c->u, ubytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if (s) return s;
    ++k;
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return LBM_OK;
}

lbm_status lbm_sync(lbm_ctx* c) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return report_faults(c);
}

lbm_status lbm_set_residual(lbm_ctx* c, int32_t enable) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  HIP_TRY(hipSetDevice(c->device));
  if (!enable) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    void* ptrs[] = {c->u_prev, c->res_partial, c->res_out};
    for (void* p : ptrs)
      if (p) HIP_TRY(hipFree(p));
    c->u_prev = nullptr;
    c->res_partial = c->res_out = nullptr;
    return LBM_OK;
  }
  if (c->u_prev) return LBM_OK;
  HIP_TRY(alloc_zero(&c->u_prev, size_t(c->D) * c->N * c->tsz(), c->stream));
  HIP_TRY(alloc_zero((void**)&c->res_partial, 2 * size_t(residual_blocks(c->N)) * sizeof(double), c->stream));
  HIP_TRY(alloc_zero((void**)&c->res_out, 3 * sizeof(double), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

lbm_status lbm_residual(lbm_ctx* c, double out[3]) {
  if (!c || !out) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (!c->u_prev) return fail(LBM_ERR_INVALID_STATE, "residual tracking is off (lbm_set_residual)");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_residual(c->D, compute_double(c->storage), c->u, c->u_prev, c->N, c->res_partial, c->res_out,
                          c->stream));
  HIP_TRY(hipMemcpyAsync(out, c->res_out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

lbm_status lbm_residual_noalloc(lbm_ctx* c, double* out_device) {
  if (!c || !out_device) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (residual_blocks(c->N) > kResidualScratchBlocks)
    return fail(LBM_ERR_NOT_SUPPORTED,
                "lbm_residual_noalloc: %u cells exceed its %u-cell maximum (scratch of %u block sums); "
                "use lbm_residual",
                c->N, kResidualScratchBlocks * kResidualCellsPerBlock, kResidualScratchBlocks);
  if (!c->u_prev) return fail(LBM_ERR_INVALID_STATE, "residual tracking is off (lbm_set_residual)");
  double* partial = g_scratch_ptr[c->device];  // looked up when the context was created
  if (!partial) return fail(LBM_ERR_INVALID_STATE, "no scratch address for device %d", c->device);
  HIP_TRY(hipSetDevice(c->device));
  // every block sum the final pass reads is written first: nothing to clear
  HIP_TRY(launch_residual(c->D, compute_double(c->storage), c->u, c->u_prev, c->N, partial, out_device,
                          c->stream));
  return LBM_OK;
}

lbm_status lbm_last_step_ms(lbm_ctx* c, double* ms) {
  if (!c || !ms) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (!c->timed) return fail(LBM_ERR_INVALID_STATE, "no lbm_step call yet");
  HIP_TRY(hipEventSynchronize(c->ev1));
  float m = 0.f;
  HIP_TRY(hipEventElapsedTime(&m, c->ev0, c->ev1));
  *ms = m;
  return LBM_OK;
}

// ---------------------------------------------------------------- step tuning
lbm_status lbm_set_step_tuning(lbm_ctx* c, const lbm_step_tuning* t) {
  if (!c || !t) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (t->block_x != 64 && t->block_x != 128 && t->block_x != 256)
    return fail(LBM_ERR_INVALID_ARG, "block_x must be 64, 128 or 256");
  for (int p = 0; p < 2; ++p) {
    if (t->form[p] < 0 || t->form[p] > 3) return fail(LBM_ERR_INVALID_ARG, "form must be 0, 1, 2 or 3");
    if (t->form[p] == 3 && (storage_bytes(c->storage) != 2 || c->n[0] % 2))
      return fail(LBM_ERR_NOT_SUPPORTED, "form 3 (two cells per thread) needs 16-bit storage and an even nx");
  }
  c->shape.bx = t->block_x;
  c->shape.by = 1;
  c->shape.form[0] = t->form[0];
  c->shape.form[1] = t->form[1];
  return LBM_OK;
}

lbm_status lbm_autotune(lbm_ctx* c, uint32_t rounds, lbm_step_tuning* chosen) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (rounds == 0 || rounds > 64) return fail(LBM_ERR_INVALID_ARG, "rounds must be in [1, 64]");
  lbm_status s = check_ready(c);
  if (s) return s;
  if (c->phase != 0) return fail(LBM_ERR_INVALID_STATE, "autotune from the natural layout (even step count)");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the trial steps advance the populations and may write the cell force and
  // porosity fields (PSM variants): those are saved and restored; the macros
  // are not written by the trials
  struct Saved {
    void* field;
    void* copy;
    size_t bytes;
  };
  std::vector<Saved> saved;
  auto release = [&]() {
    for (Saved& v : saved)
      if (v.copy) (void)hipFree(v.copy);
  };
  const size_t fbytes = size_t(c->Q) * c->Np * storage_bytes(c->storage);
  const std::pair<void*, size_t> fields[] = {{c->f, fbytes},
                                             {c->force, size_t(c->D) * c->N * c->tsz()},
                                             {c->por, size_t(c->N) * c->tsz()}};
  for (const auto& fl : fields) {
    if (!fl.first) continue;
    Saved v{fl.first, nullptr, fl.second};
    hipError_t e = hipMalloc(&v.copy, v.bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(v.copy, v.field, v.bytes, hipMemcpyDeviceToDevice, c->stream);
    saved.push_back(v);
    if (e != hipSuccess) {
      release();
      return fail(e == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_CRITICAL, "autotune save: %s",
                  hipGetErrorString(e));
    }
  }
  // each candidate runs kPairs even/odd pairs back to back (the sustained
  // rate, not that of isolated launches), every launch between two events
  constexpr int kPairs = 4;
  hipEvent_t ev[2 * kPairs + 1] = {};
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    release();
  };
#define TUNE_TRY(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      cleanup();                                                                         \
      return fail(LBM_ERR_CRITICAL, "autotune: %s: %s", #expr, hipGetErrorString(e_));   \
    }                                                                                    \
  } while (0)
  for (hipEvent_t& e : ev) TUNE_TRY(hipEventCreate(&e));
  // candidates: one-row blocks of 64, 128 or 256 threads (no wider than the
  // row needs), each phase in both addressing forms
  std::vector<int> widths;
  for (int bx = 64; bx <= 256; bx *= 2) {
    widths.push_back(bx);
    if (uint64_t(bx) >= c->n[0]) break;
  }
  const int nw = int(widths.size());
  // form 3 (two cells per thread) where the storage is 16-bit and rows even
  const int nf = (storage_bytes(c->storage) == 2 && c->n[0] % 2 == 0) ? 3 : 2;
  std::vector<double> best_us(size_t(nw) * 3 * 2, 1e30);  // [width][form-1][phase]
  const LaunchShape keep = c->shape;
  for (uint32_t r = 0; r < rounds; ++r)
    for (int wi = 0; wi < nw; ++wi)
      for (int form = 1; form <= nf; ++form) {
        LaunchShape sh = keep;
        sh.bx = widths[wi];
        sh.by = 1;
        sh.form[0] = sh.form[1] = form;
        const StepArgs a = step_args(c, 0, false);
        TUNE_TRY(hipEventRecord(ev[0], c->stream));
        for (int k = 0; k < kPairs; ++k) {
          TUNE_TRY(launch_aa_step(c->lattice, c->storage, false, a, sh, c->stream));
          TUNE_TRY(hipEventRecord(ev[2 * k + 1], c->stream));
          TUNE_TRY(launch_aa_step(c->lattice, c->storage, true, a, sh, c->stream));
          TUNE_TRY(hipEventRecord(ev[2 * k + 2], c->stream));
        }
        TUNE_TRY(hipEventSynchronize(ev[2 * kPairs]));
        double sum[2] = {0.0, 0.0};
        for (int k = 0; k < 2 * kPairs; ++k) {
          float m = 0.f;
          TUNE_TRY(hipEventElapsedTime(&m, ev[k], ev[k + 1]));
          sum[k & 1] += 1e3 * double(m);
        }
        double* b = &best_us[(size_t(wi) * 3 + (form - 1)) * 2];
        b[0] = std::min(b[0], sum[0] / kPairs);
        b[1] = std::min(b[1], sum[1] / kPairs);
      }
  // restore the state the trials advanced
  for (const Saved& v : saved)
    TUNE_TRY(hipMemcpyAsync(v.field, v.copy, v.bytes, hipMemcpyDeviceToDevice, c->stream));
  TUNE_TRY(hipStreamSynchronize(c->stream));
#undef TUNE_TRY
  cleanup();
  // per width the faster form of each phase; the width with the least total
  int bw = 0, bf[2] = {1, 1};
  double bsum = 1e30, bus[2] = {0.0, 0.0};
  for (int wi = 0; wi < nw; ++wi) {
    int f[2];
    double u[2];
    for (int p = 0; p < 2; ++p) {
      f[p] = 1;
      u[p] = best_us[(size_t(wi) * 3 + 0) * 2 + p];
      for (int form = 2; form <= nf; ++form) {
        const double t = best_us[(size_t(wi) * 3 + (form - 1)) * 2 + p];
        if (t < u[p]) {
          u[p] = t;
          f[p] = form;
        }
      }
    }
    if (u[0] + u[1] < bsum) {
      bsum = u[0] + u[1];
      bw = wi;
      bf[0] = f[0];
      bf[1] = f[1];
      bus[0] = u[0];
      bus[1] = u[1];
    }
  }
  c->shape = keep;
  c->shape.bx = widths[bw];
  c->shape.by = 1;
  c->shape.form[0] = bf[0];
  c->shape.form[1] = bf[1];
  if (chosen) {
    chosen->block_x = widths[bw];
    chosen->form[0] = bf[0];
    chosen->form[1] = bf[1];
    chosen->us[0] = bus[0];
    chosen->us[1] = bus[1];
  }
  return LBM_OK;
}

lbm_status lbm_get_step_tuning(lbm_ctx* c, lbm_step_tuning* t) {
  if (!c || !t) return fail(LBM_ERR_INVALID_ARG, "null argument");
  t->block_x = c->shape.bx;
  t->form[0] = c->shape.form[0];
  t->form[1] = c->shape.form[1];
  t->us[0] = t->us[1] = 0.0;
  return LBM_OK;
}

// ---------------------------------------------------------------- particles
lbm_status lbm_spheroid_mass(int32_t dims, int32_t storage, double r, double rho, double* mass,
                             double* inertia) {
  if (!mass || !inertia) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (dims != 2 && dims != 3) return fail(LBM_ERR_INVALID_ARG, "dims %d", dims);
  // create_spheroid_particle_group (particle/particle.hpp:79-101)
  auto run = [&](auto zero) {
    using T = decltype(zero);
    constexpr T pi = T(3.14159265358979323846);
    const T rr = T(r), de = T(rho);
    if (dims == 2) {
      auto m = T(T(T(pi * rr) * rr) * de);
      *mass = m;
      *inertia = T(0.5 * m * rr * rr);
    } else {
      auto m = (4.0 / 3.0) * pi * rr * rr * rr * de;
      *mass = T(m);
      *inertia = T(0.4 * m * rr * rr);
    }
  };
  if (storage == LBM_F64) run(0.0);
  else run(0.0f);
  return LBM_OK;
}

lbm_status lbm_particles_set(lbm_ctx* c, const lbm_particle* p, uint32_t n,
                             const lbm_particle_params* prm) {
  if (!c || (!p && n) || !prm) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (n > uint32_t(kMaxParticles)) return fail(LBM_ERR_NOT_SUPPORTED, "more than %d particles", kMaxParticles);
  if (prm->porosity_mode != LBM_POROSITY_SCATTER && prm->porosity_mode != LBM_POROSITY_PRODUCT)
    return fail(LBM_ERR_INVALID_ARG, "porosity mode %d", prm->porosity_mode);
  if (prm->cell_capacity <= 0 && prm->porosity_mode == LBM_POROSITY_PRODUCT)
    return fail(LBM_ERR_INVALID_ARG, "cell capacity must be positive");
  if (prm->coupling != LBM_COUPLING_PSM && prm->coupling != LBM_COUPLING_HLBM &&
      prm->coupling != LBM_COUPLING_FPLBM)
    return fail(LBM_ERR_INVALID_ARG, "coupling %d", prm->coupling);
  if (prm->coupling != LBM_COUPLING_PSM && prm->porosity_mode != LBM_POROSITY_SCATTER)
    return fail(LBM_ERR_NOT_SUPPORTED, "HLBM and FPLBM assign porosity per particle (scatter mode only)");
  if (prm->fixed_axis && prm->coupling == LBM_COUPLING_FPLBM)
    return fail(LBM_ERR_NOT_SUPPORTED, "FpLbmOneParticle integrates the particle itself; no fixed-axis variant");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int st = particle_storage(c->storage, *prm);
  const bool dbl = compute_double(st);
  // (re)allocate
  if (n != c->np || dbl != c->pdbl) {
    void* ptrs[] = {c->parts, c->snap, c->f64, c->t64, c->partial, c->boxes};
    for (void* q : ptrs)
      if (q) HIP_TRY(hipFree(q));
    c->parts = c->snap = nullptr;
    c->f64 = c->t64 = c->partial = nullptr;
    c->boxes = nullptr;
    c->np = 0;
    if (n > 0) {
      HIP_TRY(alloc_zero((void**)&c->boxes, size_t(n) * 12 * sizeof(int32_t), c->stream));
      HIP_TRY(alloc_zero(&c->parts, size_t(n) * kPartFields * (dbl ? 8 : 4), c->stream));
      HIP_TRY(alloc_zero(&c->snap, size_t(n) * kPartFields * (dbl ? 8 : 4), c->stream));
      HIP_TRY(alloc_zero((void**)&c->f64, size_t(n) * 3 * sizeof(double), c->stream));
      HIP_TRY(alloc_zero((void**)&c->t64, size_t(n) * 3 * sizeof(double), c->stream));
    }
  }
  c->pp = *prm;
  c->por_fresh = false;  // the next coupling starts from a full reset
  if (c->pp.cell_capacity <= 0) c->pp.cell_capacity = 4;
  // blocks per particle so that one block sweeps kGatherCells AABB cells: a
  // single large sphere (r 19, 69k cells) then spreads over 256 blocks
  // instead of waiting on a few
  double rmax = 0.0;
  for (uint32_t k = 0; k < n; ++k) rmax = std::max(rmax, p[k].radius);
  uint64_t box = 1;
  for (int d = 0; d < c->D; ++d) box *= uint64_t(2.0 * rmax + 3.0);
  uint32_t bpp = uint32_t(
      std::min<uint64_t>(256, std::max<uint64_t>(1, (box + kGatherCells - 1) / kGatherCells)));
  if (bpp != c->bpp || !c->partial) {
    if (c->partial) HIP_TRY(hipFree(c->partial));
    c->partial = nullptr;
    c->bpp = bpp;
    if (n > 0) HIP_TRY(alloc_zero((void**)&c->partial, size_t(n) * bpp * 6 * sizeof(double), c->stream));
  }
  if (n > 0) {
    lbm_status s = ensure_field(c, &c->por, c->N, 1.0);
    if (s) return s;
    s = ensure_field(c, &c->force, size_t(c->D) * c->N, 0.0);
    if (s) return s;
    if (prm->coupling == LBM_COUPLING_HLBM) {
      s = ensure_hlbm_fields(c);
      if (s) return s;
    }
    if (!c->owner && n > kLocalOwnerMax)
      HIP_TRY(alloc_zero((void**)&c->owner, size_t(c->N) * 2 * sizeof(uint32_t), c->stream));
  }
  // cell list (make_cell_list, particle/cell_list.hpp:185-206) in T
  {
    double b = dbl ? prm->max_radius * 2 : double(float(prm->max_radius) * 2.0f);
    if (!(b > 1)) b = 1;
    c->bin = b;
    uint64_t bins = 1;
    for (int d = 0; d < 3; ++d) {
      if (d < c->D) {
        double e = double(c->n[d]);
        double q = dbl ? e / b : double(float(e) / float(b));
        c->grid[d] = uint64_t(q) + 1u;
      } else {
        c->grid[d] = 1;
      }
      bins *= c->grid[d];
    }
    if (c->cl_count) HIP_TRY(hipFree(c->cl_count));
    if (c->cl_slot) HIP_TRY(hipFree(c->cl_slot));
    if (!c->cl_overflow) HIP_TRY(alloc_zero((void**)&c->cl_overflow, sizeof(uint64_t), c->stream));
    c->cl_count = c->cl_slot = nullptr;
    c->bins = bins;
    HIP_TRY(alloc_zero((void**)&c->cl_count, bins * sizeof(uint64_t), c->stream));
    HIP_TRY(alloc_zero((void**)&c->cl_slot, bins * uint64_t(c->pp.cell_capacity) * sizeof(uint64_t), c->stream));
  }
  // records in T
  if (n > 0) {
    std::vector<double> rec(size_t(n) * kPartFields, 0.0);
    for (uint32_t k = 0; k < n; ++k) {
      double* r = rec.data() + size_t(k) * kPartFields;
      for (int d = 0; d < 3; ++d) {
        r[PF_POS + d] = to_T(st, p[k].position[d]);
        r[PF_POS_OLD + d] = to_T(st, p[k].position_old[d]);
        r[PF_ROT + d] = to_T(st, p[k].rotation[d]);
        r[PF_ROT_OLD + d] = to_T(st, p[k].rotation_old[d]);
        r[PF_ACC + d] = to_T(st, p[k].acceleration[d]);
        r[PF_ANG_ACC + d] = to_T(st, p[k].angular_acceleration[d]);
        r[PF_FORCE + d] = to_T(st, p[k].force[d]);
        r[PF_TORQUE + d] = to_T(st, p[k].torque[d]);
      }
      r[PF_MASS] = to_T(st, p[k].mass);
      r[PF_INERTIA] = to_T(st, p[k].inertia);
      r[PF_DENSITY] = to_T(st, p[k].density);
      r[PF_RADIUS] = to_T(st, p[k].radius);
    }
    c->np = n;
    c->pdbl = dbl;
    lbm_status s = put_field(c, c->parts, rec.data(), rec.size(), st);
    if (s) return s;
  }
  c->np = n;
  c->pdbl = dbl;
  return LBM_OK;
}

lbm_status lbm_particles_get(lbm_ctx* c, lbm_particle* p, uint32_t n) {
  if (!c || (!p && n)) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (n > c->np) return fail(LBM_ERR_INVALID_ARG, "asked for %u of %u particles", n, c->np);
  if (n == 0) return LBM_OK;
  std::vector<double> rec(size_t(c->np) * kPartFields);
  lbm_status s = get_field(c, c->parts, rec.data(), rec.size(), c->pdbl ? LBM_F64 : LBM_F32);
  if (s) return s;
  for (uint32_t k = 0; k < n; ++k) {
    const double* r = rec.data() + size_t(k) * kPartFields;
    std::memset(&p[k], 0, sizeof(lbm_particle));
    for (int d = 0; d < 3; ++d) {
      p[k].position[d] = r[PF_POS + d];
      p[k].position_old[d] = r[PF_POS_OLD + d];
      p[k].rotation[d] = r[PF_ROT + d];
      p[k].rotation_old[d] = r[PF_ROT_OLD + d];
      p[k].acceleration[d] = r[PF_ACC + d];
      p[k].angular_acceleration[d] = r[PF_ANG_ACC + d];
      p[k].force[d] = r[PF_FORCE + d];
      p[k].torque[d] = r[PF_TORQUE + d];
    }
    p[k].mass = r[PF_MASS];
    p[k].inertia = r[PF_INERTIA];
    p[k].density = r[PF_DENSITY];
    p[k].radius = r[PF_RADIUS];
  }
  return LBM_OK;
}

lbm_status lbm_psm_couple(lbm_ctx* c, uint64_t t) {
  (void)t;
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (c->np == 0) return fail(LBM_ERR_INVALID_STATE, "no particles");
  HIP_TRY(hipSetDevice(c->device));
  lbm_status s = prepare_sweep(c);
  if (s) return s;
  ParticleArgs a = particle_args(c);
  HIP_TRY(launch_psm_couple(c->lattice, c->storage, a, c->stream));
  return LBM_OK;
}

lbm_status lbm_particle_reduce(lbm_ctx* c, double* force, double* torque) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (c->np == 0) return fail(LBM_ERR_INVALID_STATE, "no particles");
  HIP_TRY(hipSetDevice(c->device));
  if (force) HIP_TRY(hipMemcpyAsync(force, c->f64, size_t(c->np) * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (torque) HIP_TRY(hipMemcpyAsync(torque, c->t64, size_t(c->np) * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

lbm_status lbm_dem_contact(lbm_ctx* c) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (c->np == 0) return fail(LBM_ERR_INVALID_STATE, "no particles");
  HIP_TRY(hipSetDevice(c->device));
  ParticleArgs a = particle_args(c);
  HIP_TRY(launch_dem_contact(c->D, c->storage, a, c->stream));
  return LBM_OK;
}

lbm_status lbm_particle_integrate(lbm_ctx* c) {
  if (!c) return fail(LBM_ERR_INVALID_ARG, "null ctx");
  if (c->np == 0) return fail(LBM_ERR_INVALID_STATE, "no particles");
  HIP_TRY(hipSetDevice(c->device));
  ParticleArgs a = particle_args(c);
  HIP_TRY(launch_particle_integrate(c->D, c->storage, a, c->stream));
  return LBM_OK;
}

lbm_status lbm_cell_list_overflow(lbm_ctx* c, uint64_t* overflow) {
  if (!c || !overflow) return fail(LBM_ERR_INVALID_ARG, "null argument");
  *overflow = 0;
  if (!c->cl_overflow) return LBM_OK;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(overflow, c->cl_overflow, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

}  // extern "C"

// ================================================================ batches
struct lbm_batch {
  int lattice = 0, storage = 0, D = 2, device = 0;
  std::vector<lbm_ctx*> members;
  std::vector<char> active;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  uint64_t last_updates = 0;
  int phase = 0;
  BatchShape shape;
  // device work list
  StepArgs* d_members = nullptr;
  Tile* d_tiles = nullptr;
  uint32_t* d_tile_begin = nullptr;
  void** d_uprev = nullptr;
  double* d_partial = nullptr;
  double* d_res = nullptr;
  uint64_t* d_t0 = nullptr;
  std::vector<void*> uprev;
  uint32_t ntiles = 0;
  size_t tiles_cap = 0;
  bool tiles_dirty = true;
  // captured chains: key = (length, with residual)
  hipGraphExec_t chain = nullptr;
  uint64_t chain_len = 0;
  int chain_res = -1;
};

namespace {

StepArgs member_args(const lbm_ctx* c) {
  StepArgs a{};
  a.f = c->f;
  a.flags = c->flags;
  a.ops = c->ops;
  a.rho = c->rho;
  a.u = c->u;
  a.force = c->force;
  a.por = c->por;
  a.bcvel = c->bcvel;
  a.pn = c->pn;
  a.pd = c->pd;
  a.nx = int32_t(c->n[0]);
  a.ny = int32_t(c->n[1]);
  a.nz = int32_t(c->n[2]);
  a.N = c->N;
  a.Np = c->Np;
  a.periodic = c->periodic;
  a.status = c->status_d;
  return a;
}

void drop_chain(lbm_batch* b) {
  // a launch of the chain may still be queued: it must finish before the
  // executable graph goes
  if (b->chain) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(b->chain);
  }
  b->chain = nullptr;
  b->chain_len = 0;
  b->chain_res = -1;
}

// Tiles of the active members, member by member; members' argument records.
lbm_status build_tiles(lbm_batch* b) {
  const uint32_t M = uint32_t(b->members.size());
  std::vector<StepArgs> args(M);
  std::vector<Tile> tiles;
  std::vector<uint32_t> begin(M + 1, 0);
  const int bx = b->shape.block, rows = b->shape.rows;
  for (uint32_t m = 0; m < M; ++m) {
    lbm_ctx* c = b->members[m];
    args[m] = member_args(c);
    begin[m] = uint32_t(tiles.size());
    if (!b->active[m]) continue;
    const uint64_t per = uint64_t(bx) * uint64_t(rows);
    for (uint64_t c0 = 0; c0 < c->N; c0 += per) tiles.push_back(Tile{m, uint32_t(c0)});
  }
  begin[M] = uint32_t(tiles.size());
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (tiles.size() > b->tiles_cap) {
    if (b->d_tiles) HIP_TRY(hipFree(b->d_tiles));
    if (b->d_partial) HIP_TRY(hipFree(b->d_partial));
    b->tiles_cap = std::max<size_t>(tiles.size(), 1);
    HIP_TRY(hipMalloc((void**)&b->d_tiles, b->tiles_cap * sizeof(Tile)));
    HIP_TRY(hipMalloc((void**)&b->d_partial, b->tiles_cap * 2 * sizeof(double)));
  }
  b->ntiles = uint32_t(tiles.size());
  if (!tiles.empty())
    HIP_TRY(hipMemcpyAsync(b->d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_members, args.data(), M * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_tile_begin, begin.data(), (M + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->tiles_dirty = false;
  drop_chain(b);
  return LBM_OK;
}

BatchArgs batch_args(lbm_batch* b, uint32_t k, int write_macros) {
  BatchArgs a{};
  a.members = b->d_members;
  a.tiles = b->d_tiles;
  a.ntiles = b->ntiles;
  a.rows = uint32_t(b->shape.rows);
  a.t0 = b->d_t0;
  a.k = k;
  a.write_macros = write_macros;
  a.u_prev = b->d_uprev;
  return a;
}

// Enqueue a chain of n steps from the phase the batch is in; with_res also
// writes u_prev at step n-2, u at step n-1 and forms the residual sums.
lbm_status enqueue_chain(lbm_batch* b, uint64_t n, bool with_res) {
  int phase = b->phase;
  for (uint64_t k = 0; k < n; ++k) {
    int wm = 0;
    if (k + 1 == n) wm = 1;
    if (with_res && k + 2 == n) wm = 2;
    HIP_TRY(launch_aa_batch(b->lattice, b->storage, phase == 1, batch_args(b, uint32_t(k), wm), b->shape, b->stream));
    phase ^= 1;
  }
  if (with_res)
    HIP_TRY(launch_batch_residual(b->D, b->storage, batch_args(b, 0, 0), b->shape, b->d_partial,
                                  b->d_tile_begin, uint32_t(b->members.size()), b->d_res, b->stream));
  return LBM_OK;
}

// Run n steps starting at step t0 through a captured graph (rebuilt when the
// chain length, the residual flag, the phase parity or the work list change).
// `start` (if given) is recorded after any capture, right before the chain's
// work, so a timed region never holds the host's capture and instantiation.
lbm_status run_chain(lbm_batch* b, uint64_t t0, uint64_t n, bool with_res, hipEvent_t start = nullptr) {
  if (b->tiles_dirty) {
    lbm_status s = build_tiles(b);
    if (s) return s;
  }
  const bool graphable = (n % 2 == 0) && b->phase == 0;
  if (graphable && (!b->chain || b->chain_len != n || b->chain_res != int(with_res))) {
    drop_chain(b);
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    lbm_status s = enqueue_chain(b, n, with_res);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (s) return s;
    HIP_TRY(e);
    HIP_TRY(hipGraphInstantiate(&b->chain, g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(g));
    b->chain_len = n;
    b->chain_res = int(with_res);
  }
  if (start) HIP_TRY(hipEventRecord(start, b->stream));
  HIP_TRY(launch_set_u64(b->d_t0, t0, b->stream));
  if (!graphable) {
    lbm_status s = enqueue_chain(b, n, with_res);
    if (s) return s;
  } else {
    HIP_TRY(hipGraphLaunch(b->chain, b->stream));
  }
  if (n % 2) b->phase ^= 1;
  for (lbm_ctx* c : b->members) c->phase = b->phase;
  return LBM_OK;
}

lbm_status batch_ready(lbm_batch* b) {
  for (lbm_ctx* c : b->members) {
    lbm_status s = check_ready(c);
    if (s) return s;
    if (c->np > 0) return fail(LBM_ERR_NOT_SUPPORTED, "particles in a batch member");
    if (c->phase != b->phase)
      return fail(LBM_ERR_INVALID_STATE, "a member's populations were reset; reset all members together");
  }
  return LBM_OK;
}

}  // namespace

extern "C" {

lbm_status lbm_batch_create(lbm_batch** out, int32_t lattice, int32_t storage, uint32_t count,
                            const uint64_t* extents, int32_t device) {
  if (!out || !extents || count == 0) return fail(LBM_ERR_INVALID_ARG, "bad batch arguments");
  *out = nullptr;
  lbm_batch* b = new lbm_batch();
  b->lattice = lattice;
  b->storage = storage;
  b->D = dims_of(lattice);
  b->device = device;
  auto bail = [&](lbm_status s) {
    lbm_batch_destroy(b);
    return s;
  };
  for (uint32_t m = 0; m < count; ++m) {
    lbm_ctx* c = nullptr;
    lbm_status s = lbm_ctx_create(&c, lattice, storage, extents + 3 * size_t(m), device);
    if (s) return bail(s);
    b->members.push_back(c);
    b->active.push_back(1);
    void* up = nullptr;
    hipError_t e = alloc_zero(&up, size_t(c->D) * c->N * c->tsz(), c->stream);
    if (e != hipSuccess) return bail(fail(LBM_ERR_OUT_OF_MEMORY, "u_prev: %s", hipGetErrorString(e)));
    b->uprev.push_back(up);
    (void)hipStreamSynchronize(c->stream);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(LBM_ERR_CRITICAL, "stream/event creation"));
  hipError_t e = hipMalloc((void**)&b->d_members, count * sizeof(StepArgs));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_tile_begin, (count + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_uprev, count * sizeof(void*));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_res, 2 * count * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_t0, sizeof(uint64_t));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_uprev, b->uprev.data(), count * sizeof(void*), hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(fail(LBM_ERR_OUT_OF_MEMORY, "batch arrays: %s", hipGetErrorString(e)));
  *out = b;
  return LBM_OK;
}

lbm_status lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_chain(b);
  for (void* p : b->uprev) (void)hipFree(p);
  for (lbm_ctx* c : b->members) lbm_ctx_destroy(c);
  void* ptrs[] = {b->d_members, b->d_tiles, b->d_tile_begin, b->d_uprev, b->d_partial, b->d_res, b->d_t0};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

lbm_ctx* lbm_batch_member(lbm_batch* b, uint32_t i) {
  if (!b || i >= b->members.size()) return nullptr;
  // set-up through the member may change pointers (lazily allocated fields)
  b->tiles_dirty = true;
  return b->members[i];
}

lbm_status lbm_batch_step(lbm_batch* b, uint64_t t0, uint64_t nsteps) {
  if (!b) return fail(LBM_ERR_INVALID_ARG, "null batch");
  std::fill(b->active.begin(), b->active.end(), 1);
  b->tiles_dirty = true;
  lbm_status s = batch_ready(b);
  if (s) return s;
  HIP_TRY(hipSetDevice(b->device));
  if (b->tiles_dirty) {
    s = build_tiles(b);
    if (s) return s;
  }
  s = run_chain(b, t0, nsteps, false, b->ev0);
  if (s) return s;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  uint64_t cells = 0;
  for (lbm_ctx* c : b->members) cells += c->N;
  b->last_updates = cells * nsteps;
  return LBM_OK;
}

lbm_status lbm_batch_run(lbm_batch* b, uint64_t t0, double tol, uint64_t check_every,
                         uint64_t max_steps, uint64_t* steps_out, double* res_out) {
  if (!b) return fail(LBM_ERR_INVALID_ARG, "null batch");
  if (check_every < 2 || check_every % 2) return fail(LBM_ERR_INVALID_ARG, "check_every must be even and >= 2");
  if (b->phase != 0) return fail(LBM_ERR_INVALID_STATE, "batch runs start from the natural layout");
  std::fill(b->active.begin(), b->active.end(), 1);
  b->tiles_dirty = true;
  lbm_status s = batch_ready(b);
  if (s) return s;
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t M = uint32_t(b->members.size());
  std::vector<uint64_t> steps(M, 0);
  std::vector<double> res(M, 0.0), sums(2 * M);
  uint64_t done = 0, updates = 0;
  float total_ms = 0.f;
  while (done < max_steps) {
    const uint64_t n = std::min<uint64_t>(check_every, max_steps - done) & ~uint64_t(1);
    if (n < 2) break;
    uint64_t cells = 0;
    for (uint32_t m = 0; m < M; ++m)
      if (b->active[m]) cells += b->members[m]->N;
    s = run_chain(b, t0 + done, n, true, b->ev0);
    if (s) return s;
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipMemcpyAsync(sums.data(), b->d_res, 2 * M * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    total_ms += ms;
    updates += cells * n;
    done += n;
    bool changed = false, any = false;
    for (uint32_t m = 0; m < M; ++m) {
      if (!b->active[m]) continue;
      steps[m] += n;
      // a diverged member (non-finite sums) reports NaN, never "converged"
      const double num = sums[2 * m], den = sums[2 * m + 1];
      const double r = !(std::isfinite(num) && std::isfinite(den)) ? std::nan("")
                       : (den > 0 ? std::sqrt(num / den) : 0.0);
      res[m] = r;
      if (!(r >= tol)) {  // converged (or NaN: stop and report it)
        b->active[m] = 0;
        changed = true;
      } else {
        any = true;
      }
    }
    if (!any) break;
    if (changed) b->tiles_dirty = true;
  }
  b->last_ms = total_ms;
  b->last_updates = updates;
  if (steps_out) std::copy(steps.begin(), steps.end(), steps_out);
  if (res_out) std::copy(res.begin(), res.end(), res_out);
  std::fill(b->active.begin(), b->active.end(), 1);
  b->tiles_dirty = true;
  return LBM_OK;
}

lbm_status lbm_batch_set_tuning(lbm_batch* b, const lbm_batch_tuning* t) {
  if (!b || !t) return fail(LBM_ERR_INVALID_ARG, "null argument");
  if (t->block < 64 || t->block > 256 || (t->block & 63) || t->rows < 1 || t->rows > 16 || t->grid < 0)
    return fail(LBM_ERR_INVALID_ARG, "tuning out of range");
  b->shape.block = t->block;
  b->shape.rows = t->rows;
  b->shape.grid = t->grid;
  b->tiles_dirty = true;
  return LBM_OK;
}

lbm_status lbm_batch_autotune(lbm_batch* b, lbm_batch_tuning* chosen) {
  if (!b) return fail(LBM_ERR_INVALID_ARG, "null batch");
  lbm_status s = batch_ready(b);
  if (s) return s;
  if (b->phase != 0) return fail(LBM_ERR_INVALID_STATE, "autotune from the natural layout");
  HIP_TRY(hipSetDevice(b->device));
  // keep the populations and macros: trials advance the state
  std::vector<void*> saved;
  auto restore = [&]() -> lbm_status {
    size_t k = 0;
    for (lbm_ctx* c : b->members) {
      const size_t fb = size_t(c->Q) * c->Np * storage_bytes(c->storage);
      HIP_TRY(hipMemcpyAsync(c->f, saved[k++], fb, hipMemcpyDeviceToDevice, b->stream));
      HIP_TRY(hipMemcpyAsync(c->u, saved[k++], size_t(c->D) * c->N * c->tsz(), hipMemcpyDeviceToDevice, b->stream));
      HIP_TRY(hipMemcpyAsync(c->rho, saved[k++], c->N * c->tsz(), hipMemcpyDeviceToDevice, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->phase = 0;
    for (lbm_ctx* c : b->members) c->phase = 0;
    return LBM_OK;
  };
  for (lbm_ctx* c : b->members) {
    const size_t sizes[3] = {size_t(c->Q) * c->Np * storage_bytes(c->storage), size_t(c->D) * c->N * c->tsz(),
                             c->N * c->tsz()};
    void* srcs[3] = {c->f, c->u, c->rho};
    for (int k = 0; k < 3; ++k) {
      void* p = nullptr;
      HIP_TRY(hipMalloc(&p, sizes[k]));
      HIP_TRY(hipMemcpyAsync(p, srcs[k], sizes[k], hipMemcpyDeviceToDevice, b->stream));
      saved.push_back(p);
    }
  }
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  const BatchShape keep = b->shape;
  BatchShape best = keep;
  double best_ms = 1e30;
  const int blocks[3] = {64, 128, 256}, rows_c[3] = {1, 2, 4}, grids[4] = {0, 4, 8, 16};
  for (int bi = 0; bi < 3; ++bi)
    for (int ri = 0; ri < 3; ++ri)
      for (int gi = 0; gi < 4; ++gi) {
        b->shape.block = blocks[bi];
        b->shape.rows = rows_c[ri];
        b->shape.grid = grids[gi] == 0 ? 0 : grids[gi] * cus;
        b->tiles_dirty = true;
        s = build_tiles(b);
        if (s) return s;
        // warm-up with the timed chain's length: the timed call then replays
        // the graph this one captured (no capture inside the timed region)
        s = run_chain(b, 0, 8, false);
        if (s) return s;
        s = run_chain(b, 8, 8, false, b->ev0);
        if (s) return s;
        HIP_TRY(hipEventRecord(b->ev1, b->stream));
        HIP_TRY(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        if (ms / 8.0 < best_ms) {
          best_ms = ms / 8.0;
          best = b->shape;
        }
      }
  s = restore();
  if (s) return s;
  for (void* p : saved) (void)hipFree(p);
  b->shape = best;
  b->tiles_dirty = true;
  if (chosen) {
    chosen->block = best.block;
    chosen->rows = best.rows;
    chosen->grid = best.grid;
    chosen->ms_per_step = best_ms;
  }
  return LBM_OK;
}

lbm_status lbm_batch_stats(lbm_batch* b, double* ms, uint64_t* updates) {
  if (!b) return fail(LBM_ERR_INVALID_ARG, "null batch");
  if (ms) *ms = b->last_ms;
  if (updates) *updates = b->last_updates;
  return LBM_OK;
}

}  // extern "C"
