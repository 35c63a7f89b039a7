// Particle coupling kernels.
//
// One PSM coupling pass (lbm_psm_couple) is, in launch order:
//   k_reset_cells porosity := 1 (PsmReset, psm.hpp:20-32), owners := 0, HLBM N, D := 0,
//                FPLBM force := 0
//   k_owner      [more than kLocalOwnerMax particles] per particle and AABB
//                cell: which particle a serial loop over particles would have
//                written last (highest index), for the porosity and for the
//                force field; with fewer particles each gather block decides
//                this itself from the boxes that overlap its particle's
//   k_gather     per particle and AABB cell: porosity, momentum exchange
//                (psm.hpp:304-424), owner writes, and the particle's force and
//                torque summed in FP64 (wavefront shuffles, then LDS, in a
//                fixed order per block)
//   k_finalize   per particle: fixed-order sum of the block partials; force
//                reset + accumulate (integrate.hpp:195-212, psm.hpp:442-443);
//                inside lbm_step the contact kernel does this for its particle
//                instead (fuse_finalize), one launch fewer per step
//   [product mode] k_build_cells + k_product: cell list and the per-cell
//                product porosity (cell_list.hpp:81-150, couple.hpp:26-60)
// Then k_contact (contact.hpp:466-567) and k_integrate (integrate.hpp:45-83);
// inside lbm_step the contact kernel integrates its particle itself
// (fuse_integrate), reading its partners from the records the gather saved.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace lbmx {

// couplings, as gather template arguments
constexpr int kCplPsm = 0, kCplHlbm = 1, kCplFplbm = 2;

template <class S, class T>
__device__ inline T ldp(const S* p, T w);
template <>
__device__ inline double ldp<double, double>(const double* p, double) { return *p; }
template <>
__device__ inline float ldp<float, float>(const float* p, float) { return *p; }
template <>
__device__ inline float ldp<__half, float>(const __half* p, float w) { return __half2float(*p) + w; }
template <>
__device__ inline float ldp<__hip_bfloat16, float>(const __hip_bfloat16* p, float w) {
  return __bfloat162float(*p) + w;
}
// FP64 particle state over FP32-arithmetic storage: the FP32 value, widened
template <>
__device__ inline double ldp<float, double>(const float* p, double) { return double(*p); }
template <>
__device__ inline double ldp<__half, double>(const __half* p, double w) {
  return double(ldp<__half, float>(p, float(w)));
}
template <>
__device__ inline double ldp<__hip_bfloat16, double>(const __hip_bfloat16* p, double w) {
  return double(ldp<__hip_bfloat16, float>(p, float(w)));
}

// wrap a coordinate along a periodic axis; false when it leaves the lattice
__device__ inline bool wrap(int& v, int n, bool periodic) {
  if (v >= 0 && v < n) return true;
  if (!periodic) return false;
  v = (v + n) % n;
  return true;
}

// Pre-collision population f_i at (x,y,z) in the current AA phase (phase 1:
// plane opp(i) at x - c_i, 0 outside the lattice), or 0 when !ok.  The load is
// issued whatever the case (an invalid neighbour reads the cell x, y, z
// itself, which lies in the lattice): no branch around the load, so a caller's
// loads for all directions can be in flight together.
template <class L, class S, class T>
__device__ __forceinline__ T pop_sel(const ParticleArgs& a, int i, int x, int y, int z, bool ok) {
  const S* f = static_cast<const S*>(a.f);
  const uint32_t nx = uint32_t(a.nx), ny = uint32_t(a.ny);
  uint32_t c = uint32_t(x) + nx * (uint32_t(y) + ny * uint32_t(z));
  int plane = i;
  if (a.phase != 0) {
    int xs = x - L::cx[i], ys = y - L::cy[i], zs = z - L::cz[i];
    ok = ok && wrap(xs, a.nx, a.periodic & 1u) && wrap(ys, a.ny, a.periodic & 2u) &&
         wrap(zs, a.nz, a.periodic & 4u);
    if (ok) c = uint32_t(xs) + nx * (uint32_t(ys) + ny * uint32_t(zs));
    plane = L::opp[i];
  }
  const T v = ldp<S, T>(f + size_t(plane) * a.Np + c, T(L::w[i]));
  return ok ? v : T(0);
}

template <class T, int D>
struct Box {
  int lo[3], hi[3];
  uint32_t count;
};

// particle_aabb::try_calculate (aabb.hpp:64-81)
template <class T, int D>
__device__ inline Box<T, D> aabb_of(const T* P, const ParticleArgs& a) {
  Box<T, D> b;
  const int ext[3] = {a.nx, a.ny, a.nz};
  const T r = P[PF_RADIUS];
  b.count = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      const T pos = P[PF_POS + d];
      const T lower = (pos >= r) ? T(pos - r) : T(0);
      const T upper = pos + r;
      const T up1 = upper + T(1);
      // the reference casts to uint64; positions are inside the lattice here
      uint64_t lo = lower > T(0) ? uint64_t(lower) : 0;
      uint64_t hi = up1 > T(0) ? uint64_t(up1) : 0;
      if (hi > uint64_t(ext[d])) hi = uint64_t(ext[d]);
      if (lo > hi) lo = hi;
      b.lo[d] = int(lo);
      b.hi[d] = int(hi);
    } else {
      b.lo[d] = 0;
      b.hi[d] = 1;
    }
    b.count *= uint32_t(b.hi[d] - b.lo[d]);
  }
  return b;
}

// The k-th cell of the box, x fastest: neighbouring lanes take neighbouring
// cells of one lattice row, so the field and population accesses of a
// wavefront fall in runs of contiguous addresses (the box has no order of its
// own: the force sums are FP64 and the owners are maxima).
template <class T, int D>
__device__ inline void box_cell(const Box<T, D>& b, uint32_t k, int* x) {
  const uint32_t sx = uint32_t(b.hi[0] - b.lo[0]);
  const uint32_t sy = uint32_t(b.hi[1] - b.lo[1]);
  x[0] = b.lo[0] + int(k % sx);
  k /= sx;
  x[1] = b.lo[1] + int(k % sy);
  x[2] = b.lo[2] + int(k / sy);
}

template <class T, int D>
__device__ inline T porosity_at(const T* P, const int* x, T eps, T* rel) {
  for (int d = 0; d < D; ++d) rel[d] = T(x[d]) - P[PF_POS + d];
  return disc_porosity<T, D>(rel, P[PF_RADIUS], eps);
}



// PsmReset / HlbmReset / FpLbmReset over every cell (psm.hpp:20-32,
// hlbm.hpp:28-49, fplbm.hpp:21-58): porosity 1 (scatter mode), owners 0, and
// where given N = 0, D = 0, force = 0
template <class F>
__global__ void k_reset_cells(F* por, uint32_t* owner, F* pn, F* pd, F* force, uint32_t N, int dims,
                              int reset_por) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  if (reset_por) por[c] = F(1);
  if (owner) {
    owner[2 * c] = 0;
    owner[2 * c + 1] = 0;
  }
  if (pn) {
    for (int d = 0; d < dims; ++d) pn[size_t(d) * N + c] = F(0);
    pd[c] = F(0);
  }
  if (force)
    for (int d = 0; d < dims; ++d) force[size_t(d) * N + c] = F(0);
}


// LDS entries of a block's list of overlapping particles
constexpr int kOverlapCap = 256;

template <class T, int D>
__device__ __forceinline__ bool box_has(const Box<T, D>& q, const int* x) {
  for (int d = 0; d < D; ++d)
    if (x[d] < q.lo[d] || x[d] >= q.hi[d]) return false;
  return true;
}

// HlbmOneParticleStamp (hlbm.hpp:264-308): over the post-move AABB,
// D += 1 - B and N += u_p (1 - B) with B the porosity of this step's gather
// and u_p = (X - X_old) / ts after the move.  The reference stamps particle
// after particle, so a cell inside several post-move boxes sums their terms
// in particle order.  Up to kLocalOwnerMax particles each block lists the
// particles whose post-move boxes overlap its own (LDS, sorted by index; a
// block whose list overflows tests them all); a cell is written only by the
// lowest-index particle whose box holds it, which adds every holder's term in
// index order to the reset value: the reference's sums, bit for bit, with
// plain stores.  Above kLocalOwnerMax particles the terms are added
// atomically (order not fixed where boxes overlap).
template <class T, int D>
__device__ __forceinline__ void stamp_velocity(const T* P, T ts, T* vp) {
  for (int d = 0; d < D; ++d) vp[d] = (P[PF_POS + d] - P[PF_POS_OLD + d]) / ts;
}

template <class T, class F, int D>
__global__ __launch_bounds__(256) void k_hlbm_stamp(ParticleArgs a) {
  const uint32_t p = blockIdx.x / a.blocks_per_particle;
  const uint32_t chunk = blockIdx.x % a.blocks_per_particle;
  if (p >= a.n) return;  // uniform per block
  const T* parts = static_cast<const T*>(a.parts);
  const T* P = parts + size_t(p) * kPartFields;
  const Box<T, D> b = aabb_of<T, D>(P, a);
  const T ts = T(1.0) / T(a.sub_steps == 0 ? uint64_t(1) : a.sub_steps);
  T vp[3] = {T(0), T(0), T(0)};
  stamp_velocity<T, D>(P, ts, vp);
  const F* por = static_cast<const F*>(a.por);
  F* pn = static_cast<F*>(a.pn);
  F* pd = static_cast<F*>(a.pd);
  const bool ordered = a.n <= kLocalOwnerMax;  // uniform
  __shared__ int cnt;
  __shared__ uint32_t lq[kOverlapCap];
  __shared__ int lbox[kOverlapCap][6];
  if (ordered) {
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < a.n; j += blockDim.x) {
      if (j == p) continue;
      const Box<T, D> o = aabb_of<T, D>(parts + size_t(j) * kPartFields, a);
      bool hit = o.count > 0;
      for (int d = 0; d < D; ++d) hit = hit && o.lo[d] < b.hi[d] && b.lo[d] < o.hi[d];
      if (!hit) continue;
      const int e = atomicAdd(&cnt, 1);
      if (e < kOverlapCap) {
        lq[e] = j;
        for (int d = 0; d < 3; ++d) {
          lbox[e][d] = o.lo[d];
          lbox[e][3 + d] = o.hi[d];
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // by particle index (the lists are short)
      const int m = min(cnt, kOverlapCap);
      for (int u = 1; u < m; ++u)
        for (int v = u; v > 0 && lq[v - 1] > lq[v]; --v) {
          const uint32_t tq = lq[v];
          lq[v] = lq[v - 1];
          lq[v - 1] = tq;
          for (int d = 0; d < 6; ++d) {
            const int tb = lbox[v][d];
            lbox[v][d] = lbox[v - 1][d];
            lbox[v - 1][d] = tb;
          }
        }
    }
    __syncthreads();
  }
  const int listed = ordered ? cnt : 0;
  for (uint32_t k = chunk * blockDim.x + threadIdx.x; k < b.count; k += a.blocks_per_particle * blockDim.x) {
    int x[3];
    box_cell<T, D>(b, k, x);
    const uint32_t c = uint32_t(x[0]) + uint32_t(a.nx) * (uint32_t(x[1]) + uint32_t(a.ny) * uint32_t(x[2]));
    const F flip = F(1.0) - por[c];
    if (flip == F(0)) continue;
    if (!ordered) {
      atomicAdd(pd + c, flip);
      for (int d = 0; d < D; ++d) atomicAdd(pn + size_t(d) * a.N + c, F(vp[d]) * flip);
      continue;
    }
    // the cell's holders in index order: p first unless a lower index holds it
    F sd = pd[c];
    F sn[3] = {F(0), F(0), F(0)};
    for (int d = 0; d < D; ++d) sn[d] = pn[size_t(d) * a.N + c];
    bool first = true, mine_added = false;
    auto add = [&](const T* v) {
      sd = sd + flip;
      for (int d = 0; d < D; ++d) sn[d] = sn[d] + F(v[d]) * flip;
    };
    if (listed <= kOverlapCap) {
      for (int e = 0; e < listed && first; ++e) {
        bool in = true;
        for (int d = 0; d < D; ++d) in = in && x[d] >= lbox[e][d] && x[d] < lbox[e][3 + d];
        if (!in) continue;
        if (lq[e] < p) {
          first = false;
          break;
        }
        if (!mine_added) {
          add(vp);
          mine_added = true;
        }
        T vq[3] = {T(0), T(0), T(0)};
        stamp_velocity<T, D>(parts + size_t(lq[e]) * kPartFields, ts, vq);
        add(vq);
      }
    } else {
      for (uint32_t j = 0; j < a.n && first; ++j) {
        if (j == p) {
          add(vp);
          mine_added = true;
          continue;
        }
        const T* Pj = parts + size_t(j) * kPartFields;
        if (!box_has<T, D>(aabb_of<T, D>(Pj, a), x)) continue;
        if (j < p) {
          first = false;
          break;
        }
        T vq[3] = {T(0), T(0), T(0)};
        stamp_velocity<T, D>(Pj, ts, vq);
        add(vq);
      }
    }
    if (!first) continue;
    if (!mine_added) add(vp);
    pd[c] = sd;
    for (int d = 0; d < D; ++d) pn[size_t(d) * a.N + c] = sn[d];
  }
}

template <class L, class T>
__global__ __launch_bounds__(256) void k_owner(ParticleArgs a) {
  constexpr int D = L::D;
  const uint32_t p = blockIdx.x / a.blocks_per_particle;
  const uint32_t chunk = blockIdx.x % a.blocks_per_particle;
  if (p >= a.n) return;
  const T* P = static_cast<const T*>(a.parts) + size_t(p) * kPartFields;
  const Box<T, D> b = aabb_of<T, D>(P, a);
  const T eps = T(a.eps);
  for (uint32_t k = chunk * blockDim.x + threadIdx.x; k < b.count;
       k += a.blocks_per_particle * blockDim.x) {
    int x[3];
    box_cell<T, D>(b, k, x);
    T rel[3];
    const T B = porosity_at<T, D>(P, x, eps, rel);
    const uint32_t c = uint32_t(x[0]) + uint32_t(a.nx) * (uint32_t(x[1]) + uint32_t(a.ny) * uint32_t(x[2]));
    atomicMax(a.owner + 2 * c, p + 1);
    if (B < T(1.0)) atomicMax(a.owner + 2 * c + 1, p + 1);
  }
}

// Six FP64 sums over the 64-lane wavefront by one xor butterfly: each step
// exchanges all six values before adding, so the six chains share their
// shuffle round trips.  Every lane ends with the same bits (a + b = b + a).
__device__ inline void wave_sum6(double (&v)[6]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double t[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) t[k] = __shfl_xor(v[k], o, 64);
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += t[k];
  }
}

// Ownership without the per-cell owner array (a.owner == nullptr, at most
// kLocalOwnerMax particles): a cell of particle p's box takes p's porosity if
// no particle q > p has the cell in its box, and p's force if p is solid there
// (B < 1) and no q > p is solid there -- what the atomic maxima of k_owner
// decide.  Each gather block first lists the q > p whose boxes intersect p's
// (LDS, kOverlapCap entries); a block whose list overflows tests every q > p.
template <class T, int D>
struct LocalOwners {
  int* cnt;        // LDS: entries found (may exceed kOverlapCap)
  uint32_t* q;     // LDS: particle indices
  int (*box)[6];   // LDS: their boxes (lo, hi)

  // every thread of the block calls this once; all: every other particle
  // whose box overlaps b (the sweep), else only those of higher index
  __device__ void build(const ParticleArgs& a, uint32_t p, const Box<T, D>& b, bool all = false) const {
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
    for (uint32_t j = (all ? 0u : p + 1) + threadIdx.x; j < a.n; j += blockDim.x) {
      if (j == p) continue;
      const Box<T, D> o = aabb_of<T, D>(static_cast<const T*>(a.parts) + size_t(j) * kPartFields, a);
      bool hit = o.count > 0;
      for (int d = 0; d < D; ++d) hit = hit && o.lo[d] < b.hi[d] && b.lo[d] < o.hi[d];
      if (hit) {
        const int e = atomicAdd(cnt, 1);
        if (e < kOverlapCap) {
          q[e] = j;
          for (int d = 0; d < 3; ++d) {
            box[e][d] = o.lo[d];
            box[e][3 + d] = o.hi[d];
          }
        }
      }
    }
    __syncthreads();
  }

  // (p owns the cell's porosity, p owns the cell's force); solid: p's B < 1
  __device__ void test(const ParticleArgs& a, uint32_t p, const int* x, bool solid, T eps, bool& own_box,
                       bool& own_solid) const {
    own_box = true;
    own_solid = solid;
    const int n = *cnt;
    if (n <= kOverlapCap) {
      for (int e = 0; e < n; ++e) {
        if (q[e] < p) continue;  // (listed by a sweep: ownership goes to higher indices)
        bool in = true;
        for (int d = 0; d < D; ++d) in = in && x[d] >= box[e][d] && x[d] < box[e][3 + d];
        if (!in) continue;
        own_box = false;
        if (own_solid) {
          T rel[3];
          const T* Pq = static_cast<const T*>(a.parts) + size_t(q[e]) * kPartFields;
          if (porosity_at<T, D>(Pq, x, eps, rel) < T(1.0)) own_solid = false;
        }
        if (!own_solid) return;
      }
      return;
    }
    for (uint32_t j = p + 1; j < a.n; ++j) {
      const T* Pq = static_cast<const T*>(a.parts) + size_t(j) * kPartFields;
      if (!box_has<T, D>(aabb_of<T, D>(Pq, a), x)) continue;
      own_box = false;
      if (own_solid) {
        T rel[3];
        if (porosity_at<T, D>(Pq, x, eps, rel) < T(1.0)) own_solid = false;
      }
      if (!own_solid) return;
    }
  }

  // whether another particle's box holds the cell (after build(..., true)
  // over a region that contains it)
  __device__ bool held(const ParticleArgs& a, uint32_t p, const int* x) const {
    const int n = *cnt;
    if (n <= kOverlapCap) {
      for (int e = 0; e < n; ++e) {
        bool in = true;
        for (int d = 0; d < D; ++d) in = in && x[d] >= box[e][d] && x[d] < box[e][3 + d];
        if (in) return true;
      }
      return false;
    }
    for (uint32_t j = 0; j < a.n; ++j)
      if (j != p && box_has<T, D>(aabb_of<T, D>(static_cast<const T*>(a.parts) + size_t(j) * kPartFields, a), x))
        return true;
    return false;
  }
};

// S: population storage, T: particle arithmetic, F: fluid field type;
// CPL: the coupling (kCplPsm, kCplHlbm, kCplFplbm), LOCAL: ownership decided
// per block (no owner array)
template <class L, class S, class T, class F, int CPL, bool LOCAL>
__global__ __launch_bounds__(256) void k_gather(ParticleArgs a) {
  constexpr int D = L::D, Q = L::Q;
  constexpr bool hlbm = CPL == kCplHlbm, fplbm = CPL == kCplFplbm;
  const uint32_t p = blockIdx.x / a.blocks_per_particle;
  const uint32_t chunk = blockIdx.x % a.blocks_per_particle;
  if (p >= a.n) return;  // uniform per block
  const T* P = static_cast<const T*>(a.parts) + size_t(p) * kPartFields;
  if (a.fuse_integrate && chunk == 0 && threadIdx.x < kPartFields)
    static_cast<T*>(a.snap)[size_t(p) * kPartFields + threadIdx.x] = P[threadIdx.x];
  const Box<T, D> b = aabb_of<T, D>(P, a);
  const T eps = T(a.eps);
  const T one = T(1.0);
  // ts = 1/sub_steps, u_p = (X − X_old)/ts (psm.hpp:160-172)
  // PSM divides by sub_steps as given (psm.hpp:160-172); HLBM takes 0 as 1
  const T ts = one / T(hlbm && a.sub_steps == 0 ? uint64_t(1) : a.sub_steps);
  // body velocity: (X - X_old)/ts, or (X - X_old)*sub_steps for FPLBM (fplbm.hpp:198-200)
  T vp[3] = {T(0), T(0), T(0)};
  for (int d = 0; d < D; ++d)
    vp[d] = fplbm ? (P[PF_POS + d] - P[PF_POS_OLD + d]) * T(a.sub_steps) : (P[PF_POS + d] - P[PF_POS_OLD + d]) / ts;
  F* por = static_cast<F*>(a.por);
  F* fout = static_cast<F*>(a.force);
  const bool scatter = a.porosity_mode == LBM_POROSITY_SCATTER;
  __shared__ int lo_cnt;
  __shared__ uint32_t lo_q[kOverlapCap];
  __shared__ int lo_box[kOverlapCap][6];
  const LocalOwners<T, D> lo{&lo_cnt, lo_q, lo_box};
  // the sweep (PSM scatter porosity without the per-cell reset): the cells of
  // this gather's box and of the previous gather's box, whose cells outside
  // every current box go back to porosity 1
  constexpr bool kCanSweep = CPL == kCplPsm && LOCAL;
  const bool sweep = kCanSweep && a.boxes != nullptr && scatter;
  Box<T, D> bp, R = b;
  if (kCanSweep && sweep) {
    const int32_t* pb = a.boxes + size_t(p) * 12 + 6;
    bp.count = 1;
    for (int d = 0; d < 3; ++d) {
      bp.lo[d] = pb[d];
      bp.hi[d] = pb[3 + d];
      bp.count *= uint32_t(bp.hi[d] - bp.lo[d]);
    }
    if (b.count == 0) {
      R = bp;
    } else if (bp.count > 0) {
      R.count = 1;
      for (int d = 0; d < 3; ++d) {
        R.lo[d] = min(b.lo[d], bp.lo[d]);
        R.hi[d] = max(b.hi[d], bp.hi[d]);
        R.count *= uint32_t(R.hi[d] - R.lo[d]);
      }
    }
    // (constant indices only: a box indexed per lane would be placed in LDS,
    // one copy per thread, halving the gather's occupancy)
    if (chunk == 0 && threadIdx.x == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a.boxes[size_t(p) * 12 + d] = b.lo[d];
        a.boxes[size_t(p) * 12 + 3 + d] = b.hi[d];
      }
    }
  }
  if constexpr (LOCAL) lo.build(a, p, R, sweep);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t k = chunk * blockDim.x + threadIdx.x; k < R.count;
       k += a.blocks_per_particle * blockDim.x) {
    int x[3];
    box_cell<T, D>(R, k, x);
    if (kCanSweep && sweep && !box_has<T, D>(b, x)) {
      if (box_has<T, D>(bp, x) && !lo.held(a, p, x))
        por[uint32_t(x[0]) + uint32_t(a.nx) * (uint32_t(x[1]) + uint32_t(a.ny) * uint32_t(x[2]))] = F(1);
      continue;
    }
    T rel[3] = {T(0), T(0), T(0)};
    const T B = porosity_at<T, D>(P, x, eps, rel);
    const uint32_t c = uint32_t(x[0]) + uint32_t(a.nx) * (uint32_t(x[1]) + uint32_t(a.ny) * uint32_t(x[2]));
    bool own_box, own_solid;
    if constexpr (LOCAL) {
      lo.test(a, p, x, !fplbm && B < T(1.0), eps, own_box, own_solid);
    } else {
      own_box = a.owner[2 * c] == p + 1;
      own_solid = !fplbm && B < T(1.0) && a.owner[2 * c + 1] == p + 1;
    }
    if (scatter && own_box) por[c] = F(B);
    if constexpr (fplbm) {
      // FpLbmOneParticle (fplbm.hpp:204-224): m of a flag-2 cell from its
      // populations (FpLbmReset's compute_rho_momentum), else the stored macro
      T mom[3] = {T(0), T(0), T(0)};
      if (a.cflags[c] == 2) {
        T fv[Q];
#pragma unroll
        for (int j = 1; j < Q; ++j) fv[j] = pop_sel<L, S, T>(a, j, x[0], x[1], x[2], true);
#pragma unroll
        for (int j = 1; j < Q; ++j) {
          const T fj = fv[j];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            if (cdir<L>(j, d) == 1) mom[d] = mom[d] + fj;
            if (cdir<L>(j, d) == -1) mom[d] = mom[d] - fj;
          }
        }
      } else {
        const F* um = static_cast<const F*>(a.umac);
        for (int d = 0; d < D; ++d) mom[d] = T(um[size_t(d) * a.N + c]);
      }
      const T flip = one - B;
      T fc[3] = {T(0), T(0), T(0)};
      for (int d = 0; d < D; ++d) fc[d] = (vp[d] - mom[d]) * flip;
      if (own_box)
        for (int d = 0; d < D; ++d) fout[size_t(d) * a.N + c] = F(fc[d]);
      for (int d = 0; d < D; ++d) acc[d] -= double(fc[d]);
      continue;
    }
    if (B >= T(1.0)) continue;
    // every f_i at x and f_opp(i) at x + c_i first, all loads in flight at once
    T fi[Q], fo[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      fi[i] = pop_sel<L, S, T>(a, i, x[0], x[1], x[2], true);
      int xn = x[0] + L::cx[i], yn = x[1] + L::cy[i], zn = x[2] + L::cz[i];
      const bool ok = wrap(xn, a.nx, a.periodic & 1u) && wrap(yn, a.ny, a.periodic & 2u) &&
                      wrap(zn, a.nz, a.periodic & 4u);
      fo[i] = pop_sel<L, S, T>(a, L::opp[i], ok ? xn : x[0], ok ? yn : x[1], ok ? zn : x[2], ok);
    }
    T mom[3] = {T(0.0), T(0.0), T(0.0)};
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      T upe = T(0);
#pragma unroll
      for (int d = 0; d < D; ++d) upe = upe + T(cdir<L>(i, d)) * vp[d];
      const T added = fi[i] * (T(1.0) - upe) + fo[i] * (T(1.0) + upe);
#pragma unroll
      for (int d = 0; d < D; ++d) mom[d] = mom[d] + T(cdir<L>(i, d)) * added;
    }
    // PSM scales the exchange by the solid fraction; HLBM keeps it raw (hlbm.hpp:229-231)
    const T flip = hlbm ? one : one - B;
    T fc[3] = {T(0), T(0), T(0)};
    for (int d = 0; d < D; ++d) fc[d] = mom[d] * flip;
    if (own_solid)
      for (int d = 0; d < D; ++d) fout[size_t(d) * a.N + c] = F(fc[d]);
    for (int d = 0; d < D; ++d) acc[d] -= double(fc[d]);
    if (D == 2) {
      acc[3] -= double(rel[0] * fc[1] - rel[1] * fc[0]);
    } else {
      acc[3] -= double(rel[1] * fc[2] - rel[2] * fc[1]);
      acc[4] -= double(rel[2] * fc[0] - rel[0] * fc[2]);
      acc[5] -= double(rel[0] * fc[1] - rel[1] * fc[0]);
    }
  }
  // FP64 reduction: shuffles within the 64-wide wavefront, then the waves
  // of the block in order through LDS
  __shared__ double red[4][6];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  wave_sum6(acc);
  if (lane == 0)
    for (int k = 0; k < 6; ++k) red[wv][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w2 = 0; w2 < nw; ++w2) s += red[w2][threadIdx.x];
    a.partial[(size_t(p) * a.blocks_per_particle + chunk) * 6 + threadIdx.x] = s;
  }
}

// Particle p's gather partials summed in FP64 by one 64-lane wavefront, in an
// order fixed by the block count: up to 16 blocks lane 0 adds them in block
// order; above that lane l adds blocks l, l+64, ... and an xor butterfly
// combines the lanes (every lane ends with the same bits).  Lane 0 then sets
// the particle's force (and torque where the coupling has one).
template <class T, int D>
__device__ __forceinline__ void finalize_wave(const ParticleArgs& a, uint32_t p, int lane) {
  double s[6] = {0, 0, 0, 0, 0, 0};
  const double* part = a.partial + size_t(p) * a.blocks_per_particle * 6;
  if (a.blocks_per_particle <= 16) {
    if (lane != 0) return;
    for (uint32_t k = 0; k < a.blocks_per_particle; ++k)
      for (int j = 0; j < 6; ++j) s[j] += part[k * 6 + j];
  } else {
    for (uint32_t k = uint32_t(lane); k < a.blocks_per_particle; k += 64)
      for (int j = 0; j < 6; ++j) s[j] += part[k * 6 + j];
    wave_sum6(s);
    if (lane != 0) return;
  }
  T* P = static_cast<T*>(a.parts) + size_t(p) * kPartFields;
  for (int d = 0; d < 3; ++d) {
    a.f64[p * 3 + d] = s[d];
    a.t64[p * 3 + d] = s[3 + d];
    P[PF_FORCE + d] = d < D ? T(s[d]) : T(0);
    P[PF_TORQUE + d] = T(0);
  }
  if (a.psm_torque || a.hlbm) {
    if (D == 2) P[PF_TORQUE] = T(s[3]);
    else for (int d = 0; d < 3; ++d) P[PF_TORQUE + d] = T(s[3 + d]);
  }
}

// one wavefront per particle
template <class T, int D>
__global__ __launch_bounds__(64) void k_finalize(ParticleArgs a) {
  if (blockIdx.x >= a.n) return;  // uniform per block
  finalize_wave<T, D>(a, blockIdx.x, int(threadIdx.x));
}

// build_cell_list (cell_list.hpp:81-108): bins filled in particle-index order
template <class T, int D>
__global__ __launch_bounds__(1024) void k_build_cells(ParticleArgs a, uint64_t bins) {
  __shared__ uint32_t bin_of[4096];
  for (uint64_t k = threadIdx.x; k < bins; k += blockDim.x) a.cl_count[k] = 0;
  if (threadIdx.x == 0) *a.cl_overflow = 0;
  const T bin = T(a.bin);
  for (uint32_t p = threadIdx.x; p < a.n; p += blockDim.x) {
    const T* P = static_cast<const T*>(a.parts) + size_t(p) * kPartFields;
    uint64_t flat = 0, stride = 1;
    for (int d = 0; d < D; ++d) {
      const T xx = P[PF_POS + d] / bin;
      int64_t bb = int64_t(xx);
      if (bb < 0) bb = 0;
      if (uint64_t(bb) >= a.grid[d]) bb = int64_t(a.grid[d]) - 1;
      flat += uint64_t(bb) * stride;
      stride *= a.grid[d];
    }
    bin_of[p] = uint32_t(flat);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < a.n; p += blockDim.x) {
    const uint32_t bp = bin_of[p];
    uint32_t rank = 0;
    for (uint32_t q = 0; q < p; ++q) rank += (bin_of[q] == bp);
    if (rank < a.cap) {
      a.cl_slot[uint64_t(bp) * a.cap + rank] = p;
      atomicMax((unsigned long long*)(a.cl_count + bp), (unsigned long long)(rank + 1));
    } else {
      atomicAdd((unsigned long long*)a.cl_overflow, 1ull);
      if (a.status) __hip_atomic_store(a.status + kStatusOverflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// cell_porosity (couple.hpp:26-60) with for_each_bin_candidate's order
template <class L, class T, class F>
__global__ __launch_bounds__(256) void k_product(ParticleArgs a) {
  constexpr int D = L::D;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  const int x[3] = {int(c % uint32_t(a.nx)), int((c / uint32_t(a.nx)) % uint32_t(a.ny)),
                    int(c / (uint32_t(a.nx) * uint32_t(a.ny)))};
  const T bin = T(a.bin);
  const T eps = T(a.eps);
  int64_t base[3];
  for (int d = 0; d < D; ++d) base[d] = int64_t(T(x[d]) / bin);
  T B = T(1);
  uint32_t total = 1;
  for (int d = 0; d < D; ++d) total *= 3u;
  for (uint32_t off = 0; off < total; ++off) {
    uint32_t rem = off;
    uint64_t flat = 0, stride = 1;
    bool inside = true;
    for (int d = 0; d < D; ++d) {
      const int64_t bb = base[d] + int64_t(rem % 3u) - 1;
      rem /= 3u;
      if (bb < 0 || bb >= int64_t(a.grid[d])) {
        inside = false;
        break;
      }
      flat += uint64_t(bb) * stride;
      stride *= a.grid[d];
    }
    if (!inside) continue;
    const uint64_t cnt = a.cl_count[flat];
    for (uint64_t s = 0; s < cnt; ++s) {
      const uint64_t q = a.cl_slot[flat * a.cap + s];
      const T* P = static_cast<const T*>(a.parts) + size_t(q) * kPartFields;
      T rel[3];
      for (int d = 0; d < D; ++d) rel[d] = T(x[d]) - P[PF_POS + d];
      B = B * disc_porosity<T, D>(rel, P[PF_RADIUS], eps);
    }
  }
  static_cast<F*>(a.por)[c] = F(B);
}

// ---------------------------------------------------------------- DEM
template <class T, int D>
__device__ void integrate_one(const ParticleArgs& a, T* P);

template <class T, int D>
struct Man {
  T depth, n[3], pt[3];
};

template <class T, int D>
__device__ inline bool safe_norm(const T* in, T* out) {
  T l2 = T(0);
  for (int d = 0; d < D; ++d) l2 = l2 + in[d] * in[d];
  if (!(l2 > T(0))) return false;
  const T inv = T(1) / sqrt(l2);
  for (int d = 0; d < D; ++d) out[d] = in[d] * inv;
  return true;
}

template <class T, int D>
__device__ inline void cross_rf(const T* r, const T* f, T* out) {
  if (D == 2) {
    out[0] = r[0] * f[1] - r[1] * f[0];
  } else {
    out[0] = r[1] * f[2] - r[2] * f[1];
    out[1] = r[2] * f[0] - r[0] * f[2];
    out[2] = r[0] * f[1] - r[1] * f[0];
  }
}

// narrow_phase<Spheroid>::contact (contact.hpp:124-184)
template <class T, int D>
__device__ inline bool narrow(const T* lo, const T* hi, Man<T, D>& m) {
  T d[3];
  for (int k = 0; k < D; ++k) d[k] = lo[PF_POS + k] - hi[PF_POS + k];
  T len2 = T(0);
  for (int k = 0; k < D; ++k) len2 = len2 + d[k] * d[k];
  const T r_lo = lo[PF_RADIUS], r_hi = hi[PF_RADIUS];
  const T sum_r = r_lo + r_hi;
  if (!(len2 < sum_r * sum_r)) return false;
  const T len = sqrt(len2);
  T depth = sum_r - len;
  const T cap = (r_lo < r_hi) ? r_lo : r_hi;
  if (depth > cap) depth = cap;
  if (!safe_norm<T, D>(d, m.n)) {
    for (int k = 0; k < D; ++k) m.n[k] = T(0);
    m.n[0] = T(1);
  }
  m.depth = depth;
  for (int k = 0; k < D; ++k) m.pt[k] = T(hi[PF_POS + k] + m.n[k] * (r_hi - depth * 0.5));
  return true;
}

// impl::surface_velocity (contact.hpp:271-311)
template <class T, int D>
__device__ inline void surf_vel(const T* P, const T* r, T steps, T* v) {
  for (int k = 0; k < D; ++k) v[k] = (P[PF_POS + k] - P[PF_POS_OLD + k]) * steps;
  if (D == 2) {
    const T om = (P[PF_ROT] - P[PF_ROT_OLD]) * steps;
    const T t0 = -(om * r[1]), t1 = om * r[0];
    v[0] = v[0] + t0;
    v[1] = v[1] + t1;
  } else {
    T om[3], tg[3];
    for (int k = 0; k < 3; ++k) om[k] = (P[PF_ROT + k] - P[PF_ROT_OLD + k]) * steps;
    cross_rf<T, 3>(om, r, tg);
    for (int k = 0; k < 3; ++k) v[k] = v[k] + tg[k];
  }
}

// impl::soft_shell_response (contact.hpp:204-264); the damping and normal
// force carry double literals in the reference and are double here too
template <class T, int D>
__device__ inline void soft_shell(const Man<T, D>& m, const T* r_self, const T* v_rel, T m_eff,
                                  const ParticleArgs& a, T* F, T* Tq) {
  const T k_n = T(a.k_n), zeta = T(a.zeta_n), gamma_t = T(a.gamma_t), mu = T(a.mu);
  const double g_n = 2.0 * double(zeta) * double(T(sqrt(T(m_eff * k_n))));
  T v_n = T(0);
  for (int k = 0; k < D; ++k) v_n = v_n + v_rel[k] * m.n[k];
  double f_n = double(T(k_n * m.depth)) - g_n * double(v_n);
  if (f_n < 0) f_n = 0;
  T v_t[3], t_hat[3];
  for (int k = 0; k < D; ++k) v_t[k] = v_rel[k] - v_n * m.n[k];
  const bool have_t = safe_norm<T, D>(v_t, t_hat);
  T vt2 = T(0);
  for (int k = 0; k < D; ++k) vt2 = vt2 + v_t[k] * v_t[k];
  const T v_t_len = sqrt(vt2);
  T f_t = gamma_t * v_t_len;
  const double f_t_cap = double(mu) * f_n;
  if (double(f_t) > f_t_cap) f_t = T(f_t_cap);
  for (int k = 0; k < D; ++k) {
    double v = f_n * double(m.n[k]);
    if (have_t) v -= double(T(f_t * t_hat[k]));
    F[k] = T(v);
  }
  cross_rf<T, D>(r_self, F, Tq);
}

// One 64-lane wavefront per particle: the lanes run the broadphase test for
// 64 partners at a time and a ballot collects the candidates; lane 0 then
// takes them in ascending partner order, which is the serial loop's order,
// so the accumulated force and torque are the serial loop's bit for bit.
template <class T, int D>
__global__ __launch_bounds__(64) void k_contact(ParticleArgs a) {
  const uint32_t s = blockIdx.x;
  const int lane = int(threadIdx.x);
  if (s >= a.n) return;  // uniform per block
  // inside a full step the gather's finalize of this particle runs here (the
  // contact of s reads no other particle's force): one launch fewer per step
  if (a.fuse_finalize) {
    finalize_wave<T, D>(a, s, lane);
    __syncthreads();
  }
  // fused with the integrate, this block moves its particle while others
  // still read it: partners then come from the gather's snapshot
  const T* parts = static_cast<const T*>(a.fuse_integrate ? a.snap : a.parts);
  T* me = static_cast<T*>(a.parts) + size_t(s) * kPartFields;
  const T steps = T(a.sub_steps == 0 ? 1 : a.sub_steps);
  T Fs[3] = {me[PF_FORCE], me[PF_FORCE + 1], me[PF_FORCE + 2]};
  T Ts = me[PF_TORQUE];
  const T rl = me[PF_RADIUS];
  // four chunks of 64 partners per round: their broadphase loads are in
  // flight together, then lane 0 takes the candidates in ascending order
  constexpr int kChunks = 4;
  for (uint32_t g0 = 0; g0 < a.n; g0 += 64 * kChunks) {
    uint64_t masks[kChunks];
#pragma unroll
    for (int q = 0; q < kChunks; ++q) {
      const uint32_t jl = g0 + 64u * q + uint32_t(lane);
      bool cand = false;
      if (jl < a.n && jl != s) {
        // broadphase_collision_check (particle.hpp:247-279)
        const T* ot = parts + size_t(jl) * kPartFields;
        T n2 = T(0);
        for (int k = 0; k < D; ++k) {
          const T dd = me[PF_POS + k] - ot[PF_POS + k];
          n2 = n2 + dd * dd;
        }
        const T rr = ot[PF_RADIUS];
        const T rab2 = (rl * rl + rr * rr) + (rr * rl) * T(2.0);
        cand = n2 < rab2;
      }
      masks[q] = __ballot(cand);
    }
    if (lane != 0) continue;
    for (int q = 0; q < kChunks; ++q)
    for (uint64_t mask = masks[q]; mask;) {
      const uint32_t j = g0 + 64u * q + uint32_t(__ffsll((unsigned long long)mask) - 1);
      mask &= mask - 1;
      const T* ot = parts + size_t(j) * kPartFields;
      const bool self_lo = s < j;
      Man<T, D> m;
      if (!narrow<T, D>(self_lo ? me : ot, self_lo ? ot : me, m)) continue;
      // particle_contact_response (contact.hpp:321-364)
      if (!self_lo)
        for (int k = 0; k < D; ++k) m.n[k] = -m.n[k];
      T r_s[3], r_o[3], v_s[3], v_o[3], v_rel[3];
      for (int k = 0; k < D; ++k) {
        r_s[k] = m.pt[k] - me[PF_POS + k];
        r_o[k] = m.pt[k] - ot[PF_POS + k];
      }
      surf_vel<T, D>(me, r_s, steps, v_s);
      surf_vel<T, D>(ot, r_o, steps, v_o);
      for (int k = 0; k < D; ++k) v_rel[k] = v_s[k] - v_o[k];
      const T m_eff = (me[PF_MASS] * ot[PF_MASS]) / (me[PF_MASS] + ot[PF_MASS]);
      T F[3] = {T(0), T(0), T(0)}, Tq[3] = {T(0), T(0), T(0)};
      soft_shell<T, D>(m, r_s, v_rel, m_eff, a, F, Tq);
      for (int k = 0; k < D; ++k) Fs[k] = Fs[k] + F[k];
      if (D == 2) Ts = Ts + Tq[0];
    }
  }
  if (lane != 0) return;
  if (a.walls) {  // accumulate_box_wall_contacts (contact.hpp:407-456)
    const int ext[3] = {a.nx, a.ny, a.nz};
    const T r = me[PF_RADIUS];
    for (int ax = 0; ax < D; ++ax) {
      const T x = me[PF_POS + ax];
      const T hb = T(ext[ax]);
      for (int side = 0; side < 2; ++side) {
        T depth = (side == 0) ? T(r - x) : T(x - (hb - r));
        if (!(depth > T(0))) continue;
        if (depth > r) depth = r;
        Man<T, D> m;
        for (int k = 0; k < D; ++k) {
          m.n[k] = T(0);
          m.pt[k] = me[PF_POS + k];
        }
        m.n[ax] = (side == 0) ? T(1.0) : T(-1.0);
        m.pt[ax] = (side == 0) ? T(0.0) : hb;
        m.depth = depth;
        T r_s[3], v_rel[3];
        for (int k = 0; k < D; ++k) r_s[k] = m.pt[k] - me[PF_POS + k];
        surf_vel<T, D>(me, r_s, steps, v_rel);
        T F[3] = {T(0), T(0), T(0)}, Tq[3] = {T(0), T(0), T(0)};
        soft_shell<T, D>(m, r_s, v_rel, me[PF_MASS], a, F, Tq);
        for (int k = 0; k < D; ++k) Fs[k] = Fs[k] + F[k];
        if (D == 2) Ts = Ts + Tq[0];
      }
    }
  }
  // the gather's box of this step becomes the previous box of the next
  if (a.boxes && a.fuse_finalize)
    for (int k = 0; k < 6; ++k) a.boxes[size_t(s) * 12 + 6 + k] = a.boxes[size_t(s) * 12 + k];
  // only this particle's record is written; other blocks read the snapshot
  // (fused) or only positions, rotations, masses and radii (not fused)
  for (int k = 0; k < D; ++k) me[PF_FORCE + k] = Fs[k];
  if (D == 2) me[PF_TORQUE] = Ts;
  if (a.fuse_integrate) integrate_one<T, D>(a, me);
}

// integrate_particle + verlet_step_lambda (integrate.hpp:45-83, particle.hpp:171-211)
template <class T, int D>
__device__ void integrate_one(const ParticleArgs& a, T* P) {
  // FpLbmOneParticle moves only for sub_steps > 0 and sets only the linear
  // acceleration (fplbm.hpp:233-244)
  if (a.fplbm && a.sub_steps == 0) return;
  const uint64_t steps = a.sub_steps == 0 ? 1 : a.sub_steps;
  const T ts = T(1) / T(steps);
  if (a.fixed_axis) {
    // integrate_particle_fixed_axis (integrate.hpp:143-190): pinned (zero
    // acceleration), angular acceleration along the axis from axis·torque
    for (int k = 0; k < D; ++k) P[PF_ACC + k] = T(0);
    if (D == 2) {
      P[PF_ANG_ACC] = (T(a.axis[0]) * P[PF_TORQUE]) / P[PF_INERTIA];
    } else {
      T tau = T(0);
      for (int k = 0; k < 3; ++k) tau = tau + T(a.axis[k]) * P[PF_TORQUE + k];
      const T alpha = tau / P[PF_INERTIA];
      for (int k = 0; k < 3; ++k) P[PF_ANG_ACC + k] = T(a.axis[k]) * alpha;
    }
  } else {
    for (int k = 0; k < D; ++k) P[PF_ACC + k] = P[PF_FORCE + k] / P[PF_MASS] + T(a.g[k]);
    if (D == 2 && !a.fplbm) P[PF_ANG_ACC] = P[PF_TORQUE] / P[PF_INERTIA];
  }
  constexpr int R = (D == 2) ? 1 : 3;
  for (uint64_t st = 0; st < steps; ++st) {
    const T ts2 = ts * ts;
    for (int k = 0; k < D; ++k) {
      const T pn = ((P[PF_POS + k] * T(2.0)) - P[PF_POS_OLD + k]) + P[PF_ACC + k] * ts2;
      P[PF_POS_OLD + k] = P[PF_POS + k];
      P[PF_POS + k] = pn;
    }
    for (int k = 0; k < R; ++k) {
      const T rn = ((P[PF_ROT + k] * T(2.0)) - P[PF_ROT_OLD + k]) + P[PF_ANG_ACC + k] * ts2;
      P[PF_ROT_OLD + k] = P[PF_ROT + k];
      P[PF_ROT + k] = rn;
    }
  }
}

template <class T, int D>
__global__ void k_integrate(ParticleArgs a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  integrate_one<T, D>(a, static_cast<T*>(a.parts) + size_t(s) * kPartFields);
}

// the gather's boxes become the previous boxes (couplings run on their own,
// outside the full step's fused contact)
__global__ void k_commit_boxes(int32_t* boxes, uint32_t n) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int k = 0; k < 6; ++k) boxes[size_t(p) * 12 + 6 + k] = boxes[size_t(p) * 12 + k];
}

// no previous boxes: every particle's lo = hi = 0
__global__ void k_clear_prev_boxes(int32_t* boxes, uint32_t n) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int k = 0; k < 6; ++k) boxes[size_t(p) * 12 + 6 + k] = 0;
}

// ---------------------------------------------------------------- launchers
template <class L, class S, class T, class F>
static hipError_t couple_t(const ParticleArgs& a, hipStream_t s) {
  const uint32_t N = a.N;
  const bool scatter = a.porosity_mode == LBM_POROSITY_SCATTER;
  // one pass over the cells: PsmReset's porosity, the owners, HlbmReset's N
  // and D, FpLbmReset's force (the PSM sweep keeps the porosity without it)
  const bool sweep = a.boxes != nullptr && scatter && !a.owner && !a.hlbm && !a.fplbm;
  if (!sweep)
    hipLaunchKernelGGL((k_reset_cells<F>), dim3((N + 255) / 256), dim3(256), 0, s, (F*)a.por, a.owner,
                       a.hlbm ? (F*)a.pn : nullptr, a.hlbm ? (F*)a.pd : nullptr,
                       a.fplbm ? (F*)a.force : nullptr, N, L::D, scatter ? 1 : 0);
  const dim3 pg(a.n * a.blocks_per_particle);
  if (a.owner) hipLaunchKernelGGL((k_owner<L, T>), pg, dim3(256), 0, s, a);
  const int cpl = a.hlbm ? kCplHlbm : (a.fplbm ? kCplFplbm : kCplPsm);
  if (a.owner) {
    if (cpl == kCplHlbm) hipLaunchKernelGGL((k_gather<L, S, T, F, kCplHlbm, false>), pg, dim3(256), 0, s, a);
    else if (cpl == kCplFplbm) hipLaunchKernelGGL((k_gather<L, S, T, F, kCplFplbm, false>), pg, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_gather<L, S, T, F, kCplPsm, false>), pg, dim3(256), 0, s, a);
  } else {
    if (cpl == kCplHlbm) hipLaunchKernelGGL((k_gather<L, S, T, F, kCplHlbm, true>), pg, dim3(256), 0, s, a);
    else if (cpl == kCplFplbm) hipLaunchKernelGGL((k_gather<L, S, T, F, kCplFplbm, true>), pg, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_gather<L, S, T, F, kCplPsm, true>), pg, dim3(256), 0, s, a);
  }
  if (!a.fuse_finalize) {
    hipLaunchKernelGGL((k_finalize<T, L::D>), dim3(a.n), dim3(64), 0, s, a);
    if (sweep) hipLaunchKernelGGL(k_commit_boxes, dim3((a.n + 255) / 256), dim3(256), 0, s, a.boxes, a.n);
  }
  if (!scatter) {
    uint64_t bins = a.grid[0] * a.grid[1] * a.grid[2];
    hipLaunchKernelGGL((k_build_cells<T, L::D>), dim3(1), dim3(1024), 0, s, a, bins);
    hipLaunchKernelGGL((k_product<L, T, F>), dim3((N + 255) / 256), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

template <class L>
static hipError_t couple_l(int storage, const ParticleArgs& a, hipStream_t s) {
  if (a.state64) {
    switch (storage) {
      case LBM_F64: return couple_t<L, double, double, double>(a, s);
      case LBM_F32: return couple_t<L, float, double, float>(a, s);
      case LBM_F16S: return couple_t<L, __half, double, float>(a, s);
      case LBM_BF16S: return couple_t<L, __hip_bfloat16, double, float>(a, s);
    }
    return hipErrorInvalidValue;
  }
  switch (storage) {
    case LBM_F64: return couple_t<L, double, double, double>(a, s);
    case LBM_F32: return couple_t<L, float, float, float>(a, s);
    case LBM_F16S: return couple_t<L, __half, float, float>(a, s);
    case LBM_BF16S: return couple_t<L, __hip_bfloat16, float, float>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_reset(bool dbl, void* por, uint32_t N, int32_t* boxes, uint32_t n, hipStream_t s) {
  if (dbl)
    hipLaunchKernelGGL((k_reset_cells<double>), dim3((N + 255) / 256), dim3(256), 0, s, (double*)por, nullptr,
                       nullptr, nullptr, nullptr, N, 0, 1);
  else
    hipLaunchKernelGGL((k_reset_cells<float>), dim3((N + 255) / 256), dim3(256), 0, s, (float*)por, nullptr,
                       nullptr, nullptr, nullptr, N, 0, 1);
  if (n > 0) hipLaunchKernelGGL(k_clear_prev_boxes, dim3((n + 255) / 256), dim3(256), 0, s, boxes, n);
  return hipGetLastError();
}

hipError_t launch_psm_couple(int lattice, int storage, const ParticleArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  switch (lattice) {
    case LBM_D2Q9: return couple_l<D2Q9>(storage, a, s);
    case LBM_D3Q19: return couple_l<D3Q19>(storage, a, s);
    case LBM_D3Q27: return couple_l<D3Q27>(storage, a, s);
  }
  return hipErrorInvalidValue;
}

template <class T, class F>
static hipError_t stamp_t(int dims, const ParticleArgs& a, hipStream_t s) {
  const dim3 pg(a.n * a.blocks_per_particle);
  if (dims == 2) hipLaunchKernelGGL((k_hlbm_stamp<T, F, 2>), pg, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_hlbm_stamp<T, F, 3>), pg, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_hlbm_stamp(int dims, int storage, const ParticleArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (storage == LBM_F64) return stamp_t<double, double>(dims, a, s);
  return a.state64 ? stamp_t<double, float>(dims, a, s) : stamp_t<float, float>(dims, a, s);
}

hipError_t launch_dem_contact(int dims, int storage, const ParticleArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 g(a.n), b(64);
  const bool dbl = storage == LBM_F64 || a.state64;
  if (dims == 2) {
    if (dbl) hipLaunchKernelGGL((k_contact<double, 2>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_contact<float, 2>), g, b, 0, s, a);
  } else {
    if (dbl) hipLaunchKernelGGL((k_contact<double, 3>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_contact<float, 3>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_particle_integrate(int dims, int storage, const ParticleArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 g((a.n + 255) / 256), b(256);
  const bool dbl = storage == LBM_F64 || a.state64;
  if (dims == 2) {
    if (dbl) hipLaunchKernelGGL((k_integrate<double, 2>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_integrate<float, 2>), g, b, 0, s, a);
  } else {
    if (dbl) hipLaunchKernelGGL((k_integrate<double, 3>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_integrate<float, 3>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace lbmx
