// Learned warm start: a convolutional encoder-decoder trained and run with
// hand-written bf16 MFMA kernels (v_mfma_f32_16x16x32_bf16), FP32 accumulate.
//
// Layout: activations NHWC bf16 (pixel-major, channels contiguous), so that
// every 3x3 tap of 8 consecutive channels is one 16-byte load.  A 3x3 same
// convolution is the GEMM  out[p][co] = sum_{tap,ci} in[p+off(tap)][ci] *
// W[tap][ci][co]  with M = pixels, N = Cout, K = 9*Cin (implicit im2col: the
// loader gathers the A tile straight from the activation).  The data
// gradient is the same GEMM on the output gradient with the weights flipped
// and transposed (Wt[8-tap][co][ci] = W[tap][ci][co]); the weight gradient is
// the GEMM dW[(tap,ci)][co] = sum_p in[p+off(tap)][ci] * dOut[p][co], split
// over pixel ranges with a fixed-order reduction of the partial tiles, so
// training is deterministic run to run.
//
// Block tile 128 (M) x 16*NC (N) x 32 (K), NC per layer so that the
// layer's channel count is a whole number of tiles (up to 128 columns):
// four 64-wide waves, each owning 32 rows x 16*NC columns = 2 x NC MFMA
// 16x16 blocks.  A and B tiles are staged through two LDS buffers with the K
// index contiguous, so that every MFMA operand is one 16-byte LDS read (rows
// padded by 8 elements against bank conflicts); one barrier per K step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lbm_mirror_amd/lbm.h"

namespace lbmnet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> bf16 bits (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

struct Act {  // element (pixel q, channel c) at p[q*ld + off + c]
  uint16_t* p;
  int ld, off;
};

constexpr int BM = 128, BK = 32, LK = BK + 8;
// Conv K steps whose global loads are in flight across whole K steps (a
// register ring of that many slots), per tile width.  0: the next step's
// loads are issued at the start of a step and stored to LDS at its end.  2
// for the 64-, 80- and 128-channel tiles (their conv 4-7% faster); the 32-
// and 96-channel tiles are as fast or faster with 0 (the ring's registers
// cost the 96-channel tile a wave per SIMD), and 4 or more slow every tile.
constexpr int conv_prefetch(int nc) { return (nc == 4 || nc == 5 || nc == 8) ? 2 : 0; }

struct ConvArgs {
  int B, H, W;
  int cin, cout;  // GEMM K = 9*cin, N = cout; both multiples of 8
  Act in, out;
  const uint16_t* w;  // [9][cin][cout]
  const float* bias;
  int bias_n;
  float* outf;  // FP32 result (ld outf_n, columns < outf_n) instead of out
  int outf_n;
  int relu;
  Act mask;  // mask.p != null: result *= (mask > 0)
};

__device__ inline uint4 load_tap(const Act& s, int H, int W, int b, int y, int x, int tap, int c) {
  const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
  if (yy < 0 || yy >= H || xx < 0 || xx >= W) return make_uint4(0u, 0u, 0u, 0u);
  const size_t q = (size_t(b) * H + yy) * W + xx;
  return *reinterpret_cast<const uint4*>(s.p + q * s.ld + s.off + c);
}

// A wavefront's 32 rows of the block tile against NC 16-column blocks of the
// B tile (2 x NC MFMA 16x16x32 per K step of 32).
template <int NC>
__device__ inline void mma_tile_n(const uint16_t* As, const uint16_t* Bs, int wave, int lane,
                                  f32x4 (&acc)[2][NC]) {
  bf16x8 af[2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
    af[r] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(
                                           &As[(wave * 32 + r * 16 + (lane & 15)) * LK + (lane >> 4) * 8]));
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const bf16x8 bfr = __builtin_bit_cast(
        bf16x8, *reinterpret_cast<const uint4*>(&Bs[(c * 16 + (lane & 15)) * LK + (lane >> 4) * 8]));
#pragma unroll
    for (int r = 0; r < 2; ++r) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[r], bfr, acc[r][c], 0, 0, 0);
  }
}

// Forward convolution and data gradient (implicit GEMM over 3x3 taps).
// Block tile BM = 128 pixels x 16*NC output channels (NC chosen per layer so
// the channel count is a whole number of tiles), K in steps of 32 through
// two LDS buffers: the next step's A (gathered taps) and B (weights, stored
// N-major with K contiguous, so both tiles go to LDS without transposing)
// are loaded into registers while the current step's MFMAs run, then written
// to the other buffer; one barrier per K step.
template <int NC>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  constexpr int BNc = 16 * NC;
  constexpr int NB = (BNc * 4 + 255) / 256;  // B chunks of 8 per thread
  constexpr int KP = conv_prefetch(NC);
  constexpr int KS = KP > 0 ? KP : 1;  // ring slots
  // the two tile buffers; after the K loop the same LDS holds the block's
  // bf16 results for 16-byte stores (128 rows of BNc + 8 values)
  constexpr int kTiles = 2 * BM * LK + 2 * BNc * LK, kOut = BM * (BNc + 8);
  __shared__ __align__(16) uint16_t smem[kTiles > kOut ? kTiles : kOut];
  uint16_t (*As)[BM * LK] = reinterpret_cast<uint16_t (*)[BM * LK]>(smem);
  uint16_t (*Bs)[BNc * LK] = reinterpret_cast<uint16_t (*)[BNc * LK]>(smem + 2 * BM * LK);
  const int M = a.B * a.H * a.W, K = 9 * a.cin, N = a.cout, HW = a.H * a.W;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BNc;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int kc = (tid & 3) * 8;
  int rb[2], ry[2], rx[2];
  bool rv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int m = m0 + ((tid + 256 * j) >> 2);
    rv[j] = m < M;
    const int mm = rv[j] ? m : 0;
    rb[j] = mm / HW;
    const int r = mm - rb[j] * HW;
    ry[j] = r / a.W;
    rx[j] = r - ry[j] * a.W;
  }
  // slot q of the ring holds the A and B chunks of one K step from its load
  // to its LDS store
  uint4 ra[KS][2], rw[KS][NB];
  auto gload = [&](int q, int k0) {
    const int kk = k0 + kc;
    int tap = 0, c = 0;
    if (kk < K) {
      tap = kk / a.cin;
      c = kk - tap * a.cin;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
      ra[q][j] = (rv[j] && kk < K) ? load_tap(a.in, a.H, a.W, rb[j], ry[j], rx[j], tap, c) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int qq = 0; qq < NB; ++qq) {
      const int i = tid + 256 * qq;
      const int n = n0 + (i >> 2), k = k0 + (i & 3) * 8;
      rw[q][qq] = (i < BNc * 4 && n < N && k < K) ? *reinterpret_cast<const uint4*>(a.w + size_t(n) * K + k)
                                                   : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto sstore = [&](int buf, int q) {
#pragma unroll
    for (int j = 0; j < 2; ++j) *reinterpret_cast<uint4*>(&As[buf][((tid + 256 * j) >> 2) * LK + kc]) = ra[q][j];
#pragma unroll
    for (int qq = 0; qq < NB; ++qq) {
      const int i = tid + 256 * qq;
      if (i < BNc * 4) *reinterpret_cast<uint4*>(&Bs[buf][(i >> 2) * LK + (i & 3) * 8]) = rw[q][qq];
    }
  };
  f32x4 acc[2][NC];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nk = (K + BK - 1) / BK;
  gload(0, 0);
  sstore(0, 0);
#pragma unroll
  for (int j = 1; j <= KP; ++j)
    if (j < nk) gload(j % KS, j * BK);
  __syncthreads();
  if constexpr (KP == 0) {
    for (int kt = 0; kt < nk; ++kt) {
      if (kt + 1 < nk) gload(0, (kt + 1) * BK);
      mma_tile_n<NC>(As[kt & 1], Bs[kt & 1], wave, lane, acc);
      if (kt + 1 < nk) sstore((kt + 1) & 1, 0);  // read by no wave since the last barrier
      __syncthreads();
    }
  } else {
    // step kt: MFMAs on LDS buffer kt & 1; step kt + 1 from the ring into
    // the other buffer (read by no wave since the last barrier); its slot
    // then takes the loads of step kt + 1 + KP (slots are compile-time
    // indices)
    for (int kt0 = 0; kt0 < nk; kt0 += KS) {
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const int kt = kt0 + j;
        if (kt < nk) {
          mma_tile_n<NC>(As[kt & 1], Bs[kt & 1], wave, lane, acc);
          if (kt + 1 < nk) sstore((kt + 1) & 1, (j + 1) % KS);
          if (kt + 1 + KP < nk) gload((j + 1) % KS, (kt + 1 + KP) * BK);
          __syncthreads();
        }
      }
    }
  }
  if (a.outf) {  // FP32 result (the output layer)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int m = m0 + wave * 32 + r * 16 + (lane >> 4) * 4 + j;
          const int n = n0 + c * 16 + (lane & 15);
          if (m >= M || n >= a.outf_n) continue;
          float v = acc[r][c][j];
          if (a.bias && n < a.bias_n) v = v + a.bias[n];
          if (a.relu) v = v > 0.f ? v : 0.f;
          a.outf[size_t(m) * a.outf_n + n] = v;
        }
    return;
  }
  // bias and ReLU into the LDS tile (the K loop ended with a barrier) ...
  constexpr int OP = BNc + 8;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = wave * 32 + r * 16 + (lane >> 4) * 4 + j, col = c * 16 + (lane & 15);
        const int n = n0 + col;
        float v = acc[r][c][j];
        if (a.bias && n < a.bias_n) v = v + a.bias[n];
        if (a.relu) v = v > 0.f ? v : 0.f;
        smem[row * OP + col] = f2bf(v);
      }
  __syncthreads();
  // ... then 8 channels per 16-byte store, with the input-gradient mask
  // (zero where the layer input was not positive) applied per element
  constexpr int CH = BNc / 8;
  for (int t = tid; t < BM * CH; t += 256) {
    const int row = t / CH, ch = t - row * CH;
    const int m = m0 + row, n = n0 + ch * 8;
    if (m >= M || n >= N) continue;
    uint4 v = *reinterpret_cast<const uint4*>(&smem[row * OP + ch * 8]);
    if (a.mask.p) {
      const uint4 mk = *reinterpret_cast<const uint4*>(a.mask.p + size_t(m) * a.mask.ld + a.mask.off + n);
      uint32_t w[4] = {v.x, v.y, v.z, v.w};
      const uint32_t mw[4] = {mk.x, mk.y, mk.z, mk.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!(bf2f(uint16_t(mw[q] & 0xffffu)) > 0.f)) w[q] &= 0xffff0000u;
        if (!(bf2f(uint16_t(mw[q] >> 16)) > 0.f)) w[q] &= 0x0000ffffu;
      }
      v = make_uint4(w[0], w[1], w[2], w[3]);
    }
    *reinterpret_cast<uint4*>(a.out.p + size_t(m) * a.out.ld + a.out.off + n) = v;
  }
}

// 16-column blocks per conv tile for N output channels: the widest tile of
// at most 128 channels that divides N (N = 8, the output layer, takes 32)
int conv_nc(int N) {
  for (int nc : {8, 6, 5, 4, 2})
    if (N % (16 * nc) == 0) return nc;
  return 2;
}

hipError_t launch_conv(const ConvArgs& a, hipStream_t s) {
  const int M = a.B * a.H * a.W, nc = conv_nc(a.cout);
  const dim3 grid((M + BM - 1) / BM, (a.cout + 16 * nc - 1) / (16 * nc));
  switch (nc) {
    case 8: hipLaunchKernelGGL(k_conv<8>, grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_conv<6>, grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_conv<5>, grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, s, a); break;
  }
  return hipGetLastError();
}

struct WgradArgs {
  int B, H, W, cin, cout;  // M' = 9*cin (+ the bias row), N' = cout, K' = pixels
  Act in, dpre;
  float* partial;  // [splits][9*cin + 8][cout]
  int chunk;       // pixels per split, multiple of BK
};

// Two 8-element chunks (pixels k and k+1 of the same 8 rows) into LDS rows
// base[e*stride], as one 32-bit write per row.
__device__ inline void scatter8x2(uint16_t* base, int stride, uint4 v0, uint4 v1) {
  const uint32_t w0[4] = {v0.x, v0.y, v0.z, v0.w}, w1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    *reinterpret_cast<uint32_t*>(base + (2 * t) * stride) = (w1[t] << 16) | (w0[t] & 0xffffu);
    *reinterpret_cast<uint32_t*>(base + (2 * t + 1) * stride) = (w1[t] & 0xffff0000u) | (w0[t] >> 16);
  }
}

// Weight gradient: per split of the pixel range, one partial 128 x 16*NC tile
// of dW[(tap,ci)][co] = sum_p in[p+off(tap)][ci] * dpre[p][co].  Row 9*cin of
// the im2col operand is 1 on every pixel, so the same GEMM yields the bias
// gradient sum_p dpre[p][co] in that row (rows 9*cin+1 .. +7 are zero).  The
// K dimension is pixels: both operands arrive channel-contiguous and are
// transposed on their way into LDS, two pixels per 32-bit write; two LDS
// buffers, the next step's tiles loaded while the current one's MFMAs run.
template <int NC>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
  constexpr int BNc = 16 * NC;
  constexpr int BCH = BNc / 8;  // 8-channel chunks of a B row block per pixel
  __shared__ __align__(16) uint16_t As[2][BM * LK];
  __shared__ __align__(16) uint16_t Bs[2][BNc * LK];
  const int M = a.B * a.H * a.W, Mw = 9 * a.cin, Mx = Mw + 8, N = a.cout, HW = a.H * a.W;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BNc, s = blockIdx.z;
  const int p0 = s * a.chunk, p1 = min(p0 + a.chunk, M);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // The 16 lanes of a quarter-wavefront take 16 consecutive pixel pairs of
  // one 8-row group, so their transposing 32-bit LDS writes fall in 16
  // consecutive banks (row groups a multiple of 32 banks apart would
  // otherwise collide 8 ways)
  const int g = tid >> 4;
  const int m8 = m0 + g * 8;
  const bool mv = m8 < Mw, ones = m8 == Mw;
  const int tap = mv ? m8 / a.cin : 0;
  const int ci = mv ? m8 - tap * a.cin : 0;
  const int kq = 2 * (tid & 15);  // A: pixel pair (kq, kq+1) of the K step
  const bool bt = tid < 16 * BCH;  // B: 16 pixel pairs x BCH chunks
  const int kb = 2 * (tid & 15), nc = (tid >> 4) * 8;
  uint4 ra[2], rb[2];
  // (b, y, x) of this thread's first pixel kq of the current K step, advanced
  // by BK pixels per step without divisions
  int pb_ = p0 + kq;
  int cb = pb_ / HW, cy = (pb_ - cb * HW) / a.W, cx = pb_ - cb * HW - cy * a.W;
  auto gload = [&](int pb) {
    int b = cb, y = cy, x = cx;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int q = pb + kq + j;
      ra[j] = make_uint4(0u, 0u, 0u, 0u);
      if (q < p1) {
        if (mv) ra[j] = load_tap(a.in, a.H, a.W, b, y, x, tap, ci);
        else if (ones) ra[j].x = 0x3f80u;  // bf16 1.0 in row 9*cin
      }
      if (++x == a.W) {  // the pair's second pixel
        x = 0;
        if (++y == a.H) {
          y = 0;
          ++b;
        }
      }
    }
    cx += BK;
    while (cx >= a.W) {
      cx -= a.W;
      if (++cy == a.H) {
        cy = 0;
        ++cb;
      }
    }
    if (bt) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int q = pb + kb + j, n = n0 + nc;
        rb[j] = (q < p1 && n < N) ? *reinterpret_cast<const uint4*>(a.dpre.p + size_t(q) * a.dpre.ld + a.dpre.off + n)
                                  : make_uint4(0u, 0u, 0u, 0u);
      }
    }
  };
  auto sstore = [&](int buf) {
    scatter8x2(&As[buf][(g * 8) * LK + kq], LK, ra[0], ra[1]);
    if (bt) scatter8x2(&Bs[buf][nc * LK + kb], LK, rb[0], rb[1]);
  };
  f32x4 acc[2][NC];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nk = p1 > p0 ? (p1 - p0 + BK - 1) / BK : 0;
  if (nk > 0) {
    gload(p0);
    sstore(0);
    __syncthreads();
  }
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(p0 + (kt + 1) * BK);
    mma_tile_n<NC>(As[cur], Bs[cur], wave, lane, acc);
    if (kt + 1 < nk) sstore(cur ^ 1);  // read by no wave since the last barrier
    __syncthreads();
  }
  float* out = a.partial + size_t(s) * Mx * N;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = m0 + wave * 32 + r * 16 + (lane >> 4) * 4 + j;
        const int n = n0 + c * 16 + (lane & 15);
        if (m <= Mw && n < N) out[size_t(m) * N + n] = acc[r][c][j];
      }
}

hipError_t launch_wgrad(const WgradArgs& a, int splits, hipStream_t s) {
  const int nc = conv_nc(a.cout);
  const dim3 grid((9 * a.cin + 8 + BM - 1) / BM, (a.cout + 16 * nc - 1) / (16 * nc), splits);
  switch (nc) {
    case 8: hipLaunchKernelGGL(k_wgrad<8>, grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_wgrad<6>, grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_wgrad<5>, grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_wgrad<4>, grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(256), 0, s, a); break;
  }
  return hipGetLastError();
}

// grad[m*cr + n] = sum_s partial[s][m][n] for m <= 9*cin, n < cr (the
// unpadded Cout); row 9*cin is the bias, which follows the weights in the
// flat parameter layout.  A block takes 64 consecutive outputs: four groups
// of 64 threads add the splits k = g, g+4, ..., then the four sums are added
// in group order (a fixed order: training repeats bit for bit).
__global__ __launch_bounds__(256) void k_wreduce(const float* partial, int splits, int Mw, int cpad, int cr,
                                                 float* grad) {
  __shared__ float part[4][64];
  const int gq = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = blockIdx.x * 64 + l;
  const bool ok = i < (Mw + 1) * cr;
  const int m = ok ? i / cr : 0, n = ok ? i - m * cr : 0;
  // four running sums per thread (splits gq + 4(4i + u), u = 0..3) keep
  // four loads in flight; combined in a fixed order
  float su[4] = {0.f, 0.f, 0.f, 0.f};
  if (ok) {
    const size_t stride = size_t(Mw + 8) * cpad;
    const float* src = partial + size_t(m) * cpad + n;
    int k = gq;
    for (; k + 12 < splits; k += 16) {
#pragma unroll
      for (int u = 0; u < 4; ++u) su[u] += src[size_t(k + 4 * u) * stride];
    }
    for (int u = 0; k < splits; k += 4, ++u) su[u] += src[size_t(k) * stride];
  }
  part[gq][l] = (su[0] + su[1]) + (su[2] + su[3]);
  __syncthreads();
  if (gq == 0 && ok) grad[i] = ((part[0][l] + part[1][l]) + part[2][l]) + part[3][l];
}

// 8 channels (16 bytes, 16-byte aligned: channel offsets and row pitches are
// multiples of 8) as one vector access
__device__ inline void ld8(const uint16_t* p, float (&f)[8]) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f[2 * t] = bf2f(uint16_t(w[t] & 0xffffu));
    f[2 * t + 1] = bf2f(uint16_t(w[t] >> 16));
  }
}
__device__ inline void st8(uint16_t* p, const float (&f)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) w[t] = uint32_t(f2bf(f[2 * t])) | (uint32_t(f2bf(f[2 * t + 1])) << 16);
  *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

// 2x2 average pool of C channels: fine (H x W) -> coarse (H/2 x W/2).
__global__ void k_pool_fwd(Act src, Act dst, int B, int H, int W, int C) {
  const int C8 = C / 8, h = H / 2, w = W / 2;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(B) * h * w * C8) return;
  const int c = int(i % C8) * 8;
  const size_t q = i / C8;
  const int x = int(q % w), y = int((q / w) % h), b = int(q / (size_t(w) * h));
  const size_t f00 = (size_t(b) * H + 2 * y) * W + 2 * x;
  const uint16_t* p00 = src.p + f00 * src.ld + src.off + c;
  const uint16_t* p01 = p00 + src.ld;
  const uint16_t* p10 = p00 + size_t(W) * src.ld;
  const uint16_t* p11 = p10 + src.ld;
  float a00[8], a01[8], a10[8], a11[8], o[8];
  ld8(p00, a00);
  ld8(p01, a01);
  ld8(p10, a10);
  ld8(p11, a11);
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = (((a00[e] + a01[e]) + a10[e]) + a11[e]) * 0.25f;
  st8(dst.p + q * dst.ld + dst.off + c, o);
}

// Pool backward into the skip gradient: g[p] = g[p] + (act[p] > 0 ? dpool[parent]/4 : 0).
__global__ void k_pool_bwd(Act dpool, Act gskip, Act act, int B, int H, int W, int C) {
  const int C8 = C / 8;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(B) * H * W * C8) return;
  const int c = int(i % C8) * 8;
  const size_t p = i / C8;
  const int x = int(p % W), y = int((p / W) % H), b = int(p / (size_t(W) * H));
  const size_t par = (size_t(b) * (H / 2) + y / 2) * (W / 2) + x / 2;
  uint16_t* gp = gskip.p + p * gskip.ld + gskip.off + c;
  float g[8], v[8], d[8];
  ld8(gp, g);
  ld8(act.p + p * act.ld + act.off + c, v);
  ld8(dpool.p + par * dpool.ld + dpool.off + c, d);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float add = v[e] > 0.f ? d[e] * 0.25f : 0.f;
    g[e] = g[e] + add;
  }
  st8(gp, g);
}

// Nearest 2x upsample: dst fine pixel p = src coarse parent(p).
__global__ void k_up_fwd(Act src, Act dst, int B, int H, int W, int C) {
  const int C8 = C / 8;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(B) * H * W * C8) return;
  const int c = int(i % C8) * 8;
  const size_t p = i / C8;
  const int x = int(p % W), y = int((p / W) % H), b = int(p / (size_t(W) * H));
  const size_t par = (size_t(b) * (H / 2) + y / 2) * (W / 2) + x / 2;
  *reinterpret_cast<uint4*>(dst.p + p * dst.ld + dst.off + c) =
      *reinterpret_cast<const uint4*>(src.p + par * src.ld + src.off + c);
}

// Upsample backward: coarse gradient = ((g00 + g01) + g10) + g11 of its four children.
__global__ void k_up_bwd(Act gfine, Act gcoarse, int B, int H, int W, int C) {
  const int C8 = C / 8, h = H / 2, w = W / 2;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(B) * h * w * C8) return;
  const int c = int(i % C8) * 8;
  const size_t q = i / C8;
  const int x = int(q % w), y = int((q / w) % h), b = int(q / (size_t(w) * h));
  const size_t f00 = (size_t(b) * H + 2 * y) * W + 2 * x;
  const uint16_t* p00 = gfine.p + f00 * gfine.ld + gfine.off + c;
  const uint16_t* p01 = p00 + gfine.ld;
  const uint16_t* p10 = p00 + size_t(W) * gfine.ld;
  const uint16_t* p11 = p10 + gfine.ld;
  float a00[8], a01[8], a10[8], a11[8], o[8];
  ld8(p00, a00);
  ld8(p01, a01);
  ld8(p10, a10);
  ld8(p11, a11);
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = ((a00[e] + a01[e]) + a10[e]) + a11[e];
  st8(gcoarse.p + q * gcoarse.ld + gcoarse.off + c, o);
}

// Masked MSE over 3 output channels: gy[q][c] = bf16(2 (y - t) / (3 nf)) on
// counted pixels (channels 3..7 zero), per-block partial sums of (y - t)^2.
__global__ void k_loss(const float* y, const float* t, const uint8_t* m, int M, float inv3nf, uint16_t* gy,
                       double* partial) {
  __shared__ double red[256];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (q < M) {
    uint16_t g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (m[q]) {
      for (int c = 0; c < 3; ++c) {
        const float d = y[size_t(q) * 3 + c] - t[size_t(q) * 3 + c];
        s += double(d) * double(d);
        g[c] = f2bf(2.f * d * inv3nf);
      }
    }
    uint4 v;
    v.x = g[0] | (uint32_t(g[1]) << 16);
    v.y = g[2] | (uint32_t(g[3]) << 16);
    v.z = g[4] | (uint32_t(g[5]) << 16);
    v.w = g[6] | (uint32_t(g[7]) << 16);
    *reinterpret_cast<uint4*>(gy + size_t(q) * 8) = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_loss_reduce(const double* partial, int n, double scale, double* loss_out, float* loss_f) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (loss_out) *loss_out = red[0] * scale;
    if (loss_f) *loss_f = float(red[0] * scale);
  }
}

__global__ void k_adam(float* w, const float* g, float* m1, float* m2, size_t n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// FP32 master [9][cin][cr] -> the bf16 B operands of the conv GEMMs, N-major
// with K contiguous: forward W[co][tap*cin + ci], data gradient
// Wt[ci][(8 - tap)*cpad + co] (taps flipped).
__global__ void k_wconvert(const float* wm, int cin, int cr, int cpad, uint16_t* wb, uint16_t* wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * cin * cpad) return;
  const int co = i % cpad, ci = (i / cpad) % cin, tap = i / (cpad * cin);
  const uint16_t v = co < cr ? f2bf(wm[(size_t(tap) * cin + ci) * cr + co]) : uint16_t(0);
  wb[size_t(co) * (9 * cin) + tap * cin + ci] = v;
  wt[size_t(ci) * (9 * cpad) + (8 - tap) * cpad + co] = v;
}

// In-place standardisation of stored targets: t = (t - mu) / sd.
__global__ void k_standardize(float* t, size_t n, float m0, float m1, float m2, float s0, float s1, float s2) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = int(i % 3);
  const float m = c == 0 ? m0 : (c == 1 ? m1 : m2);
  const float sd = c == 0 ? s0 : (c == 1 ? s1 : s2);
  t[i] = (t[i] - m) / sd;
}

// Minibatch assembly from the on-device training set.
__global__ void k_gather(const uint16_t* xs, const float* ts, const uint8_t* ms, const uint32_t* idx, int B, int HW,
                         uint16_t* x, float* t, uint8_t* m) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(B) * HW) return;
  const int b = int(i / HW), p = int(i % HW);
  const size_t src = size_t(idx[b]) * HW + p;
  *reinterpret_cast<uint4*>(x + i * 8) = *reinterpret_cast<const uint4*>(xs + src * 8);
  t[i * 3 + 0] = ts[src * 3 + 0];
  t[i * 3 + 1] = ts[src * 3 + 1];
  t[i * 3 + 2] = ts[src * 3 + 2];
  m[i] = ms[src];
}

}  // namespace lbmnet

using namespace lbmnet;

// defined with the rest of the C ABI (lbm_last_error reads it)
lbm_status lbmx_report(lbm_status s, const char* msg);

namespace {

lbm_status nfail(lbm_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return lbmx_report(s, buf);
}

#define NET_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return nfail(e_ == hipErrorOutOfMemory ? LBM_ERR_OUT_OF_MEMORY : LBM_ERR_CRITICAL, \
                   "%s: %s", #expr, hipGetErrorString(e_));                            \
  } while (0)

// Density target: (rho - 1) / (3 u_max^2), the pressure p = (rho - 1)/3 in
// units of the inlet dynamic pressure u_max^2: O(1) from creeping to
// inertial flow.
double rho_unit(double u_max) { return 3.0 * u_max * u_max; }

struct Layer {
  int cin, cr, cpad;  // input channels, output channels (real / padded to 8)
  size_t w_off, b_off;
  uint16_t *wb = nullptr, *wt = nullptr;
  int relu;
  int level;
};

}  // namespace

struct lbm_net {
  lbm_net_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  std::vector<Layer> layers;
  size_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint64_t adam_t = 0;
  // activations (batch capacity B), level l has B*(H>>l)*(W>>l) pixels
  uint16_t* x = nullptr;
  uint16_t *a[4] = {}, *cat[3] = {}, *pool[4] = {}, *dec[3] = {}, *bott = nullptr;
  uint16_t *ga[4] = {}, *gcat[3] = {}, *gpool[4] = {}, *gdec[3] = {}, *gbott = nullptr, *gy = nullptr;
  float *y = nullptr, *t = nullptr;
  uint8_t* msk = nullptr;
  float* partial = nullptr;
  size_t partial_n = 0;
  double* lpartial = nullptr;
  double* dloss = nullptr;
  float* dlosses = nullptr;
  uint32_t dlosses_n = 0;
  uint32_t* didx = nullptr;
  size_t didx_n = 0;
  // training set
  uint16_t* xs = nullptr;
  float* ts = nullptr;
  uint8_t* ms = nullptr;
  std::vector<double> nf;  // counted pixels per sample
  // output standardisation: the network predicts (target - mu) / sd per
  // channel; mu, sd come from the training set at the first training call
  double tsum[3] = {0, 0, 0}, tsq[3] = {0, 0, 0}, tcnt = 0;
  bool frozen = false;
  float mu[3] = {0.f, 0.f, 0.f}, sd[3] = {1.f, 1.f, 1.f};
};

namespace {

int W_(const lbm_net* n, int l) { return n->cfg.width >> l; }
int H_(const lbm_net* n, int l) { return n->cfg.height >> l; }

size_t level_pixels(const lbm_net* n, int l) { return size_t(n->cfg.batch) * H_(n, l) * W_(n, l); }

template <class T>
lbm_status dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return LBM_OK;
  NET_TRY(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  NET_TRY(hipMemset(*p, 0, count * sizeof(T)));
  return LBM_OK;
}

void build_layers(lbm_net* n) {
  const int* w = n->cfg.widths;
  struct Spec {
    int cin, cout, relu, level;
  };
  const Spec spec[12] = {{8, w[0], 1, 0},        {w[0], w[0], 1, 0},        {w[0], w[1], 1, 1},
                         {w[1], w[1], 1, 1},     {w[1], w[2], 1, 2},        {w[2], w[2], 1, 2},
                         {w[2], w[3], 1, 3},     {w[3], w[3], 1, 3},        {w[3] + w[2], w[2], 1, 2},
                         {w[2] + w[1], w[1], 1, 1}, {w[1] + w[0], w[0], 1, 0}, {w[0], 3, 0, 0}};
  size_t off = 0;
  n->layers.clear();
  for (const Spec& s : spec) {
    Layer L;
    L.cin = s.cin;
    L.cr = s.cout;
    L.cpad = (s.cout + 7) / 8 * 8;
    L.relu = s.relu;
    L.level = s.level;
    L.w_off = off;
    off += size_t(9) * s.cin * s.cout;
    L.b_off = off;
    off += s.cout;
    n->layers.push_back(L);
  }
  n->nparam = off;
}

Act act(uint16_t* p, int ld, int off = 0) { return Act{p, ld, off}; }

// Splits of the pixel range for a weight-gradient launch of layer L: about
// kWgradBlocks blocks (tiles x splits), at least 256 pixels per split.
#ifndef LBMNET_WGRAD_BLOCKS
#define LBMNET_WGRAD_BLOCKS 1024
#endif
constexpr int kWgradBlocks = LBMNET_WGRAD_BLOCKS;
void wgrad_shape(const lbm_net* n, const Layer& L, int B, int& splits, int& chunk) {
  const int M = B * H_(n, L.level) * W_(n, L.level);
  const int bn = 16 * conv_nc(L.cpad);
  const int tiles = ((9 * L.cin + 8 + BM - 1) / BM) * ((L.cpad + bn - 1) / bn);
  splits = std::max(1, std::min(kWgradBlocks / std::max(1, tiles), (M + 255) / 256));
  chunk = ((M + splits - 1) / splits + BK - 1) / BK * BK;
  splits = (M + chunk - 1) / chunk;
}

lbm_status conv(lbm_net* n, int li, int B, Act in, Act out, float* outf) {
  const Layer& L = n->layers[li];
  ConvArgs a{};
  a.B = B;
  a.H = H_(n, L.level);
  a.W = W_(n, L.level);
  a.cin = L.cin;
  a.cout = L.cpad;
  a.in = in;
  a.out = out;
  a.w = L.wb;
  a.bias = n->w + L.b_off;
  a.bias_n = L.cr;
  a.outf = outf;
  a.outf_n = L.cr;
  a.relu = L.relu;
  a.mask = Act{nullptr, 0, 0};
  NET_TRY(launch_conv(a, n->stream));
  return LBM_OK;
}

// Backward of layer li given its pre-activation gradient dpre: weight and
// bias gradients into n->g, and (if dx.p) the input gradient, masked by
// (in > 0) when mask_in.
lbm_status conv_bwd(lbm_net* n, int li, int B, Act in, Act dpre, Act dx, bool mask_in) {
  const Layer& L = n->layers[li];
  const int H = H_(n, L.level), W = W_(n, L.level);
  int splits, chunk;
  wgrad_shape(n, L, B, splits, chunk);
  WgradArgs wa{B, H, W, L.cin, L.cpad, in, dpre, n->partial, chunk};
  NET_TRY(launch_wgrad(wa, splits, n->stream));
  const int Mw = 9 * L.cin;
  hipLaunchKernelGGL(k_wreduce, dim3(((Mw + 1) * L.cr + 63) / 64), dim3(256), 0, n->stream, n->partial, splits, Mw,
                     L.cpad, L.cr, n->g + L.w_off);
  NET_TRY(hipGetLastError());
  if (dx.p) {
    ConvArgs a{};
    a.B = B;
    a.H = H;
    a.W = W;
    a.cin = L.cpad;
    a.cout = L.cin;
    a.in = dpre;
    a.out = dx;
    a.w = L.wt;
    a.bias = nullptr;
    a.bias_n = 0;
    a.outf = nullptr;
    a.relu = 0;
    a.mask = mask_in ? in : Act{nullptr, 0, 0};
    NET_TRY(launch_conv(a, n->stream));
  }
  return LBM_OK;
}

unsigned blocks_for(size_t n, unsigned bs = 256) { return unsigned((n + bs - 1) / bs); }

lbm_status pool_fwd(lbm_net* n, int level, int B, Act src, Act dst, int C) {
  const int H = H_(n, level), W = W_(n, level);
  hipLaunchKernelGGL(k_pool_fwd, dim3(blocks_for(size_t(B) * (H / 2) * (W / 2) * (C / 8))), dim3(256), 0, n->stream,
                     src, dst, B, H, W, C);
  NET_TRY(hipGetLastError());
  return LBM_OK;
}
lbm_status pool_bwd(lbm_net* n, int level, int B, Act dpool, Act gskip, Act actv, int C) {
  const int H = H_(n, level), W = W_(n, level);
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(size_t(B) * H * W * (C / 8))), dim3(256), 0, n->stream, dpool, gskip,
                     actv, B, H, W, C);
  NET_TRY(hipGetLastError());
  return LBM_OK;
}
lbm_status up_fwd(lbm_net* n, int level, int B, Act src, Act dst, int C) {
  const int H = H_(n, level), W = W_(n, level);
  hipLaunchKernelGGL(k_up_fwd, dim3(blocks_for(size_t(B) * H * W * (C / 8))), dim3(256), 0, n->stream, src, dst, B, H,
                     W, C);
  NET_TRY(hipGetLastError());
  return LBM_OK;
}
lbm_status up_bwd(lbm_net* n, int level, int B, Act gfine, Act gcoarse, int C) {
  const int H = H_(n, level), W = W_(n, level);
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(size_t(B) * (H / 2) * (W / 2) * (C / 8))), dim3(256), 0, n->stream,
                     gfine, gcoarse, B, H, W, C);
  NET_TRY(hipGetLastError());
  return LBM_OK;
}

#define NS(expr)                  \
  do {                            \
    lbm_status s_ = (expr);       \
    if (s_ != LBM_OK) return s_;  \
  } while (0)

lbm_status forward(lbm_net* n, int B) {
  const int* w = n->cfg.widths;
  const int cw[3] = {w[1] + w[0], w[2] + w[1], w[3] + w[2]};  // cat widths: [up | skip]
  // encoder
  NS(conv(n, 0, B, act(n->x, 8), act(n->a[0], w[0]), nullptr));
  NS(conv(n, 1, B, act(n->a[0], w[0]), act(n->cat[0], cw[0], w[1]), nullptr));
  NS(pool_fwd(n, 0, B, act(n->cat[0], cw[0], w[1]), act(n->pool[1], w[0]), w[0]));
  NS(conv(n, 2, B, act(n->pool[1], w[0]), act(n->a[1], w[1]), nullptr));
  NS(conv(n, 3, B, act(n->a[1], w[1]), act(n->cat[1], cw[1], w[2]), nullptr));
  NS(pool_fwd(n, 1, B, act(n->cat[1], cw[1], w[2]), act(n->pool[2], w[1]), w[1]));
  NS(conv(n, 4, B, act(n->pool[2], w[1]), act(n->a[2], w[2]), nullptr));
  NS(conv(n, 5, B, act(n->a[2], w[2]), act(n->cat[2], cw[2], w[3]), nullptr));
  NS(pool_fwd(n, 2, B, act(n->cat[2], cw[2], w[3]), act(n->pool[3], w[2]), w[2]));
  NS(conv(n, 6, B, act(n->pool[3], w[2]), act(n->a[3], w[3]), nullptr));
  NS(conv(n, 7, B, act(n->a[3], w[3]), act(n->bott, w[3]), nullptr));
  // decoder
  NS(up_fwd(n, 2, B, act(n->bott, w[3]), act(n->cat[2], cw[2], 0), w[3]));
  NS(conv(n, 8, B, act(n->cat[2], cw[2]), act(n->dec[2], w[2]), nullptr));
  NS(up_fwd(n, 1, B, act(n->dec[2], w[2]), act(n->cat[1], cw[1], 0), w[2]));
  NS(conv(n, 9, B, act(n->cat[1], cw[1]), act(n->dec[1], w[1]), nullptr));
  NS(up_fwd(n, 0, B, act(n->dec[1], w[1]), act(n->cat[0], cw[0], 0), w[1]));
  NS(conv(n, 10, B, act(n->cat[0], cw[0]), act(n->dec[0], w[0]), nullptr));
  NS(conv(n, 11, B, act(n->dec[0], w[0]), act(nullptr, 0), n->y));
  return LBM_OK;
}

lbm_status loss_fwd(lbm_net* n, int B, double nf, double* dloss, float* dlossf) {
  const int M = B * n->cfg.height * n->cfg.width;
  const float inv = float(1.0 / (3.0 * std::max(1.0, nf)));
  const unsigned nb = blocks_for(size_t(M));
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, n->stream, n->y, n->t, n->msk, M, inv, n->gy, n->lpartial);
  NET_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, n->stream, n->lpartial, int(nb),
                     1.0 / (3.0 * std::max(1.0, nf)), dloss, dlossf);
  NET_TRY(hipGetLastError());
  return LBM_OK;
}

lbm_status backward(lbm_net* n, int B) {
  const int* w = n->cfg.widths;
  const int cw[3] = {w[1] + w[0], w[2] + w[1], w[3] + w[2]};
  const Act none{nullptr, 0, 0};
  NS(conv_bwd(n, 11, B, act(n->dec[0], w[0]), act(n->gy, 8), act(n->gdec[0], w[0]), true));
  NS(conv_bwd(n, 10, B, act(n->cat[0], cw[0]), act(n->gdec[0], w[0]), act(n->gcat[0], cw[0]), true));
  NS(up_bwd(n, 0, B, act(n->gcat[0], cw[0], 0), act(n->gdec[1], w[1]), w[1]));
  NS(conv_bwd(n, 9, B, act(n->cat[1], cw[1]), act(n->gdec[1], w[1]), act(n->gcat[1], cw[1]), true));
  NS(up_bwd(n, 1, B, act(n->gcat[1], cw[1], 0), act(n->gdec[2], w[2]), w[2]));
  NS(conv_bwd(n, 8, B, act(n->cat[2], cw[2]), act(n->gdec[2], w[2]), act(n->gcat[2], cw[2]), true));
  NS(up_bwd(n, 2, B, act(n->gcat[2], cw[2], 0), act(n->gbott, w[3]), w[3]));
  NS(conv_bwd(n, 7, B, act(n->a[3], w[3]), act(n->gbott, w[3]), act(n->ga[3], w[3]), true));
  NS(conv_bwd(n, 6, B, act(n->pool[3], w[2]), act(n->ga[3], w[3]), act(n->gpool[3], w[2]), false));
  NS(pool_bwd(n, 2, B, act(n->gpool[3], w[2]), act(n->gcat[2], cw[2], w[3]), act(n->cat[2], cw[2], w[3]), w[2]));
  NS(conv_bwd(n, 5, B, act(n->a[2], w[2]), act(n->gcat[2], cw[2], w[3]), act(n->ga[2], w[2]), true));
  NS(conv_bwd(n, 4, B, act(n->pool[2], w[1]), act(n->ga[2], w[2]), act(n->gpool[2], w[1]), false));
  NS(pool_bwd(n, 1, B, act(n->gpool[2], w[1]), act(n->gcat[1], cw[1], w[2]), act(n->cat[1], cw[1], w[2]), w[1]));
  NS(conv_bwd(n, 3, B, act(n->a[1], w[1]), act(n->gcat[1], cw[1], w[2]), act(n->ga[1], w[1]), true));
  NS(conv_bwd(n, 2, B, act(n->pool[1], w[0]), act(n->ga[1], w[1]), act(n->gpool[1], w[0]), false));
  NS(pool_bwd(n, 0, B, act(n->gpool[1], w[0]), act(n->gcat[0], cw[0], w[1]), act(n->cat[0], cw[0], w[1]), w[0]));
  NS(conv_bwd(n, 1, B, act(n->a[0], w[0]), act(n->gcat[0], cw[0], w[1]), act(n->ga[0], w[0]), true));
  NS(conv_bwd(n, 0, B, act(n->x, 8), act(n->ga[0], w[0]), none, false));
  return LBM_OK;
}

lbm_status refresh_weights(lbm_net* n) {
  for (Layer& L : n->layers) {
    hipLaunchKernelGGL(k_wconvert, dim3(blocks_for(size_t(9) * L.cin * L.cpad)), dim3(256), 0, n->stream,
                       n->w + L.w_off, L.cin, L.cr, L.cpad, L.wb, L.wt);
    NET_TRY(hipGetLastError());
  }
  return LBM_OK;
}

// Network input per cell (8 channels):
//   0 fluid indicator (flag >= 2), 1 u_max/0.08, 2 (tau-1/2)/0.7,
//   3 inlet profile shape 4(y-1/2)(ny-3/2-y)/(ny-2)^2, 4 x/(nx-1), 5 y/(ny-1),
//   6 log1p of a 1D hydraulic estimate of the pressure from column x to the
//     outlet in target units: sum over columns x' > x of the plane-Poiseuille
//     gradient 12 nu Q / h(x')^3 (h = open cells of the column, Q the inlet
//     flux), divided by u_max^2,
//   7 log((ny-2)/h(x)), the column's blockage.
// Channels 6-7 carry the domain-wide resistance a 3x3-convolution stack
// cannot integrate over its receptive field.
void host_features(const lbm_net_config& cfg, const uint8_t* flags, double u_max, double tau, float* out) {
  const int H = cfg.height, W = cfg.width;
  const double nu = (tau - 0.5) / 3.0;
  double S = 0.0;  // inlet flux / u_max
  for (int y = 1; y < H - 1; ++y) S += 4.0 * (y - 0.5) * (H - 1.5 - y) / ((H - 2.0) * (H - 2.0));
  std::vector<double> hcol(W), press(W);
  for (int x = 0; x < W; ++x) {
    int h = 0;
    for (int y = 0; y < H; ++y) h += flags[size_t(y) * W + x] >= 2 ? 1 : 0;
    hcol[x] = h > 0 ? double(h) : 1.0;
  }
  double acc = 0.0;
  for (int x = W - 1; x >= 0; --x) {
    press[x] = 12.0 * nu * S * acc / u_max;
    acc += 1.0 / (hcol[x] * hcol[x] * hcol[x]);
  }
  const double hfull = H > 2 ? double(H - 2) : 1.0;
  for (int y = 0; y < H; ++y) {
    const double yy = y;
    const double prof = 4.0 * (yy - 0.5) * (H - 1.5 - yy) / ((H - 2.0) * (H - 2.0));
    for (int x = 0; x < W; ++x) {
      float* f = out + (size_t(y) * W + x) * 8;
      const uint8_t fl = flags[size_t(y) * W + x];
      f[0] = fl >= 2 ? 1.f : 0.f;
      f[1] = float(u_max / 0.08);
      f[2] = float((tau - 0.5) / 0.7);
      f[3] = float(prof);
      f[4] = float(double(x) / double(W - 1));
      f[5] = float(double(y) / double(H - 1));
      f[6] = float(std::log1p(press[x]));
      f[7] = float(std::log(hfull / hcol[x]));
    }
  }
}

lbm_status upload_x(lbm_net* n, const float* x, size_t count) {
  std::vector<uint16_t> h(count);
  for (size_t i = 0; i < count; ++i) h[i] = f2bf(x[i]);
  NET_TRY(hipMemcpyAsync(n->x, h.data(), count * sizeof(uint16_t), hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipStreamSynchronize(n->stream));
  return LBM_OK;
}

void free_all(lbm_net* n) {
  for (Layer& L : n->layers) {
    (void)hipFree(L.wb);
    (void)hipFree(L.wt);
  }
  void* ps[] = {n->w,      n->g,       n->m1,      n->m2,       n->x,       n->bott,      n->gbott, n->gy,
                n->y,      n->t,       n->msk,     n->partial,  n->lpartial, n->dloss,  n->dlosses,
                n->didx,   n->xs,      n->ts,      n->ms};
  for (void* p : ps) (void)hipFree(p);
  for (int l = 0; l < 4; ++l) {
    (void)hipFree(n->a[l]);
    (void)hipFree(n->ga[l]);
    (void)hipFree(n->pool[l]);
    (void)hipFree(n->gpool[l]);
  }
  for (int l = 0; l < 3; ++l) {
    (void)hipFree(n->cat[l]);
    (void)hipFree(n->gcat[l]);
    (void)hipFree(n->dec[l]);
    (void)hipFree(n->gdec[l]);
  }
  if (n->ev0) (void)hipEventDestroy(n->ev0);
  if (n->ev1) (void)hipEventDestroy(n->ev1);
  if (n->stream) (void)hipStreamDestroy(n->stream);
}

lbm_status create_impl(lbm_net* n) {
  const lbm_net_config& c = n->cfg;
  const int* w = c.widths;
  NET_TRY(hipSetDevice(n->device));
  NET_TRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  NET_TRY(hipEventCreate(&n->ev0));
  NET_TRY(hipEventCreate(&n->ev1));
  build_layers(n);
  NS(dalloc(&n->w, n->nparam));
  NS(dalloc(&n->g, n->nparam));
  NS(dalloc(&n->m1, n->nparam));
  NS(dalloc(&n->m2, n->nparam));
  for (Layer& L : n->layers) {
    NS(dalloc(&L.wb, size_t(9) * L.cin * L.cpad));
    NS(dalloc(&L.wt, size_t(9) * L.cin * L.cpad));
  }
  const int cw[3] = {w[1] + w[0], w[2] + w[1], w[3] + w[2]};
  NS(dalloc(&n->x, level_pixels(n, 0) * 8));
  for (int l = 0; l < 4; ++l) {
    NS(dalloc(&n->a[l], level_pixels(n, l) * w[l]));
    NS(dalloc(&n->ga[l], level_pixels(n, l) * w[l]));
    if (l > 0) {
      NS(dalloc(&n->pool[l], level_pixels(n, l) * w[l - 1]));
      NS(dalloc(&n->gpool[l], level_pixels(n, l) * w[l - 1]));
    }
  }
  for (int l = 0; l < 3; ++l) {
    NS(dalloc(&n->cat[l], level_pixels(n, l) * cw[l]));
    NS(dalloc(&n->gcat[l], level_pixels(n, l) * cw[l]));
    NS(dalloc(&n->dec[l], level_pixels(n, l) * w[l]));
    NS(dalloc(&n->gdec[l], level_pixels(n, l) * w[l]));
  }
  NS(dalloc(&n->bott, level_pixels(n, 3) * w[3]));
  NS(dalloc(&n->gbott, level_pixels(n, 3) * w[3]));
  NS(dalloc(&n->gy, level_pixels(n, 0) * 8));
  NS(dalloc(&n->y, level_pixels(n, 0) * 3));
  NS(dalloc(&n->t, level_pixels(n, 0) * 3));
  NS(dalloc(&n->msk, level_pixels(n, 0)));
  size_t pmax = 0;
  for (const Layer& L : n->layers) {
    int splits, chunk;
    wgrad_shape(n, L, c.batch, splits, chunk);
    pmax = std::max(pmax, size_t(splits) * (9 * L.cin + 8) * L.cpad);
  }
  n->partial_n = pmax;
  NS(dalloc(&n->partial, pmax));
  NS(dalloc(&n->lpartial, size_t(blocks_for(level_pixels(n, 0)))));
  NS(dalloc(&n->dloss, 1));
  const size_t HW = size_t(c.height) * c.width;
  if (c.capacity > 0) {
    NS(dalloc(&n->xs, size_t(c.capacity) * HW * 8));
    NS(dalloc(&n->ts, size_t(c.capacity) * HW * 3));
    NS(dalloc(&n->ms, size_t(c.capacity) * HW));
  }
  // He-normal initialisation, biases zero
  std::vector<float> hw(n->nparam, 0.f);
  std::mt19937_64 rng(c.seed);
  for (const Layer& L : n->layers) {
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * L.cin)));
    for (size_t i = 0; i < size_t(9) * L.cin * L.cr; ++i) hw[L.w_off + i] = float(nd(rng));
  }
  NET_TRY(hipMemcpy(n->w, hw.data(), n->nparam * sizeof(float), hipMemcpyHostToDevice));
  NS(refresh_weights(n));
  NET_TRY(hipStreamSynchronize(n->stream));
  return LBM_OK;
}

bool valid_config(const lbm_net_config* c) {
  if (c->height < 8 || c->width < 8 || c->height % 8 || c->width % 8) return false;
  for (int l = 0; l < 4; ++l)
    if (c->widths[l] <= 0 || c->widths[l] % 8) return false;
  if (c->batch < 1 || c->capacity < 0) return false;
  const size_t px = size_t(c->batch) * c->height * c->width;
  return px <= size_t(1) << 31;
}

}  // namespace

extern "C" {

lbm_status lbm_net_create(lbm_net** out, const lbm_net_config* cfg, int32_t device) {
  if (!out || !cfg) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!valid_config(cfg))
    return nfail(LBM_ERR_INVALID_ARG, "net config: height/width must be multiples of 8 (>= 8), widths multiples of 8, batch >= 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return nfail(LBM_ERR_NOT_FOUND, "device %d not present", device);
  lbm_net* n = new lbm_net();
  n->cfg = *cfg;
  n->device = device;
  lbm_status s = create_impl(n);
  if (s != LBM_OK) {
    free_all(n);
    delete n;
    return s;
  }
  *out = n;
  return LBM_OK;
}

lbm_status lbm_net_destroy(lbm_net* n) {
  if (!n) return LBM_OK;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  free_all(n);
  delete n;
  return LBM_OK;
}

lbm_status lbm_net_param_count(lbm_net* n, uint64_t* count) {
  if (!n || !count) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  *count = n->nparam;
  return LBM_OK;
}

lbm_status lbm_net_get_params(lbm_net* n, float* p) {
  if (!n || !p) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  NET_TRY(hipSetDevice(n->device));
  NET_TRY(hipMemcpyAsync(p, n->w, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  NET_TRY(hipStreamSynchronize(n->stream));
  return LBM_OK;
}

lbm_status lbm_net_set_params(lbm_net* n, const float* p) {
  if (!n || !p) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  NET_TRY(hipSetDevice(n->device));
  NET_TRY(hipMemcpyAsync(n->w, p, n->nparam * sizeof(float), hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipMemsetAsync(n->m1, 0, n->nparam * sizeof(float), n->stream));
  NET_TRY(hipMemsetAsync(n->m2, 0, n->nparam * sizeof(float), n->stream));
  n->adam_t = 0;
  NS(refresh_weights(n));
  NET_TRY(hipStreamSynchronize(n->stream));
  return LBM_OK;
}

lbm_status lbm_net_features(lbm_net* n, const uint8_t* flags, double u_max, double tau, float* features) {
  if (!n || !flags || !features) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  host_features(n->cfg, flags, u_max, tau, features);
  return LBM_OK;
}

lbm_status lbm_net_add_sample(lbm_net* n, const uint8_t* flags, double u_max, double tau, const double* rho,
                              const double* u) {
  if (!n || !flags || !rho || !u) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  if (int(n->nf.size()) >= n->cfg.capacity)
    return nfail(LBM_ERR_OVERFLOW, "training set full (capacity %d)", n->cfg.capacity);
  if (!(u_max > 0.0) || !(tau > 0.5)) return nfail(LBM_ERR_INVALID_ARG, "need u_max > 0 and tau > 1/2");
  const size_t HW = size_t(n->cfg.height) * n->cfg.width, k = n->nf.size();
  for (size_t i = 0; i < HW; ++i)
    if (!std::isfinite(rho[i]) || !std::isfinite(u[i]) || !std::isfinite(u[HW + i]))
      return nfail(LBM_ERR_INVALID_ARG, "sample has a non-finite value at cell %zu", i);
  std::vector<float> f(HW * 8);
  host_features(n->cfg, flags, u_max, tau, f.data());
  std::vector<uint16_t> xb(HW * 8);
  for (size_t i = 0; i < HW * 8; ++i) xb[i] = f2bf(f[i]);
  std::vector<float> t(HW * 3);
  std::vector<uint8_t> m(HW);
  double cnt = 0.0;
  const double ru = rho_unit(u_max);
  for (size_t i = 0; i < HW; ++i) {
    m[i] = flags[i] >= 2 ? 1 : 0;
    cnt += m[i];
    t[i * 3 + 0] = float(u[i] / u_max);
    t[i * 3 + 1] = float(u[HW + i] / u_max);
    t[i * 3 + 2] = float((rho[i] - 1.0) / ru);
    if (m[i])
      for (int c = 0; c < 3; ++c) {
        n->tsum[c] += t[i * 3 + c];
        n->tsq[c] += double(t[i * 3 + c]) * t[i * 3 + c];
      }
  }
  n->tcnt += cnt;
  if (n->frozen)
    for (size_t i = 0; i < HW * 3; ++i) t[i] = (t[i] - n->mu[i % 3]) / n->sd[i % 3];
  NET_TRY(hipSetDevice(n->device));
  NET_TRY(hipMemcpyAsync(n->xs + k * HW * 8, xb.data(), HW * 8 * sizeof(uint16_t), hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipMemcpyAsync(n->ts + k * HW * 3, t.data(), HW * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipMemcpyAsync(n->ms + k * HW, m.data(), HW, hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipStreamSynchronize(n->stream));
  n->nf.push_back(cnt);
  return LBM_OK;
}

lbm_status lbm_net_clear_samples(lbm_net* n) {
  if (!n) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  n->nf.clear();
  return LBM_OK;
}

lbm_status lbm_net_sample_count(lbm_net* n, uint32_t* count) {
  if (!n || !count) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  *count = uint32_t(n->nf.size());
  return LBM_OK;
}

lbm_status lbm_net_train(lbm_net* n, uint32_t steps, float lr, uint64_t seed, float* losses) {
  if (!n) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  const uint32_t S = uint32_t(n->nf.size());
  const int B = n->cfg.batch;
  if (S < uint32_t(B)) return nfail(LBM_ERR_INVALID_STATE, "training set holds %u samples, batch is %d", S, B);
  if (steps == 0) return LBM_OK;
  NET_TRY(hipSetDevice(n->device));
  if (!n->frozen && n->tcnt > 0) {
    for (int c = 0; c < 3; ++c) {
      const double m = n->tsum[c] / n->tcnt;
      const double v = std::max(n->tsq[c] / n->tcnt - m * m, 0.0);
      n->mu[c] = float(m);
      n->sd[c] = float(std::max(std::sqrt(v), 1e-6));
    }
    const size_t cnt = size_t(S) * n->cfg.height * n->cfg.width * 3;
    hipLaunchKernelGGL(k_standardize, dim3(blocks_for(cnt)), dim3(256), 0, n->stream, n->ts, cnt, n->mu[0], n->mu[1],
                       n->mu[2], n->sd[0], n->sd[1], n->sd[2]);
    NET_TRY(hipGetLastError());
    n->frozen = true;
  }
  // minibatch indices: a fresh seeded permutation per epoch
  std::vector<uint32_t> idx(size_t(steps) * B), perm(S);
  std::mt19937_64 rng(seed);
  size_t pos = S;
  std::vector<double> nfb(steps, 0.0);
  for (uint32_t k = 0; k < steps; ++k)
    for (int b = 0; b < B; ++b) {
      if (pos + 1 > S) {
        for (uint32_t i = 0; i < S; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        pos = 0;
      }
      idx[size_t(k) * B + b] = perm[pos++];
      nfb[k] += n->nf[idx[size_t(k) * B + b]];
    }
  if (n->didx_n < idx.size()) {
    (void)hipFree(n->didx);
    n->didx = nullptr;
    NS(dalloc(&n->didx, idx.size()));
    n->didx_n = idx.size();
  }
  if (n->dlosses_n < steps) {
    (void)hipFree(n->dlosses);
    n->dlosses = nullptr;
    NS(dalloc(&n->dlosses, steps));
    n->dlosses_n = steps;
  }
  NET_TRY(hipMemcpyAsync(n->didx, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, n->stream));
  const int HW = n->cfg.height * n->cfg.width;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  NET_TRY(hipEventRecord(n->ev0, n->stream));
  for (uint32_t k = 0; k < steps; ++k) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(size_t(B) * HW)), dim3(256), 0, n->stream, n->xs, n->ts, n->ms,
                       n->didx + size_t(k) * B, B, HW, n->x, n->t, n->msk);
    NET_TRY(hipGetLastError());
    NS(forward(n, B));
    NS(loss_fwd(n, B, nfb[k], nullptr, n->dlosses + k));
    NS(backward(n, B));
    n->adam_t += 1;
    const float c1 = float(1.0 - std::pow(0.9, double(n->adam_t)));
    const float c2 = float(1.0 - std::pow(0.999, double(n->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n->nparam)), dim3(256), 0, n->stream, n->w, n->g, n->m1, n->m2,
                       n->nparam, lr, b1, b2, c1, c2, eps);
    NET_TRY(hipGetLastError());
    NS(refresh_weights(n));
  }
  NET_TRY(hipEventRecord(n->ev1, n->stream));
  if (losses)
    NET_TRY(hipMemcpyAsync(losses, n->dlosses, steps * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  NET_TRY(hipStreamSynchronize(n->stream));
  n->timed = true;
  return LBM_OK;
}

lbm_status lbm_net_forward_raw(lbm_net* n, uint32_t count, const float* x, float* y) {
  if (!n || !x || !y) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  NET_TRY(hipSetDevice(n->device));
  const size_t HW = size_t(n->cfg.height) * n->cfg.width;
  for (uint32_t s = 0; s < count; s += n->cfg.batch) {
    const int B = int(std::min<uint32_t>(n->cfg.batch, count - s));
    NS(upload_x(n, x + size_t(s) * HW * 8, size_t(B) * HW * 8));
    NS(forward(n, B));
    NET_TRY(hipMemcpyAsync(y + size_t(s) * HW * 3, n->y, size_t(B) * HW * 3 * sizeof(float), hipMemcpyDeviceToHost,
                           n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
  }
  return LBM_OK;
}

lbm_status lbm_net_predict(lbm_net* n, uint32_t count, const uint8_t* flags, const double* u_max, const double* tau,
                           double* rho, double* u) {
  if (!n || !flags || !u_max || !tau || !rho || !u) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  NET_TRY(hipSetDevice(n->device));
  const size_t HW = size_t(n->cfg.height) * n->cfg.width;
  std::vector<float> xf(size_t(n->cfg.batch) * HW * 8), yf(size_t(n->cfg.batch) * HW * 3);
  NET_TRY(hipEventRecord(n->ev0, n->stream));
  for (uint32_t s = 0; s < count; s += n->cfg.batch) {
    const int B = int(std::min<uint32_t>(n->cfg.batch, count - s));
    for (int b = 0; b < B; ++b) host_features(n->cfg, flags + (s + b) * HW, u_max[s + b], tau[s + b], xf.data() + b * HW * 8);
    NS(upload_x(n, xf.data(), size_t(B) * HW * 8));
    NS(forward(n, B));
    NET_TRY(hipMemcpyAsync(yf.data(), n->y, size_t(B) * HW * 3 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    for (int b = 0; b < B; ++b) {
      const size_t k = s + b;
      const uint8_t* fl = flags + k * HW;
      double* r = rho + k * HW;
      double* ux = u + k * 2 * HW;
      double* uy = ux + HW;
      for (size_t i = 0; i < HW; ++i) {
        const float* o = yf.data() + (b * HW + i) * 3;
        if (fl[i] >= 2) {
          ux[i] = (double(o[0]) * n->sd[0] + n->mu[0]) * u_max[k];
          uy[i] = (double(o[1]) * n->sd[1] + n->mu[1]) * u_max[k];
          r[i] = 1.0 + (double(o[2]) * n->sd[2] + n->mu[2]) * rho_unit(u_max[k]);
        } else {
          ux[i] = 0.0;
          uy[i] = 0.0;
          r[i] = 1.0;
        }
      }
    }
  }
  NET_TRY(hipEventRecord(n->ev1, n->stream));
  NET_TRY(hipEventSynchronize(n->ev1));
  n->timed = true;
  return LBM_OK;
}

lbm_status lbm_net_correct_flux(lbm_net* n, uint32_t count, const uint8_t* flags, const double* u_max, double* u) {
  if (!n || !flags || !u_max || !u) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  const int H = n->cfg.height, W = n->cfg.width;
  const size_t HW = size_t(H) * W;
  double S = 0.0;  // the inlet parabola's flux per unit u_max (the features' profile)
  for (int y = 1; y < H - 1; ++y) S += 4.0 * (y - 0.5) * (H - 1.5 - y) / ((H - 2.0) * (H - 2.0));
  std::vector<double> q(W);
  for (uint32_t k = 0; k < count; ++k) {
    const uint8_t* fl = flags + k * HW;
    double* ux = u + k * 2 * HW;
    double* uy = ux + HW;
    const double q_in = u_max[k] * S;
    std::fill(q.begin(), q.end(), 0.0);
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        if (fl[size_t(y) * W + x] >= 2) q[x] += ux[size_t(y) * W + x];
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const size_t i = size_t(y) * W + x;
        if (fl[i] < 2 || !(q[x] > 0.2 * q_in)) continue;
        const double s = q_in / q[x];
        ux[i] *= s;
        uy[i] *= s;
      }
  }
  return LBM_OK;
}

lbm_status lbm_net_loss_grad_raw(lbm_net* n, uint32_t count, const float* x, const float* t, const float* m,
                                 double* loss, float* grad) {
  if (!n || !x || !t || !m || !loss || !grad) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  if (count < 1 || int(count) > n->cfg.batch) return nfail(LBM_ERR_INVALID_ARG, "count must be 1..batch");
  NET_TRY(hipSetDevice(n->device));
  const int B = int(count);
  const size_t HW = size_t(n->cfg.height) * n->cfg.width, P = size_t(B) * HW;
  NS(upload_x(n, x, P * 8));
  std::vector<uint8_t> mb(P);
  double nf = 0.0;
  for (size_t i = 0; i < P; ++i) {
    mb[i] = m[i] > 0.f ? 1 : 0;
    nf += mb[i];
  }
  NET_TRY(hipMemcpyAsync(n->t, t, P * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
  NET_TRY(hipMemcpyAsync(n->msk, mb.data(), P, hipMemcpyHostToDevice, n->stream));
  NS(forward(n, B));
  NS(loss_fwd(n, B, nf, n->dloss, nullptr));
  NS(backward(n, B));
  NET_TRY(hipMemcpyAsync(loss, n->dloss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
  NET_TRY(hipMemcpyAsync(grad, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  NET_TRY(hipStreamSynchronize(n->stream));
  return LBM_OK;
}

lbm_status lbm_net_get_output_scale(lbm_net* n, float mean[3], float stdev[3]) {
  if (!n || !mean || !stdev) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  for (int c = 0; c < 3; ++c) {
    mean[c] = n->mu[c];
    stdev[c] = n->sd[c];
  }
  return LBM_OK;
}

lbm_status lbm_net_set_output_scale(lbm_net* n, const float mean[3], const float stdev[3]) {
  if (!n || !mean || !stdev) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  if (n->frozen && !n->nf.empty())
    return nfail(LBM_ERR_INVALID_STATE, "training set already standardised with other statistics");
  for (int c = 0; c < 3; ++c) {
    if (!(stdev[c] > 0.f)) return nfail(LBM_ERR_INVALID_ARG, "stdev must be positive");
    n->mu[c] = mean[c];
    n->sd[c] = stdev[c];
  }
  n->frozen = true;
  return LBM_OK;
}

lbm_status lbm_net_last_ms(lbm_net* n, double* ms) {
  if (!n || !ms) return nfail(LBM_ERR_INVALID_ARG, "null argument");
  if (!n->timed) return nfail(LBM_ERR_INVALID_STATE, "no train or predict call yet");
  float f = 0.f;
  NET_TRY(hipEventElapsedTime(&f, n->ev0, n->ev1));
  *ms = f;
  return LBM_OK;
}

}  // extern "C"
