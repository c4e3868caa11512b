// Periodic 3x3x3 convolutions of channels-last meshes on the matrix cores (BF16 MFMA,
// float32 accumulation): the learned potential correction's layers and their gradients.
//
// Implicit GEMM with M = output voxels, N = 16 output channels, K = 27 taps x CIN input
// channels, on v_mfma_f32_16x16x32_bf16.  Work is split into 8x8x8 blocks of output voxels;
// a block's input with a one-voxel periodic halo (10^3 x CIN bf16) is staged in LDS, and each
// wave runs 16-voxel M tiles over the K steps (32 = 32/CIN taps x CIN).
// Lane layout of the MFMA (checked on the device by the conv tests): lane l feeds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8, and receives
// D[4 (l / 16) + i][l % 16], i < 4.
//
// Layouts: activations bf16 [X][Y][Z][C], pre-activations / gradients float32 [X][Y][Z][16],
// weights bf16 [tap][cout][cin] (tap = (dx+1)*9 + (dy+1)*3 + (dz+1)), bias float32 [16].
#include <hip/hip_bf16.h>

#include <algorithm>

#include "pm_common.h"

namespace pm {
namespace conv {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kB = 8;              // output block edge
constexpr int kH = kB + 2;         // staged block edge (one-voxel halo)
constexpr int kVox = kB * kB * kB;  // 512 output voxels per workgroup
constexpr int kThreads = 256;      // 4 waves; each runs 8 of the 32 M tiles
constexpr int kCout = 16;
constexpr int kTaps = 27;

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Block q of a persistent walk over the (nbx, nby, nbz) block grid (hidden layer, fused stem,
// one-channel output layer).  The workgroups one XCD runs at the same time take consecutive q,
// so q is decoded in 4 x 4 x 4 patches of blocks (patches in z-fastest order): the resident
// blocks of an XCD form a compact patch that shares its x and y halo planes in that XCD's L2,
// not a single z line of blocks (1024^3: hidden layer 25.8 -> 24.8 ms; output layer 19.8 ->
// 16.2 ms, its HBM fetch 85 -> 71 GiB; the standalone stem measured slower with it and keeps
// the z-fastest order).  Grids that do not divide into patches keep the z-fastest order.
__device__ __forceinline__ void block_of(int q, int nbx, int nby, int nbz, int& bx, int& by, int& bz) {
  if ((nbx & 3) == 0 && (nby & 3) == 0 && (nbz & 3) == 0) {
    const int p = q >> 6, w = q & 63;
    const int npz = nbz >> 2, npy = nby >> 2;
    const int pz = p % npz, py = (p / npz) % npy, px = p / (npz * npy);
    bz = 4 * pz + (w & 3);
    by = 4 * py + ((w >> 2) & 3);
    bx = 4 * px + (w >> 4);
  } else {
    bz = q % nbz;
    by = (q / nbz) % nby;
    bx = q / (nbz * nby);
  }
}

__device__ __forceinline__ float gelu(float x) {
  // tanh form of GELU (jax.nn.gelu's default approximate=True)
  const float c = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}

// Derivative of the tanh form: with s = sigmoid(2u) = 0.5 (1 + tanh u) and 1 - tanh^2 u =
// 4 s (1 - s), d/dx [0.5 x (1 + tanh u)] = s + 2 x s (1 - s) u', u' = sqrt(2/pi) (1 + 3 *
// 0.044715 x^2): one hardware exp2 and one reciprocal, as gelu_fast, instead of the library
// tanh (about a third of the instructions).  Saturates to 0 and 1.
__device__ __forceinline__ float gelu_grad(float x) {
  constexpr float c = 0.7978845608028654f;
  constexpr float k1 = -2.f * c * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
  constexpr float k3 = k1 * 0.044715f;
  const float x2 = x * x;
  const float s = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * (k1 + k3 * x2)));
  const float du = c + (3.f * 0.044715f * c) * x2;
  return s + 2.f * x * (s * (1.f - s)) * du;
}

// Voxel m (0..15) of M tile `tile`: tiles are 2 y-rows x 8 z of one x plane.
__device__ __forceinline__ void tile_voxel(int tile, int m, int& vx, int& vy, int& vz) {
  vx = tile / 4;
  vy = (tile % 4) * 2 + m / 8;
  vz = m % 8;
}

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;

// GELU's tanh form (jax.nn.gelu's default approximate=True) as x * sigmoid(2u):
// 0.5 x (1 + tanh(u)) = x / (1 + exp(-2u)), u = sqrt(2/pi) (x + 0.044715 x^3), with one
// hardware exp2 and one reciprocal.  Saturates to x / inf -> -0 and x / 1 -> x.
__device__ __forceinline__ float gelu_fast(float x) {
  constexpr float k1 = -2.f * 0.7978845608028654f * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
  constexpr float k3 = k1 * 0.044715f;
  const float e = __builtin_amdgcn_exp2f(x * (k1 + k3 * (x * x)));
  return x * __builtin_amdgcn_rcpf(1.f + e);
}

// ----------------------------------------------------------------------------------------
// Forward convolution, walking x planes.
//
// Blocks are BX x 8 x BZ = 512 output voxels, staged with a one-voxel halo (SX x SY x SZ).
// With BZ = 8 an M tile is 2 y-rows x 8 z of one x plane; with BZ = 16 (BX = 4) it is one
// y-row of 16 consecutive z, so the 16 lanes of a fragment read 256 contiguous bytes of a
// channel half (all 64 LDS banks once; the 2 x 8 tile's rows sit SZ voxels apart and collide)
// and an epilogue store covers 512 contiguous bytes.
// Epilogue modes of the forward: kPlain writes pre and/or act; kOutResidual is the output
// layer fused with the correction's residual stage (inference): rho_add = gain (rho - 1) +
// amp tanh(pre channel 0), the same arithmetic as k_residual of pm_correction.hip; kBwdData*
// are a 16 -> 16 layer's data gradient run as a forward convolution (see k_conv_fwd): dz and
// dx float32 (kBwdData), both bf16 (kBwdData16), or dz bf16 and dx float32 (kBwdData16F).
enum Epilogue { kPlain = 0, kOutResidual = 1, kBwdData = 2, kBwdData16 = 3, kBwdData16F = 4 };

template <int BZ_>
struct FwdGeom {
  static constexpr int BZ = BZ_;
  static constexpr int BY = 8;
  static constexpr int BX = kVox / (BY * BZ);  // 8 or 4
  static constexpr int SX = BX + 2, SY = BY + 2, SZ = BZ + 2;
  static constexpr int NS = SX * SY * SZ;  // staged voxels
  static constexpr int ROWS = BZ / 8;      // y-rows a wave walks one after the other
  static_assert(BZ == 8 || BZ == 16, "block z extent");
};

// The 27 taps are grouped by dx: within one dx the 9 (dy, dz) taps j = 3 dy' + dz' (dy', dz'
// in 0..2) are spread over KS K steps of TPS taps (TPS = 32 / CIN; padded slots past j = 8 have
// zero weights).  A wave walks the SX staged x planes of its tile rows: staged plane p feeds
// output planes p (dx = -1), p - 1 (dx = 0) and p - 2 (dx = +1), so the wave reads plane p's KS
// fragments from LDS once and issues 3 KS MFMAs into three rotating accumulators, and reads
// plane p + 1 while those run: KS LDS reads per M tile and plane instead of one per K step,
// with the read latency covered by the previous plane's matrix work.
template <int CIN, int SZ>
struct PlaneTaps {
  static constexpr int TPS = 32 / CIN;
  static constexpr int KS = (9 + TPS - 1) / TPS;
  // staged (y, z) offset of in-plane tap j
  static constexpr int off2(int j) { return (j / 3) * SZ + j % 3; }
  // offset of slot t of step k relative to the step's slot 0 (padded slots: 0)
  static constexpr int diff(int k, int t) { return k * TPS + t < 9 ? off2(k * TPS + t) - off2(k * TPS) : 0; }
  // steps whose slots have the same relative offsets share a per-lane LDS base (a class)
  int cls[KS] = {};
  int rep[KS] = {};
  int ncls = 0;
  constexpr PlaneTaps() {
    for (int k = 0; k < KS; ++k) {
      int c = -1;
      for (int q = 0; q < ncls && c < 0; ++q) {
        bool same = true;
        for (int t = 1; t < TPS; ++t) same = same && diff(k, t) == diff(rep[q], t);
        if (same) c = q;
      }
      if (c < 0) {
        c = ncls;
        rep[ncls++] = k;
      }
      cls[k] = c;
    }
  }
};

template <int CIN, bool GELU, bool RES, int EPI, int BZ>
struct PlaneWalk {
  using Geo = FwdGeom<BZ>;
  using T = PlaneTaps<CIN, Geo::SZ>;
  static constexpr int KS = T::KS;
  static constexpr int BX = Geo::BX;
  static constexpr T taps{};
  // weights of (dx' = 0..2, step k): A[cout = m][k = 8 g + j]
  bf16x8 wa[3 * KS];
  float bsv[4];
  int cbase[4];  // per-lane LDS element offset of each step class (half, lane voxel, slot)
  bf16x8 frag[2][KS];
  f32x4 acc[3];
  const __bf16* xin;
  // epilogue
  int64_t el;     // lane's output voxel in plane 0 (row 0) of the block
  int64_t plane;  // voxels per x plane of the mesh
  int64_t rowz;   // voxels per y row of the mesh (n2)
  int g;
  const __bf16* res;
  float* pre;
  __bf16* act;
  int pch;
  // per-tile values loaded at the start of the block (before the next block's prefetch, so
  // that waiting for them does not wait for the prefetch): the residual's 4 channels (RES)
  // or the density (kOutResidual)
  typedef uint32_t Early __attribute__((ext_vector_type(2 * kB)));
  Early early;
  const float* rho;
  float* rho_add;
  float gain, amp;
  const float* addf;  // kBwdData*: gradient added before the GELU derivative (or null; bf16 data
                      // when dz is bf16)
  const float* gact;  // kBwdData*: pre-activation whose GELU derivative scales the result (or null)

  __device__ __forceinline__ int64_t voxel(int r, int t) const { return el + r * rowz + t * plane; }

  __device__ __forceinline__ void load_early() {
#pragma unroll
    for (int r = 0; r < Geo::ROWS; ++r) {
#pragma unroll
      for (int t = 0; t < BX; ++t) {
        const int q = r * BX + t;
        if constexpr (RES) {
          const uint2 v = *reinterpret_cast<const uint2*>(res + voxel(r, t) * kCout + 4 * g);
          early[2 * q] = v.x;
          early[2 * q + 1] = v.y;
        } else if constexpr (EPI == kOutResidual) {
          early[q] = __float_as_uint(rho[voxel(r, t)]);
        }
      }
    }
  }

  __device__ __forceinline__ void read_plane(int buf, int r, int p) {
#pragma unroll
    for (int k = 0; k < KS; ++k)
      frag[buf][k] = *reinterpret_cast<const bf16x8*>(
          xin + cbase[taps.cls[k]] + (r * Geo::SZ + p * Geo::SY * Geo::SZ + T::off2(k * T::TPS)) * 8);
  }

  __device__ __forceinline__ void finish(int r, int t, const f32x4& c) {
    // lane holds output channels 4g .. 4g+3 of its voxel in output plane t of row r
    const int64_t e = voxel(r, t);
    const int q = r * BX + t;
    if constexpr (EPI >= kBwdData) {
      // dx = (sum + add) * gelu'(gact), channels 4g .. 4g+3 of the voxel; with pch = 4 only
      // channels 0 .. 3 are written (dx has 4 channels per voxel)
      if (pch != kCout && g != 0) return;
      f32x4 v = c;
      if (addf) {
        if constexpr (EPI == kBwdData) {
          const f32x4 a = *reinterpret_cast<const f32x4*>(addf + e * kCout + 4 * g);
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] += a[i];
        } else {
          const bf16x4 a = *reinterpret_cast<const bf16x4*>(reinterpret_cast<const __bf16*>(addf) + e * kCout + 4 * g);
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] += (float)a[i];
        }
      }
      if (gact) {
        const f32x4 ga = *reinterpret_cast<const f32x4*>(gact + e * kCout + 4 * g);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] *= gelu_grad(ga[i]);
      }
      if constexpr (EPI == kBwdData16) {
        bf16x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = (__bf16)v[i];
        *reinterpret_cast<bf16x4*>(act + e * pch + 4 * g) = o;
      } else {
        *reinterpret_cast<f32x4*>(pre + e * pch + 4 * g) = v;
      }
      return;
    }
    float zv[4];
    bf16x4 rv;
    if constexpr (RES) rv = __builtin_bit_cast(bf16x4, make_uint2(early[2 * q], early[2 * q + 1]));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      zv[i] = c[i];  // the bias was the accumulator's initial value
      if constexpr (RES) zv[i] += (float)rv[i];
    }
    if constexpr (EPI == kOutResidual) {
      if (g == 0) rho_add[e] = gain * (__uint_as_float(early[q]) - 1.f) + amp * tanhf(zv[0]);
      return;
    }
    if (pre) {
      if (pch == kCout)
        *reinterpret_cast<f32x4*>(pre + e * kCout + 4 * g) = f32x4{zv[0], zv[1], zv[2], zv[3]};
      else if (g == 0)
        pre[e] = zv[0];
    }
    if (act) {
      bf16x4 o;
#pragma unroll
      for (int i = 0; i < 4; ++i) o[i] = (__bf16)(GELU ? gelu_fast(zv[i]) : zv[i]);
      *reinterpret_cast<bf16x4*>(act + e * kCout + 4 * g) = o;
    }
  }

  // staged plane P of row R: its fragments are in frag[P % 2]; plane P + 1 is read into the
  // other buffer (and after the last plane of a row, plane 0 of the next row)
  template <int R, int P>
  __device__ __forceinline__ void walk() {
    if constexpr (P + 1 < Geo::SX)
      read_plane((P + 1) % 2, R, P + 1);
    else if constexpr (R + 1 < Geo::ROWS)
      read_plane((P + 1) % 2, R + 1, 0);
    if constexpr (P < BX) acc[P % 3] = f32x4{bsv[0], bsv[1], bsv[2], bsv[3]};  // the bias enters as C
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      // output plane t = P - dx' takes dx' = P - t
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int t = P - d;
        if (t >= 0 && t < BX)
          acc[(P - d + 3) % 3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[d * KS + k], frag[P % 2][k],
                                                                       acc[(P - d + 3) % 3], 0, 0, 0);
      }
    }
    if constexpr (P >= 2) finish(R, P - 2, acc[(P - 2) % 3]);
    if constexpr (P + 1 < Geo::SX)
      walk<R, P + 1>();
    else if constexpr (R + 1 < Geo::ROWS)
      walk<R + 1, 0>();  // the plane count per row is even, so frag parity carries over
  }
};

// Forward: pre = conv(x) + bias (+ res), act = gelu(pre) (or pre when !GELU), written as
// float32 pre-activations (when pre != null; the first pch of the 16 channels per voxel) and
// bf16 activations (when act != null).
//
// A persistent grid walks the output blocks; gridDim.x is a multiple of 8 and workgroup w (on
// XCD w % 8) takes block it * G + (w % 8) * (G / 8) + w / 8 of iteration it, so the blocks of
// one iteration on one XCD are consecutive (z neighbours share their halo planes in that XCD's
// L2).  The weights stay in registers as MFMA A fragments for the whole kernel
// (D[cout][voxel] = W[cout][K] X[K][voxel]); the next block's input tile is loaded into
// registers while the current block is computed from LDS, so the loads of one block are in
// flight together and overlap the matrix work of the previous one.  The staged tile is stored
// as one plane per 8-channel half ([half][voxel][8]).  The global addresses of a block are the
// thread's fixed offsets plus the block origin, except for blocks whose halo wraps around the
// mesh.
//
// EPI = kBwdData runs a 16 -> 16 layer's data gradient through the same walk: dx[v] =
// sum_tap dz[v - off(tap)] w[tap]^T is the forward convolution of dz with the taps mirrored
// (tap -> 26 - tap, off -> -off) and each tap's weight matrix transposed, so the A fragments
// are loaded as A[ci][(tap, co)] = w[26 - tap][co][ci] (once per workgroup), x is the float32
// dz (rounded to bf16 when staged, as the matrix cores read it), there is no bias, and the
// epilogue writes dx = (sum + addf) * gelu'(gact) to pre as float32 (pch = 16 channels per
// voxel, or 4: only the first four).  kBwdData16 reads a bf16 dz (and add) and writes dx to
// act as bf16; kBwdData16F reads a bf16 dz and writes float32: a gradient that goes on to
// another data gradient only is rounded to bf16 there anyway, so it is stored as bf16.
template <int CIN, bool GELU, bool RES, int EPI = kPlain, int BZ = 8>
__global__ void __launch_bounds__(kThreads, 2) k_conv_fwd(const __bf16* __restrict__ x, const __bf16* __restrict__ w,
                                                          const float* __restrict__ bias,
                                                          const __bf16* __restrict__ res, float* __restrict__ pre,
                                                          __bf16* __restrict__ act, int n0, int n1, int n2, int pch,
                                                          const float* __restrict__ rho = nullptr,
                                                          const float* __restrict__ stats = nullptr, float c = 0.f,
                                                          float r = 0.f, float* __restrict__ rho_add = nullptr,
                                                          const float* __restrict__ addf = nullptr,
                                                          const float* __restrict__ gact = nullptr) {
  using Geo = FwdGeom<BZ>;
  using Walk = PlaneWalk<CIN, GELU, RES, EPI, BZ>;
  constexpr bool BWD = EPI >= kBwdData;
  constexpr bool F32IN = EPI == kBwdData;  // dz staged from float32
  static_assert(!BWD || (CIN == 16 && !GELU && !RES), "data gradient: 16 -> 16, plain");
  constexpr int PW = F32IN ? 8 : 4;  // prefetch words per staged 8-channel vector (float32 or bf16)
  const float* __restrict__ xf = reinterpret_cast<const float*>(x);
  using Taps = PlaneTaps<CIN, Geo::SZ>;
  constexpr int TPS = Taps::TPS, KS = Taps::KS;
  constexpr int V = CIN / 8;  // 8-channel halves of a voxel (LDS planes)
  constexpr int NS = Geo::NS, SZ = Geo::SZ, SY = Geo::SY;
  constexpr int BX = Geo::BX, BY = Geo::BY;
  constexpr int ITEMS = NS * V;  // 16-byte vectors staged per block
  constexpr int PER = (ITEMS + kThreads - 1) / kThreads;
  constexpr Taps taps{};
  static_assert(taps.ncls <= 4, "step classes");
  static_assert(Geo::SX % 2 == 0, "even plane count (fragment buffer parity across rows)");
  __shared__ __align__(16) __bf16 xin[V * NS * 8];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  // K group g of a step: tap slot gt = g / V, channel half gh = g % V
  const int gt = g / V, gh = g % V;
  Walk wk;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int j = k * TPS + gt;
      const int tap = d * 9 + (j < 9 ? j : 0);
      if constexpr (BWD) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
          wk.wa[d * KS + k][i] = j < 9 ? w[((kTaps - 1 - tap) * kCout + 8 * gh + i) * kCout + m] : (__bf16)0.f;
      } else {
        bf16x8 v = *reinterpret_cast<const bf16x8*>(w + (tap * kCout + m) * CIN + 8 * gh);
#pragma unroll
        for (int i = 0; i < 8; ++i) wk.wa[d * KS + k][i] = j < 9 ? v[i] : (__bf16)0.f;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) wk.bsv[i] = BWD ? 0.f : bias[4 * g + i];
  // weights and bias are in registers before the block loop starts (otherwise the wait for
  // them at the loop head also waits for the prefetch in flight)
  __builtin_amdgcn_s_waitcnt(0);
  // this lane's voxel in the block's plane 0 (row 0): BZ = 8: y = 2 wave + m / 8, z = m % 8;
  // BZ = 16: y = 2 wave (+ the row walked), z = m
  const int ly = BZ == 8 ? 2 * wave + m / 8 : 2 * wave;
  const int lz = BZ == 8 ? m % 8 : m;
  {
    const int corner = ly * SZ + lz;  // staged corner of its taps
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      int d = 0;
#pragma unroll
      for (int t = 1; t < TPS; ++t) d = gt == t ? Taps::diff(taps.rep[cc < taps.ncls ? cc : 0], t) : d;
      wk.cbase[cc] = (gh * NS + corner + d) * 8;
    }
  }
  wk.xin = xin;
  wk.plane = (int64_t)n1 * n2;
  wk.rowz = n2;
  wk.g = g;
  wk.res = res;
  wk.pre = pre;
  wk.act = act;
  wk.pch = pch;
  wk.rho = rho;
  wk.rho_add = rho_add;
  wk.addf = addf;
  wk.gact = gact;
  if constexpr (EPI == kOutResidual) {
    wk.gain = stats[1];
    wk.amp = r * stats[0] / c;
  }
  // this thread's staged vectors: element offset from the block's (-1, -1, -1) corner when
  // the halo does not wrap (fits int32: the caller checks SX * n1 * n2 * CIN < 2^31)
  typedef int Offsets __attribute__((ext_vector_type(PER)));
  Offsets loc;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int it = min((int)threadIdx.x + u * kThreads, ITEMS - 1);
    const int v = it % V, vox = it / V;
    const int z = vox % SZ, y = (vox / SZ) % SY, xx = vox / (SZ * SY);
    loc[u] = ((xx * n1 + y) * n2 + z) * CIN + 8 * v;
  }
  const int nbz = n2 / BZ, nby = n1 / BY;
  const int nblk = (n0 / BX) * nby * nbz;  // < 2^31: the caller checks the mesh size
  const int G = gridDim.x, per = G / 8, first = (blockIdx.x % 8) * per + blockIdx.x / 8;
  // iteration i loads block nxt into registers, computes block cur (staged in LDS by the
  // previous iteration) while those loads are in flight, then stages nxt
  typedef uint32_t Prefetch __attribute__((ext_vector_type(PW * PER)));  // a value, never a private array
  // one staged 8-channel vector at element offset o of x: 16 bytes of bf16, or 32 of float32
  auto fetch = [&](Prefetch& pf, int u, int64_t o) {
    if constexpr (F32IN) {
      const float4 t0 = *reinterpret_cast<const float4*>(xf + o);
      const float4 t1 = *reinterpret_cast<const float4*>(xf + o + 4);
      pf[8 * u] = __float_as_uint(t0.x);
      pf[8 * u + 1] = __float_as_uint(t0.y);
      pf[8 * u + 2] = __float_as_uint(t0.z);
      pf[8 * u + 3] = __float_as_uint(t0.w);
      pf[8 * u + 4] = __float_as_uint(t1.x);
      pf[8 * u + 5] = __float_as_uint(t1.y);
      pf[8 * u + 6] = __float_as_uint(t1.z);
      pf[8 * u + 7] = __float_as_uint(t1.w);
    } else {
      const uint4 t = *reinterpret_cast<const uint4*>(x + o);
      pf[4 * u] = t.x;
      pf[4 * u + 1] = t.y;
      pf[4 * u + 2] = t.z;
      pf[4 * u + 3] = t.w;
    }
  };
  int cur = -1, nxt = first;
  for (;;) {
    if (cur >= 0) {
      int bx, by, bz;
      block_of(cur, n0 / BX, nby, nbz, bx, by, bz);
      const int64_t e0 = (((int64_t)bx * BX * n1 + by * BY) * n2 + bz * BZ);  // block origin voxel
      wk.el = e0 + (int64_t)ly * n2 + lz;
      wk.load_early();
    }
    Prefetch pf;
    if (nxt < nblk) {
      int bx, by, bz;
      block_of(nxt, n0 / BX, nby, nbz, bx, by, bz);
      const int ox = bx * BX, oy = by * BY, oz = bz * BZ;
      const bool inner = ox > 0 && ox + BX < n0 && oy > 0 && oy + BY < n1 && oz > 0 && oz + BZ < n2;
      if (inner) {
        const int64_t base = (((int64_t)(ox - 1) * n1 + (oy - 1)) * n2 + (oz - 1)) * CIN;
#pragma unroll
        for (int u = 0; u < PER; ++u) fetch(pf, u, base + loc[u]);  // unconditional (clamped) loads, all in flight
      } else {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int it = min((int)threadIdx.x + u * kThreads, ITEMS - 1);
          const int v = it % V, vox = it / V;
          const int z = vox % SZ, y = (vox / SZ) % SY, xx = vox / (SZ * SY);
          const int gx = wrap(ox + xx - 1, n0), gy = wrap(oy + y - 1, n1), gz = wrap(oz + z - 1, n2);
          fetch(pf, u, (((int64_t)gx * n1 + gy) * n2 + gz) * CIN + 8 * v);
        }
      }
    }
    if (cur >= 0) {
      wk.read_plane(0, 0, 0);
      wk.template walk<0, 0>();
    }
    if (nxt >= nblk) break;
    __syncthreads();  // the MFMAs of block cur are done with the tile
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = threadIdx.x + u * kThreads;
      if (i < ITEMS) {
        if constexpr (F32IN) {
          bf16x8 b;
#pragma unroll
          for (int q = 0; q < 8; ++q) b[q] = (__bf16)__uint_as_float(pf[8 * u + q]);
          *reinterpret_cast<bf16x8*>(xin + ((i % V) * NS + i / V) * 8) = b;
        } else {
          *reinterpret_cast<uint4*>(xin + ((i % V) * NS + i / V) * 8) =
              make_uint4(pf[4 * u], pf[4 * u + 1], pf[4 * u + 2], pf[4 * u + 3]);
        }
      }
    }
    __syncthreads();
    cur = nxt;
    nxt += G;
  }
}

// Inference stem straight from the density mesh: act = gelu(conv(x0; w) + bias) where x0 is
// the 8-channel input [bf16((rho - 1) c / sigma), bf16(a), bf16(Omega_m), bf16(sigma8), 0 ...]
// that pm_correction_features writes (PHI: [source, bf16(phi / sigma_p), a, Omega_m, sigma8,
// 0 ...] with the potential channel).  The a, Omega_m and sigma8 channels are the same at
// every voxel, so their periodic convolution is sum_tap w[tap][cout][ch] * value, a constant
// per output channel that is folded into the bias; the last channels are zero.  What remains
// is a one-channel convolution (two with PHI) whose 27 taps fit one K = 32 MFMA step per
// channel (5 padded taps with zero weights): lane (g, m) feeds taps 8g .. 8g+7 of voxel m, read
// as eight 2-byte LDS values.  The staged tiles are the bf16 inputs of the block's 10^3 halo,
// computed from the float32 meshes on load with the same arithmetic as the feature kernel.
// Same persistent, prefetching block walk as k_conv_fwd.
constexpr int kStemCin = 8;  // channels of the stem's weight tensor

// the stem's bias with the constant input channels folded in (fixed order), for output channel co
template <bool PHI>
__device__ __forceinline__ float stem_bias(const __bf16* __restrict__ w, const float* __restrict__ bias, int co,
                                           float a, float om, float s8) {
  constexpr int C0 = PHI ? 2 : 1;  // first constant channel
  const __bf16 ba = (__bf16)a, bo = (__bf16)om, bs = (__bf16)s8;
  float t = 0.f;
  for (int tap = 0; tap < kTaps; ++tap) {
    const __bf16* wt = w + (tap * kCout + co) * kStemCin;
    t += (float)wt[C0] * (float)ba + (float)wt[C0 + 1] * (float)bo + (float)wt[C0 + 2] * (float)bs;
  }
  return bias[co] + t;
}

// A fragment of the stem's one-channel MFMA: channel ch's weights of taps 8g .. 8g+7 for output
// channel m (the padded taps 27 .. 31 are zero)
__device__ __forceinline__ bf16x8 stem_fragment(const __bf16* __restrict__ w, int ch, int g, int m) {
  bf16x8 f;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tap = 8 * g + j;
    f[j] = tap < kTaps ? w[(tap * kCout + m) * kStemCin + ch] : (__bf16)0.f;
  }
  return f;
}

// The same one-channel walk computes the data gradient of the one-output-channel layer (MODE
// kStemBwdOut1): dx[v][ci] = gelu'(gact[v][ci]) sum_tap dz0[v - off(tap)] w[tap][0][ci] is a
// one-channel convolution of dz0 with the taps mirrored, so lane (g, m) holds
// w[26 - tap][0][m] for taps 8g .. 8g+7, the staged tile is bf16(dz0), and the epilogue writes
// float32 (kStemBwdOut1Bf16: bf16, to act; no bias, no GELU).
enum StemMode { kStemDensity = 0, kStemPotential = 1, kStemBwdOut1 = 2, kStemBwdOut1Bf16 = 3 };

template <int MODE>
__global__ void __launch_bounds__(kThreads, 4) k_conv_stem(const float* __restrict__ rho,
                                                           const float* __restrict__ phi, float c,
                                                           const float* __restrict__ stats, float a, float om,
                                                           float s8, const __bf16* __restrict__ w,
                                                           const float* __restrict__ bias, __bf16* __restrict__ act,
                                                           int n0, int n1, int n2, float* __restrict__ dxo = nullptr,
                                                           const float* __restrict__ gact = nullptr) {
  constexpr bool PHI = MODE == kStemPotential;
  constexpr bool BWD1 = MODE == kStemBwdOut1 || MODE == kStemBwdOut1Bf16;
  constexpr int NS = kH * kH * kH;
  constexpr int PER = (NS + kThreads - 1) / kThreads;
  constexpr int NT = PHI ? 2 : 1;  // staged input channels
  __shared__ __align__(16) __bf16 xin[NT][PER * kThreads];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  bf16x8 wa;
  if constexpr (BWD1) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int tap = 8 * g + j;
      wa[j] = tap < kTaps ? w[((kTaps - 1 - tap) * kCout) * kCout + m] : (__bf16)0.f;
    }
  } else {
    wa = stem_fragment(w, 0, g, m);
  }
  const bf16x8 wp = PHI ? stem_fragment(w, 1, g, m) : bf16x8{};
  float bsv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bsv[i] = BWD1 ? 0.f : stem_bias<PHI>(w, bias, 4 * g + i, a, om, s8);
  const float inv = BWD1 ? 1.f : 1.f / stats[0];
  const float pinv = PHI ? 1.f / stats[3] : 0.f;
  __builtin_amdgcn_s_waitcnt(0);
  // staged offsets of this lane's 8 taps (padded taps read tap 26's voxel; zero weights)
  typedef int TapOffsets __attribute__((ext_vector_type(8)));
  TapOffsets toff;
  {
    const int corner = (2 * wave + m / 8) * kH + m % 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int t = min(8 * g + j, kTaps - 1);
      toff[j] = corner + ((t / 9) * kH + (t / 3) % 3) * kH + t % 3;
    }
  }
  typedef int Offsets __attribute__((ext_vector_type(PER)));
  Offsets loc;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int vox = min((int)threadIdx.x + u * kThreads, NS - 1);
    const int z = vox % kH, y = (vox / kH) % kH, xx = vox / (kH * kH);
    loc[u] = (xx * n1 + y) * n2 + z;
  }
  const int nbz = n2 / kB, nby = n1 / kB;
  const int nblk = (n0 / kB) * nby * nbz;
  const int G = gridDim.x, per = G / 8, first = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int64_t plane = (int64_t)n1 * n2;
  typedef float Prefetch __attribute__((ext_vector_type(PER)));
  int cur = -1, nxt = first;
  for (;;) {
    Prefetch pf, pp;
    if (nxt < nblk) {
      const int bz = nxt % nbz, by = (nxt / nbz) % nby, bx = nxt / (nbz * nby);
      const int ox = bx * kB, oy = by * kB, oz = bz * kB;
      const bool inner = ox > 0 && ox + kB < n0 && oy > 0 && oy + kB < n1 && oz > 0 && oz + kB < n2;
      if (inner) {
        const int64_t o = ((int64_t)(ox - 1) * n1 + (oy - 1)) * n2 + (oz - 1);
#pragma unroll
        for (int u = 0; u < PER; ++u) pf[u] = rho[o + loc[u]];
        if constexpr (PHI) {
#pragma unroll
          for (int u = 0; u < PER; ++u) pp[u] = phi[o + loc[u]];
        }
      } else {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int vox = min((int)threadIdx.x + u * kThreads, NS - 1);
          const int z = vox % kH, y = (vox / kH) % kH, xx = vox / (kH * kH);
          const int gx = wrap(ox + xx - 1, n0), gy = wrap(oy + y - 1, n1), gz = wrap(oz + z - 1, n2);
          const int64_t o = ((int64_t)gx * n1 + gy) * n2 + gz;
          pf[u] = rho[o];
          if constexpr (PHI) pp[u] = phi[o];
        }
      }
    }
    if (cur >= 0) {
      const int bz = cur % nbz, by = (cur / nbz) % nby, bx = cur / (nbz * nby);
      const int64_t e0 = (((int64_t)bx * kB * n1 + by * kB) * n2 + bz * kB);
      const int64_t el = e0 + (int64_t)(2 * wave + m / 8) * n2 + m % 8;
#pragma unroll
      for (int q = 0; q < kB; ++q) {  // output x plane q: staged planes q .. q+2
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = xin[0][toff[j] + q * kH * kH];
        f32x4 cv = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, b, f32x4{bsv[0], bsv[1], bsv[2], bsv[3]}, 0, 0, 0);
        if constexpr (PHI) {
          bf16x8 bp;
#pragma unroll
          for (int j = 0; j < 8; ++j) bp[j] = xin[1][toff[j] + q * kH * kH];
          cv = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wp, bp, cv, 0, 0, 0);
        }
        const int64_t e = el + q * plane;
        if constexpr (BWD1) {
          f32x4 v = cv;
          if (gact) {
            const f32x4 ga = *reinterpret_cast<const f32x4*>(gact + e * kCout + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] *= gelu_grad(ga[i]);
          }
          if constexpr (MODE == kStemBwdOut1Bf16) {
            bf16x4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) o[i] = (__bf16)v[i];
            *reinterpret_cast<bf16x4*>(act + e * kCout + 4 * g) = o;
          } else {
            *reinterpret_cast<f32x4*>(dxo + e * kCout + 4 * g) = v;
          }
        } else {
          bf16x4 o;
#pragma unroll
          for (int i = 0; i < 4; ++i) o[i] = (__bf16)gelu_fast(cv[i]);
          *reinterpret_cast<bf16x4*>(act + e * kCout + 4 * g) = o;
        }
      }
    }
    if (nxt >= nblk) break;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = threadIdx.x + u * kThreads;
      if (i < NS) {
        xin[0][i] = BWD1 ? (__bf16)pf[u] : (__bf16)(((pf[u] - 1.f) * c) * inv);
        if constexpr (PHI) xin[1][i] = (__bf16)(pp[u] * pinv);
      }
    }
    __syncthreads();
    cur = nxt;
    nxt += G;
  }
}

// Inference stem fused with a first hidden layer that is a plain GELU conv:
// act = gelu(conv16(gelu(stem(rho))) + b1).  The hidden layer's staged input tile (6 x 10 x 18
// voxels of 16 channels around a 4 x 8 x 16 block) is computed in LDS from the density instead
// of being read from HBM: the block stages the bf16 source of its 8 x 12 x 20 density tile
// (two-voxel halo), runs the stem's one-step MFMA over the 1080 staged voxels (68 tiles of 16
// voxels; lane (g, m) feeds taps 8g .. 8g+7 of voxel m with the same arithmetic as
// k_conv_stem, so the activations are k_conv_stem's bit for bit), writes their GELU as bf16
// into the tile and walks the hidden conv over it exactly as k_conv_fwd does (PHI: the
// potential's tile is staged beside the density's and the stem adds its second MFMA step, as
// k_conv_stem<kStemPotential>).  The stem
// activation (32 bytes per voxel written, then read back 2.8x over by the hidden layer's halo
// tiles) never reaches HBM; the price is the stem's MFMA and GELU on more voxels than the
// block has.  The stem's GELUs, not the matrix cores, bound the kernel (per 16 voxels a stem
// tile issues ~38 VALU instructions and 8 quarter-rate transcendentals against one MFMA), so
// the grid walks runs of consecutive blocks along x: block bx's staged planes 0 and 1 are
// block bx - 1's planes 4 and 5, copied within LDS, and only planes 2 .. 5 (720 voxels, 1.41
// per output voxel instead of 2.11) go through the stem.  A work item is one (y, z) column
// of blocks and a range of `run` x blocks; the workgroups one XCD runs at once take
// consecutive items (neighbouring columns).  Persistent and prefetching (the next block's
// density tile in registers while the current block is computed).
constexpr int kFRX = 8, kFRY = 12, kFRZ = 20;                // density tile: block + two-voxel halo
constexpr int kFRN = kFRX * kFRY * kFRZ;                     // 1920
constexpr int kFRPer = (kFRN + kThreads - 1) / kThreads;     // 8
// LDS plane stride of the density tile (bf16 elements; rows stay kFRZ apart): with the
// unpadded 240 the four lane groups of a stem gather (taps 8g + j, one x plane apart at the
// extremes) overlap in bank windows, 2.14 distinct dwords per bank on average over all tiles
// and taps, up to 3; with 292 it is 1.31 (an exhaustive search over row and plane strides)
constexpr int kFRP = 292;
constexpr int kFRL = kFRX * kFRP;

template <bool PHI>
__global__ void __launch_bounds__(kThreads, 2) k_conv_stem_hidden(const float* __restrict__ rho,
                                                                  const float* __restrict__ phi, float c,
                                                                  const float* __restrict__ stats, float a, float om,
                                                                  float s8, const __bf16* __restrict__ ws,
                                                                  const float* __restrict__ bs,
                                                                  const __bf16* __restrict__ wh,
                                                                  const float* __restrict__ bh,
                                                                  __bf16* __restrict__ act, int n0, int n1, int n2,
                                                                  int run) {
  using Geo = FwdGeom<16>;
  using Walk = PlaneWalk<16, true, false, kPlain, 16>;
  using Taps = PlaneTaps<16, Geo::SZ>;
  constexpr int TPS = Taps::TPS, KS = Taps::KS;
  constexpr int NS = Geo::NS, SZ = Geo::SZ, SY = Geo::SY;
  constexpr int BX = Geo::BX, BY = Geo::BY, BZ = Geo::BZ, SX = Geo::SX;
  constexpr int STEM_TILES = (NS + 15) / 16;  // 68
  constexpr int NT = PHI ? 2 : 1;             // staged input channels of the stem
  constexpr Taps taps{};
  static_assert(taps.ncls <= 4, "step classes");
  __shared__ __align__(16) __bf16 xin[2 * NS * 8];
  __shared__ __align__(16) __bf16 rt[NT][kFRL];
  uint32_t* rtw = reinterpret_cast<uint32_t*>(&rt[0][0]);  // PHI: kFRL (source, potential) words
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  const int gt = g / 2, gh = g % 2;  // hidden K group: tap slot, channel half
  Walk wk;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int j = k * TPS + gt;
      const int tap = d * 9 + (j < 9 ? j : 0);
      bf16x8 v = *reinterpret_cast<const bf16x8*>(wh + (tap * kCout + m) * kCout + 8 * gh);
#pragma unroll
      for (int i = 0; i < 8; ++i) wk.wa[d * KS + k][i] = j < 9 ? v[i] : (__bf16)0.f;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) wk.bsv[i] = bh[4 * g + i];
  // stem: A fragments and the bias of output channels 4g .. 4g+3 with the constant input
  // channels folded in, as k_conv_stem
  const bf16x8 wsa = stem_fragment(ws, 0, g, m);
  const bf16x8 wsp = PHI ? stem_fragment(ws, 1, g, m) : bf16x8{};
  float sbv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) sbv[i] = stem_bias<PHI>(ws, bs, 4 * g + i, a, om, s8);
  const float inv = 1.f / stats[0];
  const float pinv = PHI ? 1.f / stats[3] : 0.f;
  __builtin_amdgcn_s_waitcnt(0);
  // density-tile offsets of this lane's 8 stem taps relative to a staged voxel's own cell
  // (padded taps read tap 26's cell; their weights are zero)
  typedef int TapOffsets __attribute__((ext_vector_type(8)));
  TapOffsets soff;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int t = min(8 * g + j, kTaps - 1);
    soff[j] = (t / 9 - 1) * kFRP + ((t / 3) % 3 - 1) * kFRZ + (t % 3 - 1);
  }
  const int ly = 2 * wave, lz = m;
  {
    const int corner = ly * SZ + lz;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      int d = 0;
#pragma unroll
      for (int t = 1; t < TPS; ++t) d = gt == t ? Taps::diff(taps.rep[cc < taps.ncls ? cc : 0], t) : d;
      wk.cbase[cc] = (gh * NS + corner + d) * 8;
    }
  }
  wk.xin = xin;
  wk.plane = (int64_t)n1 * n2;
  wk.rowz = n2;
  wk.g = g;
  wk.res = nullptr;
  wk.pre = nullptr;
  wk.act = act;
  wk.pch = kCout;
  wk.rho = nullptr;
  wk.rho_add = nullptr;
  // this thread's density-tile cells: offset from the tile's (-2, -2, -2) corner
  typedef int Offsets __attribute__((ext_vector_type(kFRPer)));
  Offsets loc;
#pragma unroll
  for (int u = 0; u < kFRPer; ++u) {
    const int it = min((int)threadIdx.x + u * kThreads, kFRN - 1);
    const int z = it % kFRZ, y = (it / kFRZ) % kFRY, xx = it / (kFRZ * kFRY);
    loc[u] = (xx * n1 + y) * n2 + z;
  }
  const int nbz = n2 / BZ, nby = n1 / BY, nbx = n0 / BX;
  const int ncol = nby * nbz;
  const int nitems = ncol * ((nbx + run - 1) / run);
  const int G = gridDim.x, per = G / 8;
  // the next block to stage: item `item` (column item % ncol, x range item / ncol), x block nbxi
  int item = (blockIdx.x % 8) * per + blockIdx.x / 8;
  int nbxi = (item / ncol) * run;
  constexpr int CP = 2 * SY * SZ;  // voxels of the two carried planes (per channel half)
  typedef float Prefetch __attribute__((ext_vector_type(kFRPer)));
  bool have_cur = false;
  int64_t ecur = 0;  // origin voxel of the block being walked
  for (;;) {
    Prefetch pf, pp;
    const bool have_nxt = item < nitems;
    bool carry = false;
    int64_t enxt = 0;
    if (have_nxt) {
      const int col = item % ncol;
      const int by = col / nbz, bz = col % nbz, bx = nbxi;
      carry = nbxi > (item / ncol) * run;  // block bx - 1 of the same run was staged last
      const int ox = bx * BX, oy = by * BY, oz = bz * BZ;
      enxt = ((int64_t)ox * n1 + oy) * n2 + oz;
      const bool inner = ox >= 2 && ox + BX + 2 <= n0 && oy >= 2 && oy + BY + 2 <= n1 && oz >= 2 && oz + BZ + 2 <= n2;
      if (inner) {
        const int64_t o = ((int64_t)(ox - 2) * n1 + (oy - 2)) * n2 + (oz - 2);
#pragma unroll
        for (int u = 0; u < kFRPer; ++u) pf[u] = rho[o + loc[u]];
        if constexpr (PHI) {
#pragma unroll
          for (int u = 0; u < kFRPer; ++u) pp[u] = phi[o + loc[u]];
        }
      } else {
#pragma unroll
        for (int u = 0; u < kFRPer; ++u) {
          const int it = min((int)threadIdx.x + u * kThreads, kFRN - 1);
          const int z = it % kFRZ, y = (it / kFRZ) % kFRY, xx = it / (kFRZ * kFRY);
          const int gx = wrap(ox + xx - 2, n0), gy = wrap(oy + y - 2, n1), gz = wrap(oz + z - 2, n2);
          const int64_t o = ((int64_t)gx * n1 + gy) * n2 + gz;
          pf[u] = rho[o];
          if constexpr (PHI) pp[u] = phi[o];
        }
      }
    }
    if (have_cur) {
      wk.el = ecur + (int64_t)ly * n2 + lz;
      wk.read_plane(0, 0, 0);
      wk.template walk<0, 0>();
    }
    if (!have_nxt) break;
    __syncthreads();  // the hidden layer's MFMAs of the current block are done with the tile
#pragma unroll
    for (int u = 0; u < kFRPer; ++u) {
      const int i = threadIdx.x + u * kThreads;
      if (i < kFRN) {
        const int l = i + (i / (kFRY * kFRZ)) * (kFRP - kFRY * kFRZ);  // padded planes
        if constexpr (PHI) {  // (source, potential) bf16 pairs: one 4-byte LDS word per cell
          const __bf16 sv = (__bf16)(((pf[u] - 1.f) * c) * inv), pv = (__bf16)(pp[u] * pinv);
          rtw[l] = (uint32_t)__builtin_bit_cast(uint16_t, sv) | ((uint32_t)__builtin_bit_cast(uint16_t, pv) << 16);
        } else {
          rt[0][l] = (__bf16)(((pf[u] - 1.f) * c) * inv);
        }
      }
    }
    if (carry) {
      // staged planes SX - 2, SX - 1 of the block just walked are planes 0, 1 of this one
      for (int i = threadIdx.x; i < 2 * CP; i += kThreads) {
        const int h = i / CP, v = i % CP;
        *reinterpret_cast<uint4*>(xin + (h * NS + v) * 8) =
            *reinterpret_cast<const uint4*>(xin + (h * NS + (SX - 2) * SY * SZ + v) * 8);
      }
    }
    __syncthreads();
    // the stem on the staged voxels of the next block (from plane 2 on when planes 0, 1 were
    // carried), into the hidden layer's tile: a wave takes the tiles t0 + wave, t0 + wave + 4,
    // ... two at a time, so one tile's LDS reads and MFMA overlap the other's GELU (the same
    // arithmetic per voxel as one at a time)
    const int v0 = carry ? CP : 0;
    auto stem_mfma = [&](int t) {
      // density-tile cell of staged voxel vc = (sx, sy, sz), q = vc / SZ = sx SY + sy:
      // (sx + 1) kFRP + (sy + 1) kFRZ + sz + 1 = vc + q (kFRZ - SZ) + sx (kFRP - SY kFRZ) + c0
      const int vc = min(16 * t + m, NS - 1);
      const int q = vc / SZ;
      const int c = vc + q * (kFRZ - SZ) + (q / SY) * (kFRP - SY * kFRZ) + (kFRP + kFRZ + 1);
      if constexpr (PHI) {
        // taps 2i, 2i+1: the low halves of their two words are the source fragment's pair,
        // the high halves the potential's
        typedef uint32_t Words __attribute__((ext_vector_type(4)));
        Words lo, hi;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t w0 = rtw[c + soff[2 * i]], w1 = rtw[c + soff[2 * i + 1]];
          lo[i] = __builtin_amdgcn_perm(w1, w0, 0x05040100u);
          hi[i] = __builtin_amdgcn_perm(w1, w0, 0x07060302u);
        }
        const f32x4 cv = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wsa, __builtin_bit_cast(bf16x8, lo),
                                                               f32x4{sbv[0], sbv[1], sbv[2], sbv[3]}, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(wsp, __builtin_bit_cast(bf16x8, hi), cv, 0, 0, 0);
      } else {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = rt[0][c + soff[j]];
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(wsa, b, f32x4{sbv[0], sbv[1], sbv[2], sbv[3]}, 0, 0, 0);
      }
    };
    auto stem_store = [&](int t, const f32x4& cv) {
      const int v = 16 * t + m;
      bf16x4 o;
#pragma unroll
      for (int i = 0; i < 4; ++i) o[i] = (__bf16)gelu_fast(cv[i]);
      if (v >= v0 && v < NS) *reinterpret_cast<bf16x4*>(xin + ((g / 2) * NS + v) * 8 + 4 * (g % 2)) = o;
    };
    int t = v0 / 16 + wave;
    for (; t + 4 < STEM_TILES; t += 8) {
      const f32x4 c0 = stem_mfma(t);
      const f32x4 c1 = stem_mfma(t + 4);
      stem_store(t, c0);
      stem_store(t + 4, c1);
    }
    if (t < STEM_TILES) stem_store(t, stem_mfma(t));
    __syncthreads();
    have_cur = true;
    ecur = enxt;
    if (++nbxi == min(nbx, (item / ncol + 1) * run)) {
      item += G;
      nbxi = (item / ncol) * run;
    }
  }
}

// Output layer: one output channel from 16 input channels.  The 16 -> 16 implicit GEMM would
// spend 15 of its 16 matrix rows on channels that are thrown away, so this kernel puts the
// taps on the rows instead: for every staged (x, y) column u of the block and each of the nine
// in-plane offsets t = (dx, dy),
//   D_t[u][z] = sum_{dz, ci} w[(dx, dy, dz)][0][ci] x[u][z + dz][ci]
// is one MFMA row (K = 3 dz x 16 channels = 48, two K steps; rows 9..15 and the second step's
// upper half carry zero weights), with the 16 z voxels of the column on the N axis.  The
// output is then out[v] = sum_t D_t[v + (dx, dy)][z]: the D values go to LDS (over the staged
// tile, which is no longer needed) and each output voxel sums its nine, in tap order.  Two
// MFMAs per staged column replace 15 per 16 output voxels.
//
// Blocks are 8 x 8 x 16 output voxels (the mesh must divide into them) staged as 10 x 10 x 18
// with a one-voxel periodic halo: 1800 voxels of 32 bytes, stored as one plane per 8-channel
// half; the 100 staged columns x 9 offsets x 16 z of D take exactly the same 57600 bytes.  The
// layer is bound by HBM, and the halo sets its traffic: 8 x 8 x 16 blocks on 512-thread
// workgroups stage 1.76 voxels per output voxel against 2.11 for 4 x 8 x 16 blocks on 256
// threads, at the same two waves per SIMD (1024^3: 16.2 -> 15.5 ms, fetch 72 -> 53 GiB).
// Epilogues: kPlain writes the pre-activation out + bias[0] (pre, one channel); kOutResidual
// writes rho_add = gain (rho - 1) + amp tanh(out + bias[0]) (the residual stage).
constexpr int kOBX = 8, kOBY = 8, kOBZ = 16;
constexpr int kOThreads = 512;                              // 8 waves
constexpr int kOWaves = kOThreads / 64;
constexpr int kOSX = kOBX + 2, kOSY = kOBY + 2, kOSZ = kOBZ + 2;
constexpr int kOCols = kOSX * kOSY;                         // 100 staged columns
constexpr int kONS = kOCols * kOSZ;                         // 1800 staged voxels
constexpr int kOItems = 2 * kONS;                           // 16-byte vectors staged
constexpr int kOPer = (kOItems + kOThreads - 1) / kOThreads;  // 8 per thread
constexpr int kOColsPerWave = (kOCols + kOWaves - 1) / kOWaves;  // 13 (the last round: waves 0..3)
static_assert(kOCols * 9 * kOBZ == kONS * 8, "D fills the staged tile exactly");
static_assert((kOBX * kOBY * kOBZ) % kOThreads == 0, "output voxels per thread");

template <int EPI>
__global__ void __launch_bounds__(kOThreads) k_conv_out1(const __bf16* __restrict__ x, const __bf16* __restrict__ w,
                                                        const float* __restrict__ bias, float* __restrict__ pre,
                                                        const float* __restrict__ rho, const float* __restrict__ stats,
                                                        float c, float r, float* __restrict__ rho_add, int n0, int n1,
                                                        int n2) {
  __shared__ __align__(16) float lds[kONS * 8];
  __bf16* xin = reinterpret_cast<__bf16*>(lds);
  const int nbz = n2 / kOBZ, nby = n1 / kOBY;
  const int nblk = (n0 / kOBX) * nby * nbz;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  // A fragments: row m = in-plane offset t = 3 (dx + 1) + (dy + 1), K = 8 g + j.
  // step 0: g = 0, 1: dz = -1, channels 8 g + j; g = 2, 3: dz = 0, channels 8 (g - 2) + j
  // step 1: g = 0, 1: dz = +1; g = 2, 3: zero
  bf16x8 wa0, wa1;
  {
    const int ch = 8 * (g % 2);
    const int dz0 = g < 2 ? 0 : 1;  // dz + 1 of step 0
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      wa0[j] = m < 9 ? w[((3 * m + dz0) * kCout) * kCout + ch + j] : (__bf16)0.f;
      wa1[j] = (m < 9 && g < 2) ? w[((3 * m + 2) * kCout) * kCout + ch + j] : (__bf16)0.f;
    }
  }
  const float b0v = bias[0];
  float gain = 0.f, amp = 0.f;
  if constexpr (EPI == kOutResidual) {
    gain = stats[1];
    amp = r * stats[0] / c;
  }
  __builtin_amdgcn_s_waitcnt(0);  // weights in registers before the prefetching loop
  // staged items of this thread: item i is half (i % 2) of staged voxel i / 2 (global order, so
  // consecutive threads read consecutive 16 bytes); element offset from the block's
  // (-1, -1, -1) corner when the halo does not wrap (the caller checks it fits int32)
  typedef int Offsets __attribute__((ext_vector_type(kOPer)));
  Offsets loc;
#pragma unroll
  for (int u = 0; u < kOPer; ++u) {
    const int it = min((int)threadIdx.x + u * kOThreads, kOItems - 1);
    const int h = it % 2, vox = it / 2;
    const int sz = vox % kOSZ, sy = (vox / kOSZ) % kOSY, sx = vox / (kOSZ * kOSY);
    loc[u] = ((sx * n1 + sy) * n2 + sz) * kCout + 8 * h;
  }
  // B fragments of staged column col: lane (g, m) feeds voxel z = m (staged z = m + 1 + dz),
  // channel half g % 2; step 0 reads dz = -1 (g < 2) or 0, step 1 reads dz = +1
  const int h = g % 2;
  const int sz0 = g < 2 ? m : m + 1, sz1 = m + 2;
  constexpr int VPT = (kOBX * kOBY * kOBZ) / kOThreads;  // output voxels per thread
  const int G = gridDim.x, per = G / 8, first = (blockIdx.x % 8) * per + blockIdx.x / 8;
  typedef uint32_t Prefetch __attribute__((ext_vector_type(4 * kOPer)));  // a value, never a private array
  int cur = -1, nxt = first;
  int64_t ecur = 0;  // origin voxel of block cur
  for (;;) {
    // the current block's epilogue loads go first: the vector memory counter waits in order
    float rv[VPT];
    if (cur >= 0) {
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        const int v = threadIdx.x + s * kOThreads;
        const int z = v % kOBZ, y = (v / kOBZ) % kOBY, xx = v / (kOBZ * kOBY);
        const int64_t e = ecur + ((int64_t)xx * n1 + y) * n2 + z;
        rv[s] = EPI == kOutResidual ? rho[e] : 0.f;
      }
    }
    Prefetch pf;
    if (nxt < nblk) {
      int bx, by, bz;
      block_of(nxt, n0 / kOBX, nby, nbz, bx, by, bz);
      const int ox = bx * kOBX, oy = by * kOBY, oz = bz * kOBZ;
      const bool inner = ox > 0 && ox + kOBX < n0 && oy > 0 && oy + kOBY < n1 && oz > 0 && oz + kOBZ < n2;
      if (inner) {
        const __bf16* base = x + (((int64_t)(ox - 1) * n1 + (oy - 1)) * n2 + (oz - 1)) * kCout;
#pragma unroll
        for (int u = 0; u < kOPer; ++u) {  // unconditional (clamped) loads, all in flight together
          const uint4 t = *reinterpret_cast<const uint4*>(base + loc[u]);
          pf[4 * u] = t.x;
          pf[4 * u + 1] = t.y;
          pf[4 * u + 2] = t.z;
          pf[4 * u + 3] = t.w;
        }
      } else {
#pragma unroll
        for (int u = 0; u < kOPer; ++u) {
          const int it = min((int)threadIdx.x + u * kOThreads, kOItems - 1);
          const int hh = it % 2, vox = it / 2;
          const int sz = vox % kOSZ, sy = (vox / kOSZ) % kOSY, sx = vox / (kOSZ * kOSY);
          const int gx = wrap(ox + sx - 1, n0), gy = wrap(oy + sy - 1, n1), gz = wrap(oz + sz - 1, n2);
          const uint4 t = *reinterpret_cast<const uint4*>(x + (((int64_t)gx * n1 + gy) * n2 + gz) * kCout + 8 * hh);
          pf[4 * u] = t.x;
          pf[4 * u + 1] = t.y;
          pf[4 * u + 2] = t.z;
          pf[4 * u + 3] = t.w;
        }
      }
    }
    if (cur >= 0) {
      f32x4 acc[kOColsPerWave];
#pragma unroll
      for (int k = 0; k < kOColsPerWave; ++k) {
        const int col = wave + kOWaves * k;
        if (col >= kOCols) break;  // wave-uniform: the last round has fewer columns than waves
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(xin + (h * kONS + col * kOSZ + sz0) * 8);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(xin + (h * kONS + col * kOSZ + sz1) * 8);
        f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa0, b0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, b1, d, 0, 0, 0);
      }
      __syncthreads();  // every wave is done reading the staged tile
      // D over the tile: lds[(col * 9 + t) * 16 + z]; lane (g, m) holds t = 4 g + i of voxel z = m
#pragma unroll
      for (int k = 0; k < kOColsPerWave; ++k) {
        const int col = wave + kOWaves * k;
        if (col >= kOCols) break;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (4 * g + i < 9) lds[(col * 9 + 4 * g + i) * kOBZ + m] = acc[k][i];
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < VPT; ++s) {
        const int v = threadIdx.x + s * kOThreads;
        const int z = v % kOBZ, y = (v / kOBZ) % kOBY, xx = v / (kOBZ * kOBY);
        float o = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const int col = (xx + t / 3) * kOSY + (y + t % 3);  // staged column (x + 1 + dx, y + 1 + dy)
          o += lds[(col * 9 + t) * kOBZ + z];
        }
        const float zv = o + b0v;
        const int64_t e = ecur + ((int64_t)xx * n1 + y) * n2 + z;
        if constexpr (EPI == kOutResidual)
          rho_add[e] = gain * (rv[s] - 1.f) + amp * tanhf(zv);
        else
          pre[e] = zv;
      }
    }
    if (nxt >= nblk) break;
    __syncthreads();  // the reduction is done with D
#pragma unroll
    for (int u = 0; u < kOPer; ++u) {
      const int i = threadIdx.x + u * kOThreads;
      if (i < kOItems)
        *reinterpret_cast<uint4*>(xin + ((i % 2) * kONS + i / 2) * 8) =
            make_uint4(pf[4 * u], pf[4 * u + 1], pf[4 * u + 2], pf[4 * u + 3]);
    }
    __syncthreads();
    cur = nxt;
    {
      int bx, by, bz;
      block_of(cur, n0 / kOBX, nby, nbz, bx, by, bz);
      ecur = ((int64_t)bx * kOBX * n1 + by * kOBY) * n2 + bz * kOBZ;
    }
    nxt += G;
  }
}


// Weight and bias gradients: dw[tap][co][ci] = sum_v dz[v][co] x[v + off(tap)][ci],
// db[co] = sum_v dz[v][co].  A persistent grid walks the voxel blocks; each workgroup keeps
// its partial sums in registers (MFMA accumulators, K = voxels) and writes them to
// partial[blockIdx.x]; k_reduce_partials adds the partials in a fixed order.
template <int CIN>
__global__ void __launch_bounds__(kThreads) k_conv_bwd_weight(const float* __restrict__ dz,
                                                              const __bf16* __restrict__ x, int n0, int n1, int n2,
                                                              float* __restrict__ partial) {
  // channel-major staging so the 8 voxels of an MFMA K group are contiguous along z
  __shared__ __align__(16) __bf16 xs[CIN * kH * kH * kH];
  __shared__ __align__(16) __bf16 gs[kCout * kVox];
  __shared__ float bsum[kThreads / 64][kCout];
  constexpr int NW = kTaps * kCout * CIN;  // weight gradients per workgroup
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  // wave w owns taps w, w+4, ...: up to 7 taps, one 16x16 (co x ci) block each (CIN <= 16)
  static_assert(CIN == 8 || CIN == 16, "CIN must be 8 or 16");
  f32x4 acc[7];
#pragma unroll
  for (int t = 0; t < 7; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f;
  const int nbz = n2 / kB, nby = n1 / kB, nblk = (n0 / kB) * nby * nbz;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
    const int ox = bx * kB, oy = by * kB, oz = bz * kB;
    __syncthreads();
    for (int it = threadIdx.x; it < kH * kH * kH * CIN; it += kThreads) {
      const int c = it % CIN, vox = it / CIN;
      const int z = vox % kH, y = (vox / kH) % kH, xx = vox / (kH * kH);
      const int gx = wrap(ox + xx - 1, n0), gy = wrap(oy + y - 1, n1), gz = wrap(oz + z - 1, n2);
      xs[c * kH * kH * kH + vox] = x[(((int64_t)gx * n1 + gy) * n2 + gz) * CIN + c];
    }
    for (int it = threadIdx.x; it < kVox * kCout; it += kThreads) {
      const int c = it % kCout, vox = it / kCout;
      const int z = vox % kB, y = (vox / kB) % kB, xx = vox / (kB * kB);
      const float v = dz[(((int64_t)(ox + xx) * n1 + (oy + y)) * n2 + (oz + z)) * kCout + c];
      gs[c * kVox + vox] = (__bf16)v;
      bacc += (threadIdx.x % kCout == c) ? v : 0.f;  // c == threadIdx.x % 16 since 256 % 16 == 0
    }
    __syncthreads();
    // K step k covers voxels 32k .. 32k+31: x plane k/2, y rows 4*(k%2) .. +3, all 8 z
#pragma unroll 1
    for (int k = 0; k < kVox / 32; ++k) {
      // lane group g covers voxels 32k + 8g .. +7: x plane k/2, y row 4*(k%2) + g, z 0..7
      const int vx = k / 2, gy = 4 * (k % 2) + g;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(gs + m * kVox + (vx * kB + gy) * kB);
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const int tap = wave + 4 * t;
        if (tap >= kTaps) break;
        const int dx_ = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz_ = tap % 3 - 1;
        // x[v + off]: staged voxel (vx+1+dx, gy+1+dy, z+1+dz), z = 0..7 contiguous
        const int base = ((vx + 1 + dx_) * kH + (gy + 1 + dy)) * kH + (1 + dz_);
        const int ci = m < CIN ? m : 0;  // CIN = 8: lanes 8..15 feed zero columns
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = m < CIN ? xs[ci * kH * kH * kH + base + j] : (__bf16)0.f;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // bias partials: lanes with the same channel (threadIdx.x % 16) are summed per wave
  for (int off = 16; off < 64; off <<= 1) bacc += __shfl_xor(bacc, off);
  if (lane < 16) bsum[wave][lane] = bacc;
  __syncthreads();
  float* out = partial + (int64_t)blockIdx.x * (NW + kCout);
  // acc[t][cb] lane holds rows co = 4g+i, column ci = cb*16 + m
#pragma unroll
  for (int t = 0; t < 7; ++t) {
    const int tap = wave + 4 * t;
    if (tap >= kTaps) break;
    if (m < CIN) {
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(tap * kCout + 4 * g + i) * CIN + m] = acc[t][i];
    }
  }
  if (threadIdx.x < kCout) {
    float b = 0.f;
    for (int w2 = 0; w2 < kThreads / 64; ++w2) b += bsum[w2][threadIdx.x];
    out[NW + threadIdx.x] = b;
  }
}

// sum over partial rows in a fixed order: out[i] = sum_r partial[r][i]
__global__ void k_reduce_partials(const float* __restrict__ partial, int rows, int width, float* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < width; i += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += partial[(int64_t)r * width + i];
    out[i] = s;
  }
}

}  // namespace
}  // namespace conv
}  // namespace pm

using namespace pm;

namespace {

int check_shape(const int32_t* shape) {
  PM_RETURN_IF(shape == nullptr, PM_ERROR_INVALID_ARGUMENT, "shape is NULL");
  for (int a = 0; a < 3; ++a)
    PM_RETURN_IF(shape[a] < conv::kB || shape[a] % conv::kB != 0, PM_ERROR_UNSUPPORTED,
                 "conv3d mesh axes must be positive multiples of 8");
  return PM_SUCCESS;
}

int check_ptr16(const void* p, const char* what) {
  PM_RETURN_IF(p == nullptr, PM_ERROR_INVALID_ARGUMENT, std::string(what) + " is NULL");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(p) % 16 != 0, PM_ERROR_INVALID_ARGUMENT,
               std::string(what) + " must be 16-byte aligned");
  return PM_SUCCESS;
}

// the forward conv's 4 x 8 x 16 blocks (conflict-free fragment reads) when the mesh divides
// into them; 8^3 blocks otherwise
bool wide_blocks(const int32_t* shape) { return shape[0] % 4 == 0 && shape[2] % 16 == 0; }

// blocks of 512 output voxels (8^3, or 4 x 8 x 16)
unsigned nblocks(const int32_t* shape) {
  return (unsigned)((int64_t)(shape[0] / conv::kB) * (shape[1] / conv::kB) * (shape[2] / conv::kB));
}


// Workgroups of a persistent conv grid: what fits on the device at once (the kernel's
// occupancy times the CU count), at most the number of blocks, rounded up to a multiple of 8
// (one share per XCD).
int persistent_grid(const void* kern, int64_t nblk, unsigned* grid, int threads = conv::kThreads) {
  int dev = 0, cus = 0, per_cu = 0;
  PM_HIP_TRY(hipGetDevice(&dev));
  PM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
  int64_t g = (int64_t)std::max(cus, 1) * std::max(per_cu, 1);
  g = std::min<int64_t>(g, nblk);
  *grid = (unsigned)(ceil_div(g, 8) * 8);
  return PM_SUCCESS;
}

// x runs of a persistent grid over (y, z) columns of blocks (k_conv_stem_hidden):
// whole x lines when there are enough columns for every workgroup the grid holds, halved until
// there are work items for all of them; the grid is cut to the work items there are.
int64_t x_run(int64_t nbx, int64_t ncol, unsigned* grid) {
  int64_t run = nbx;
  while (run > 1 && ncol * ceil_div(nbx, run) < (int64_t)*grid) run = ceil_div(run, 2);
  *grid = (unsigned)std::min<int64_t>(*grid, ceil_div(ncol * ceil_div(nbx, run), 8) * 8);
  return run;
}

// The one-output-channel layer with the taps on the matrix rows (k_conv_out1): a persistent
// grid over the 8 x 8 x 16 blocks (the next block's tile prefetched while one is computed)
int launch_out1(const __bf16* x, const __bf16* w, const float* bias, float* pre, const float* rho,
                const float* stats, float c, float r, float* rho_add, const int32_t* shape, hipStream_t s) {
  const int64_t nblk = (int64_t)shape[0] / conv::kOBX * (shape[1] / conv::kOBY) * (shape[2] / conv::kOBZ);
  PM_RETURN_IF(nblk >= (int64_t(1) << 31) ||
                   (int64_t)conv::kOSX * shape[1] * shape[2] * conv::kCout >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  unsigned grid = 0;
  if (rho_add) {
    auto kern = conv::k_conv_out1<conv::kOutResidual>;
    PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblk, &grid, conv::kOThreads));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kOThreads), 0, s, x, w, bias, pre, rho, stats, c, r, rho_add,
                       shape[0], shape[1], shape[2]);
  } else {
    auto kern = conv::k_conv_out1<conv::kPlain>;
    PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblk, &grid, conv::kOThreads));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kOThreads), 0, s, x, w, bias, pre, rho, stats, c, r, rho_add,
                       shape[0], shape[1], shape[2]);
  }
  PM_LAUNCHED("k_conv_out1", s);
  return PM_SUCCESS;
}

bool out1_blocks(const int32_t* shape) { return shape[0] % conv::kOBX == 0 && shape[2] % conv::kOBZ == 0; }

constexpr int kWGrid = 1024;  // persistent workgroups of the weight-gradient pass

}  // namespace

extern "C" {

int pm_conv3d_fwd(const void* x, int32_t cin, const void* w, const float* bias, const void* res, int32_t gelu,
                  float* pre, int32_t pre_channels, void* act, const int32_t shape[3], void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(pre_channels != 1 && pre_channels != 16, PM_ERROR_INVALID_ARGUMENT, "pre_channels must be 1 or 16");
  PM_RETURN_IF(cin != 8 && cin != 16, PM_ERROR_UNSUPPORTED, "conv3d input channels must be 8 or 16");
  PM_TRY(check_ptr16(x, "x"));
  PM_RETURN_IF(w == nullptr || bias == nullptr, PM_ERROR_INVALID_ARGUMENT, "weights / bias NULL");
  PM_RETURN_IF(pre == nullptr && act == nullptr, PM_ERROR_INVALID_ARGUMENT, "no output requested");
  hipStream_t s = as_stream(stream);
  const __bf16* xb = static_cast<const __bf16*>(x);
  const __bf16* wb = static_cast<const __bf16*>(w);
  const __bf16* rb = static_cast<const __bf16*>(res);
  __bf16* ab = static_cast<__bf16*>(act);
  if (res) PM_TRY(check_ptr16(res, "res"));
  PM_RETURN_IF(pre && reinterpret_cast<uintptr_t>(pre) % 16 != 0, PM_ERROR_INVALID_ARGUMENT,
               "pre must be 16-byte aligned");
  PM_RETURN_IF(act && reinterpret_cast<uintptr_t>(act) % 16 != 0, PM_ERROR_INVALID_ARGUMENT,
               "act must be 16-byte aligned");
  PM_RETURN_IF((int64_t)conv::kH * shape[1] * shape[2] * cin >= (int64_t(1) << 31) ||
                   (int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  if (cin == 16 && pre_channels == 1 && act == nullptr && !gelu && res == nullptr && pre && out1_blocks(shape))
    return launch_out1(xb, wb, bias, pre, nullptr, nullptr, 0.f, 0.f, nullptr, shape, s);
  const int64_t nblk = nblocks(shape);
  const int key = (cin == 16 ? 4 : 0) + (gelu ? 2 : 0) + (res ? 1 : 0) + (wide_blocks(shape) ? 8 : 0);
  void (*kern)(const __bf16*, const __bf16*, const float*, const __bf16*, float*, __bf16*, int, int, int, int,
               const float*, const float*, float, float, float*, const float*, const float*) = nullptr;
  switch (key) {
#define PM_CONV_CASE(K, C, G, R, Z) \
  case K: kern = conv::k_conv_fwd<C, G, R, conv::kPlain, Z>; break;
    PM_CONV_CASE(0, 8, false, false, 8)
    PM_CONV_CASE(1, 8, false, true, 8)
    PM_CONV_CASE(2, 8, true, false, 8)
    PM_CONV_CASE(3, 8, true, true, 8)
    PM_CONV_CASE(4, 16, false, false, 8)
    PM_CONV_CASE(5, 16, false, true, 8)
    PM_CONV_CASE(6, 16, true, false, 8)
    PM_CONV_CASE(7, 16, true, true, 8)
    PM_CONV_CASE(8, 8, false, false, 16)
    PM_CONV_CASE(9, 8, false, true, 16)
    PM_CONV_CASE(10, 8, true, false, 16)
    PM_CONV_CASE(11, 8, true, true, 16)
    PM_CONV_CASE(12, 16, false, false, 16)
    PM_CONV_CASE(13, 16, false, true, 16)
    PM_CONV_CASE(14, 16, true, false, 16)
    default: kern = conv::k_conv_fwd<16, true, true, conv::kPlain, 16>; break;
#undef PM_CONV_CASE
  }
  unsigned grid = 0;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblk, &grid));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, xb, wb, bias, rb, pre, ab, shape[0], shape[1],
                     shape[2], (int)pre_channels, nullptr, nullptr, 0.f, 0.f, nullptr, nullptr, nullptr);
  PM_LAUNCHED("k_conv_fwd", s);
  return PM_SUCCESS;
}

int pm_conv3d_stem(const float* rho, const float* phi, const int32_t shape[3], float omega_m, float a, float sigma8,
                   const float* stats, const void* w, const float* bias, void* act, void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(rho == nullptr || stats == nullptr || w == nullptr || bias == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "NULL argument");
  PM_TRY(check_ptr16(act, "act"));
  PM_RETURN_IF((int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31) ||
                   (int64_t)conv::kH * shape[1] * shape[2] >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  hipStream_t s = as_stream(stream);
  unsigned grid = 0;
  auto kern = phi ? conv::k_conv_stem<conv::kStemPotential> : conv::k_conv_stem<conv::kStemDensity>;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblocks(shape), &grid));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, rho, phi, 1.5f * omega_m, stats, a, omega_m,
                     sigma8, static_cast<const __bf16*>(w), bias, static_cast<__bf16*>(act), shape[0], shape[1],
                     shape[2], nullptr, nullptr);
  PM_LAUNCHED("k_conv_stem", s);
  return PM_SUCCESS;
}

int pm_conv3d_stem_hidden(const float* rho, const float* phi, const int32_t shape[3], float omega_m, float a,
                          float sigma8,
                          const float* stats, const void* w_stem, const float* b_stem, const void* w_hidden,
                          const float* b_hidden, void* act, void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(!wide_blocks(shape), PM_ERROR_UNSUPPORTED,
               "fused stem + hidden conv: the mesh's x extent must be a multiple of 4 and z of 16");
  PM_RETURN_IF(rho == nullptr || stats == nullptr || w_stem == nullptr || b_stem == nullptr || b_hidden == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_TRY(check_ptr16(w_hidden, "w_hidden"));
  PM_TRY(check_ptr16(act, "act"));
  PM_RETURN_IF((int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31) ||
                   (int64_t)conv::kFRX * shape[1] * shape[2] >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  hipStream_t s = as_stream(stream);
  unsigned grid = 0;
  auto kern = phi ? conv::k_conv_stem_hidden<true> : conv::k_conv_stem_hidden<false>;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblocks(shape), &grid));
  const int64_t run = x_run(shape[0] / 4, (int64_t)(shape[1] / 8) * (shape[2] / 16), &grid);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, rho, phi, 1.5f * omega_m, stats, a, omega_m,
                     sigma8, static_cast<const __bf16*>(w_stem), b_stem, static_cast<const __bf16*>(w_hidden),
                     b_hidden, static_cast<__bf16*>(act), shape[0], shape[1], shape[2], (int)run);
  PM_LAUNCHED("k_conv_stem_hidden", s);
  return PM_SUCCESS;
}

int pm_conv3d_out_residual(const void* x, const void* w, const float* bias, const float* rho,
                           const int32_t shape[3], float omega_m, float max_residual, const float* stats,
                           float* rho_add, void* stream) {
  PM_TRY(check_shape(shape));
  PM_TRY(check_ptr16(x, "x"));
  PM_RETURN_IF(w == nullptr || bias == nullptr || rho == nullptr || stats == nullptr || rho_add == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF((int64_t)conv::kH * shape[1] * shape[2] * 16 >= (int64_t(1) << 31) ||
                   (int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  hipStream_t s = as_stream(stream);
  if (out1_blocks(shape))
    return launch_out1(static_cast<const __bf16*>(x), static_cast<const __bf16*>(w), bias, nullptr, rho, stats,
                       1.5f * omega_m, max_residual, rho_add, shape, s);
  auto kern = wide_blocks(shape) ? conv::k_conv_fwd<16, false, false, conv::kOutResidual, 16>
                                 : conv::k_conv_fwd<16, false, false, conv::kOutResidual, 8>;
  unsigned grid = 0;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblocks(shape), &grid));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, static_cast<const __bf16*>(x),
                     static_cast<const __bf16*>(w), bias, nullptr, nullptr, nullptr, shape[0], shape[1], shape[2], 1,
                     rho, stats, 1.5f * omega_m, max_residual, rho_add, nullptr, nullptr);
  PM_LAUNCHED("k_conv_out_residual", s);
  return PM_SUCCESS;
}

int pm_conv3d_bwd_data(const void* dz, const void* w, const void* add, const float* gact, void* dx,
                       int32_t dx_channels, int32_t bf16_io, const int32_t shape[3], void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(dz == nullptr || w == nullptr || dx == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(dx_channels != 16 && dx_channels != 4, PM_ERROR_INVALID_ARGUMENT, "dx_channels must be 16 or 4");
  PM_RETURN_IF(bf16_io != 0 && bf16_io != 1 && bf16_io != 3, PM_ERROR_INVALID_ARGUMENT,
               "bf16_io must be 0 (float32 dz and dx), 1 (bf16 dz, float32 dx) or 3 (bf16 dz and dx)");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(dz) % 16 != 0 || reinterpret_cast<uintptr_t>(dx) % 16 != 0 ||
                   reinterpret_cast<uintptr_t>(add) % 16 != 0 || reinterpret_cast<uintptr_t>(gact) % 16 != 0,
               PM_ERROR_INVALID_ARGUMENT, "conv3d data gradient: meshes must be 16-byte aligned");
  PM_RETURN_IF((int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31) ||
                   (int64_t)(conv::kB + 2) * shape[1] * shape[2] * conv::kCout >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  PM_RETURN_IF(dz == dx, PM_ERROR_INVALID_ARGUMENT, "conv3d data gradient: dx must not alias dz");
  hipStream_t s = as_stream(stream);
  const auto wb = static_cast<const __bf16*>(w);
  const auto xb = static_cast<const __bf16*>(dz);  // read as float32 by the kBwdData walk
  const bool wide = wide_blocks(shape);
  decltype(&conv::k_conv_fwd<16, false, false, conv::kBwdData, 8>) kern;
  if (bf16_io == 0)
    kern = wide ? conv::k_conv_fwd<16, false, false, conv::kBwdData, 16> : conv::k_conv_fwd<16, false, false, conv::kBwdData, 8>;
  else if (bf16_io == 1)
    kern = wide ? conv::k_conv_fwd<16, false, false, conv::kBwdData16F, 16>
                : conv::k_conv_fwd<16, false, false, conv::kBwdData16F, 8>;
  else
    kern = wide ? conv::k_conv_fwd<16, false, false, conv::kBwdData16, 16>
                : conv::k_conv_fwd<16, false, false, conv::kBwdData16, 8>;
  float* dxf = bf16_io == 3 ? nullptr : static_cast<float*>(dx);
  __bf16* dxb = bf16_io == 3 ? static_cast<__bf16*>(dx) : nullptr;
  unsigned grid = 0;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblocks(shape), &grid));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, xb, wb, nullptr, nullptr, dxf, dxb, shape[0],
                     shape[1], shape[2], (int)dx_channels, nullptr, nullptr, 0.f, 0.f, nullptr,
                     static_cast<const float*>(add), gact);
  PM_LAUNCHED("k_conv_bwd_data", s);
  return PM_SUCCESS;
}

int pm_conv3d_bwd_data_out1(const float* dz0, const void* w, const float* gact, void* dx, int32_t dx_bf16,
                            const int32_t shape[3], void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(dz0 == nullptr || w == nullptr || dx == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(dx) % 16 != 0 || reinterpret_cast<uintptr_t>(gact) % 16 != 0,
               PM_ERROR_INVALID_ARGUMENT, "dx and gact must be 16-byte aligned");
  PM_RETURN_IF(dz0 == dx, PM_ERROR_INVALID_ARGUMENT, "dx must not alias dz0");
  PM_RETURN_IF((int64_t)shape[0] * shape[1] * shape[2] / (conv::kB * conv::kB * conv::kB) >= (int64_t(1) << 31) ||
                   (int64_t)conv::kH * shape[1] * shape[2] >= (int64_t(1) << 31),
               PM_ERROR_UNSUPPORTED, "conv3d: mesh too large for 32-bit tile offsets");
  hipStream_t s = as_stream(stream);
  auto kern = dx_bf16 ? conv::k_conv_stem<conv::kStemBwdOut1Bf16> : conv::k_conv_stem<conv::kStemBwdOut1>;
  unsigned grid = 0;
  PM_TRY(persistent_grid(reinterpret_cast<const void*>(kern), nblocks(shape), &grid));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(conv::kThreads), 0, s, dz0, nullptr, 0.f, nullptr, 0.f, 0.f, 0.f,
                     static_cast<const __bf16*>(w), nullptr, dx_bf16 ? static_cast<__bf16*>(dx) : nullptr, shape[0],
                     shape[1], shape[2], dx_bf16 ? nullptr : static_cast<float*>(dx), gact);
  PM_LAUNCHED("k_conv_bwd_out1", s);
  return PM_SUCCESS;
}

int pm_conv3d_bwd_weight_workspace(const int32_t shape[3], int32_t cin, size_t* bytes) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(cin != 8 && cin != 16, PM_ERROR_UNSUPPORTED, "conv3d input channels must be 8 or 16");
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  const int64_t grid = std::min<int64_t>(nblocks(shape), kWGrid);
  *bytes = sizeof(float) * grid * (conv::kTaps * conv::kCout * cin + conv::kCout);
  return PM_SUCCESS;
}

int pm_conv3d_bwd_weight(const float* dz, const void* x, int32_t cin, const int32_t shape[3], float* dw_db,
                         void* workspace, size_t workspace_bytes, void* stream) {
  size_t need = 0;
  PM_TRY(pm_conv3d_bwd_weight_workspace(shape, cin, &need));
  PM_RETURN_IF(dz == nullptr || x == nullptr || dw_db == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < need, PM_ERROR_WORKSPACE, "conv3d workspace too small");
  hipStream_t s = as_stream(stream);
  const unsigned grid = (unsigned)std::min<int64_t>(nblocks(shape), kWGrid);
  float* partial = static_cast<float*>(workspace);
  const __bf16* xb = static_cast<const __bf16*>(x);
  if (cin == 8)
    hipLaunchKernelGGL(conv::k_conv_bwd_weight<8>, dim3(grid), dim3(conv::kThreads), 0, s, dz, xb, shape[0],
                       shape[1], shape[2], partial);
  else
    hipLaunchKernelGGL(conv::k_conv_bwd_weight<16>, dim3(grid), dim3(conv::kThreads), 0, s, dz, xb, shape[0],
                       shape[1], shape[2], partial);
  PM_LAUNCHED("k_conv_bwd_weight", s);
  const int width = conv::kTaps * conv::kCout * cin + conv::kCout;
  hipLaunchKernelGGL(conv::k_reduce_partials, dim3((unsigned)ceil_div(width, 256)), dim3(256), 0, s, partial,
                     (int)grid, width, dw_db);
  PM_LAUNCHED("k_reduce_partials", s);
  return PM_SUCCESS;
}

}  // extern "C"
