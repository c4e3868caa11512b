// The single-device PM force step and its reverse (force VJP).
//
// Forward (nbody/gravity.py:786-814 on the streamed single-device branch :597-606, 506-543):
//   1. cell sort (pm_sort.hip) and a gather of the particle payload into key order;
//   2. CIC scatter of v_p = valid * N_m/N_p from sorted particles (spatially coherent atomics);
//   3. z pass r2c with S = (rho - 1) * (1.5 Omega_m) fused into the load;
//   4. y pass forward;
//   5. x pass forward, mask x (-1/k^2), then two inverse x transforms in registers:
//        A <- IFFT_x(phi),  B <- IFFT_x(-i k_x phi)
//   6. y pass inverse on A with two outputs: A <- IFFT_y(A), C <- IFFT_y(-i k_y A);
//      y pass inverse on B;
//   7. z pass c2r of B -> F_x, C -> F_y, (-i k_z) A -> F_z  (1/N normalisation fused);
//   8. CIC gather of (F_x, F_y, F_z) from sorted particles, written back through the
//      permutation in caller order.
// The derivative operators commute with the transforms along the other axes, so the three
// inverse transforms of the reference (one per component, :535-543) share their x and y
// passes: 4R + 14C bytes of mesh traffic instead of 4R + 20C.
#include <algorithm>

#include "pm_common.h"
#include "pm_fft_launch.h"
#include "pm_sort.h"

namespace pm {
namespace grav {
namespace {

constexpr int kBlock = 256;

__global__ void k_fill(float* __restrict__ p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// dst = src for a mesh handed back to the caller: 16-byte accesses, four in flight per
// thread (a device-to-device hipMemcpyAsync of the 4 GiB density at 1024^3 ran at 2.1 TB/s)
__global__ void __launch_bounds__(kBlock) k_copy_mesh(const float* __restrict__ src, float* __restrict__ dst,
                                                      int64_t n) {
  const int64_t n4 = n / 4;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const float4* s4 = reinterpret_cast<const float4*>(src);
  float4* d4 = reinterpret_cast<float4*>(dst);
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 a = s4[i], b = s4[i + stride], c = s4[i + 2 * stride], d = s4[i + 3 * stride];
    d4[i] = a;
    d4[i + stride] = b;
    d4[i + 2 * stride] = c;
    d4[i + 3 * stride] = d;
  }
  for (; i < n4; i += stride) d4[i] = s4[i];
  for (int64_t j = 4 * n4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) dst[j] = src[j];
}

__global__ void __launch_bounds__(kBlock) k_scatter_sorted(const int16_t* __restrict__ pmid,
                                                           const float* __restrict__ disp,
                                                           const uint8_t* __restrict__ valid, int64_t n,
                                                           float val, Geom g, float* __restrict__ mesh) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !valid[j]) return;
  AxisCic ax[3];
  particle_cic(g, pmid, disp, j, ax);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int64_t flat;
    corner_index(g, ax, c, flat);
    atomic_add_f32(mesh + flat, val * corner_weight(ax, c));
  }
}

__global__ void __launch_bounds__(kBlock) k_gather3_sorted(const int16_t* __restrict__ pmid,
                                                           const float* __restrict__ disp,
                                                           const uint8_t* __restrict__ valid,
                                                           const int32_t* __restrict__ perm, int64_t n,
                                                           Geom g, const float* __restrict__ fx,
                                                           const float* __restrict__ fy,
                                                           const float* __restrict__ fz,
                                                           float* __restrict__ acc) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float r[3] = {0.f, 0.f, 0.f};
  if (valid[j]) {
    AxisCic ax[3];
    particle_cic(g, pmid, disp, j, ax);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      corner_index(g, ax, c, flat);
      float w = corner_weight(ax, c);
      r[0] = r[0] + fx[flat] * w;
      r[1] = r[1] + fy[flat] * w;
      r[2] = r[2] + fz[flat] * w;
    }
  }
  int64_t i = perm[j];
  acc[i * 3 + 0] = r[0];
  acc[i * 3 + 1] = r[1];
  acc[i * 3 + 2] = r[2];
}

// ----------------------------------------------------------------------------------------
// Brick-tiled CIC on sorted particles.  The mesh is tiled by bricks of (BX,BY,BZ) cells; the
// cell sort groups particles by the brick of their base (corner-0) cell, so every corner a
// particle touches lies in its brick's tile extended by one cell on the upper side.  One
// workgroup walks a column of bricks (fixed x and y brick coordinates) along z.
//
// Scatter (owner computes): the column's bricks are accumulated one after the other into an
// LDS tile of (BX+1)(BY+1)(BZ+1) int64 fixed-point cells with LDS atomics.  The tile's upper z
// plane belongs to the next brick of the column and is carried to it inside LDS (the first
// brick's lower plane waits for the last brick's carry), so the z halo never leaves the CU.
// Owned cells are stored once with plain stores (no zero fill, no global atomics); the upper
// x and y planes go to z-contiguous plane buffers (XS, YS) that k_fold_xy adds into the lower
// x/y planes of the neighbouring columns (pull, fixed order: deterministic).
// Gather: the column is walked downward in z.  Each force component's tile rows z in [0, BZ)
// are read as aligned rows; the z = BZ plane is the lower plane of the brick above, kept in
// LDS from the previous step, so no row costs a second cache line.
constexpr int kTileMaxFloats = 19100;  // gather f32 tile: (32+1)*(16+1)*(32+2) = 19074 fits
constexpr int kTileMaxWords = 9900;    // scatter int64 tile: (16+1)*(16+1)*(32+2) = 9826 (78.6 KB: 2 WG/CU)
constexpr int kPlaneMax = 1100;        // (b0+1)*(b1+1) cells of one z plane of a tile
// Scatter accumulation is exact 64-bit fixed point (2^-32 units): gfx950's LDS integer atomics run
// an order of magnitude faster than its LDS float atomics, and integer sums are order independent,
// so the density is deterministic; each cell is rounded to float32 once when stored.
constexpr float kFix = 4294967296.0f;         // 2^32
constexpr float kUnfix = 2.3283064365386963e-10f;  // 2^-32

struct Bricks {
  int n[3];   // mesh shape
  int b[3];   // brick shape
  int nb[3];  // bricks per axis
};

inline Bricks make_bricks(const int32_t* mesh, const int32_t* brick) {
  Bricks k;
  for (int a = 0; a < 3; ++a) {
    k.n[a] = mesh[a];
    k.b[a] = brick[a];
    k.nb[a] = mesh[a] / brick[a];
  }
  return k;
}

inline bool tiled_ok(const int32_t* mesh, const int32_t* brick) {
  for (int a = 0; a < 3; ++a)
    if (brick[a] < 1 || mesh[a] % brick[a] != 0) return false;
  // LDS tile rows have a z pitch of b2 + 2 words
  const int64_t plane = (int64_t)(brick[0] + 1) * (brick[1] + 1);
  const int64_t words = plane * (brick[2] + 2);
  return words <= kTileMaxFloats && words <= kTileMaxWords && plane <= kPlaneMax;
}

inline bool is_brick(const Bricks& k, int x, int y, int z) { return k.b[0] == x && k.b[1] == y && k.b[2] == z; }

// Upper-plane buffers of the tiled scatter, z-contiguous:
//   XS[(col * (b1+1) + ly) * nz + z]  the x == b0 plane of column col (ly in [0, b1])
//   YS[(col * b0 + lx) * nz + z]      the y == b1 plane of column col (lx in [0, b0))
inline int64_t xs_floats(const Bricks& k) { return (int64_t)k.nb[0] * k.nb[1] * (k.b[1] + 1) * k.n[2]; }
inline int64_t ys_floats(const Bricks& k) { return (int64_t)k.nb[0] * k.nb[1] * k.b[0] * k.n[2]; }

// Workgroups are dispatched round-robin over the 8 XCDs; logical column q runs on workgroup
// (q % g8) * 8 + q / g8 so neighbouring columns (which share halo rows) share an L2.
__device__ __forceinline__ int column_of(int L, int G) {
  if (G % 8 != 0) return L;
  return (L % 8) * (G / 8) + L / 8;
}

// The column walks that read a tile plus its one-cell x/y halo (gather, displacement cotangent):
// XCD x (workgroups L with L % 8 == x) takes the columns cx in [x R, (x + 1) R), R = nb0 / 8, and
// visits them in patches of PX x 8 columns, patch after patch, so the workgroups an XCD runs at
// the same time (about 64) form a compact patch of the (cx, cy) plane and read the halo rows
// they share into one L2 at nearly the same time, instead of a 64-column strip whose x-halo
// rows are read again, long after, by the strip beside it.  Other shapes keep column_of.
__device__ __forceinline__ int patch_column_of(int L, int G, int nb0, int nb1) {
  const int R = nb0 / 8;
  const int PX = R < 8 ? R : 8;
  if (G % 8 != 0 || G != nb0 * nb1 || nb0 % 8 != 0 || nb1 % 8 != 0 || PX == 0 || R % PX != 0) return column_of(L, G);
  const int x = L % 8, j = L / 8;
  const int per = PX * 8;
  const int p = j / per, w = j - p * per;
  const int npy = nb1 / 8;
  const int pxi = p / npy, pyi = p - pxi * npy;
  const int cx = x * R + pxi * PX + w / 8;
  const int cy = pyi * 8 + (w & 7);
  return cx * nb1 + cy;
}

constexpr int kTileThreads = 512;
constexpr int kBatch = 8;  // independent loads kept in flight per thread in the runtime-shape tile loops

// Compile-time brick shape when BX > 0 (integer divisions by the tile extents become
// shifts/multiplies); BX = 0 reads the shape from the runtime descriptor.
//
// LDS tile layout: tile[(x * t1 + y) * pitch + z], z in [0, b2], pitch = b2 + 2, so every
// corner of a particle is its corner-0 address plus a compile-time offset, and rows are
// 8-byte aligned for float2 LDS stores when b2 is even.  The spare word z = b2 + 1 of each
// scatter row holds the z carry.
template <int BX, int BY, int BZ>
struct BrickShape {
  int b0, b1, b2;
  static constexpr bool kVec = BX > 0 && BZ % 4 == 0;
  __device__ __forceinline__ explicit BrickShape(const Bricks& k)
      : b0(BX ? BX : k.b[0]), b1(BY ? BY : k.b[1]), b2(BZ ? BZ : k.b[2]) {}
  __device__ __forceinline__ int t0() const { return b0 + 1; }
  __device__ __forceinline__ int t1() const { return b1 + 1; }
  __device__ __forceinline__ int t2() const { return b2 + 1; }
  __device__ __forceinline__ int pitch() const { return b2 + 2; }
  __device__ __forceinline__ int rows() const { return (b0 + 1) * (b1 + 1); }
  __device__ __forceinline__ int tfloats() const { return (b0 + 1) * (b1 + 1) * (b2 + 2); }
  __device__ __forceinline__ int cell(int x, int y, int z) const { return (x * (b1 + 1) + y) * (b2 + 2) + z; }
  // offset of corner c relative to corner 0
  __device__ __forceinline__ int corner_off(int c) const {
    return ((c & 1) * (b1 + 1) + ((c >> 1) & 1)) * (b2 + 2) + ((c >> 2) & 1);
  }
};

// Base cell and fractional offset of a particle for the offset-0, same-cell-size CIC of the
// force step: bit-identical to axis_cic_same_cell with offset 0 (rem = 0 there).
__device__ __forceinline__ void base_frac(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                                          int64_t j, float a1, const int n[3], int base[3], float f[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = disp[j * 3 + a] / a1;
    float lf = floorf(t);
    base[a] = pos_mod((int)pmid[j * 3 + a] + (int)lf, n[a]);
    f[a] = t - lf;
  }
}

// Raw particle record (int16 anchor, float32 displacement) held in registers.
struct PRaw {
  int p[3];
  float d[3];
};

// Loads particles start + threadIdx.x + q * kTileThreads (q < PPT), clamped to end - 1 so every
// load is unconditional (callers make sure end > start of the particle range).
template <int PPT>
__device__ __forceinline__ void load_raw(PRaw* raw, const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                                         int64_t start, int64_t end) {
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    int64_t j = start + threadIdx.x + (int64_t)q * kTileThreads;
    j = j < end ? j : end - 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      raw[q].p[a] = pmid[j * 3 + a];
      raw[q].d[a] = disp[j * 3 + a];
    }
  }
}

// base_frac on a raw record (same arithmetic)
__device__ __forceinline__ void raw_base_frac(const PRaw& r, float a1, const int n[3], int base[3], float f[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = r.d[a] / a1;
    float lf = floorf(t);
    base[a] = pos_mod(r.p[a] + (int)lf, n[a]);
    f[a] = t - lf;
  }
}

// Tile cell of a particle's base (corner 0) cell.  The particle was binned by the sort with
// the brick of this same base cell (pm_sort.hip cell_key: the same float32 division, floor and
// periodic wrap), so the local index lies inside the brick; the clamp only keeps the LDS
// access in bounds should that invariant ever be broken.
template <int BX, int BY, int BZ>
__device__ __forceinline__ int local_cell(const BrickShape<BX, BY, BZ>& B, const int bs[3], const int o[3]) {
  int l0 = min(max(bs[0] - o[0], 0), B.b0 - 1);
  int l1 = min(max(bs[1] - o[1], 0), B.b1 - 1);
  int l2 = min(max(bs[2] - o[2], 0), B.b2 - 1);
  return B.cell(l0, l1, l2);
}

// CIC weight of corner c from the fractional offsets, in the reference's product order
__device__ __forceinline__ float frac_weight(const float f[3], int c) {
  float wx = (c & 1) ? 1.f - fabsf(f[0] - 1.f) : 1.f - fabsf(f[0]);
  float wy = ((c >> 1) & 1) ? 1.f - fabsf(f[1] - 1.f) : 1.f - fabsf(f[1]);
  float wz = ((c >> 2) & 1) ? 1.f - fabsf(f[2] - 1.f) : 1.f - fabsf(f[2]);
  return wx * wy * wz;
}

__device__ __forceinline__ float unfix(unsigned long long v) { return (float)(long long)v * kUnfix; }

template <int BX, int BY, int BZ>
__global__ void __launch_bounds__(kTileThreads) k_scatter_columns(
    const int16_t* __restrict__ pmid, const float* __restrict__ disp, const int32_t* __restrict__ bin_off,
    float val, Geom g, Bricks k, float* __restrict__ mesh, float* __restrict__ xs, float* __restrict__ ys,
    unsigned long long* __restrict__ first, const float* __restrict__ vals, const float* __restrict__ vscale) {
  __shared__ __align__(16) unsigned long long tile[kTileMaxWords];
  // per-particle values (sorted order) are brought to |v| <= 1 by vscale[0] before the fixed
  // point and scaled back on output; a constant value is used as it is
  const float vs = vals ? vscale[0] : 1.f;
  const float us = vals ? kUnfix / vs : kUnfix;
  auto unfix_s = [&](unsigned long long x) { return (float)(long long)x * us; };
  const BrickShape<BX, BY, BZ> B(k);
  const int col = column_of(blockIdx.x, gridDim.x);
  const int cx = col / k.nb[1], cy = col - cx * k.nb[1];
  const int P = B.pitch(), rows = B.rows(), carry = B.b2 + 1, tf = B.tfloats();
  const int64_t n1 = k.n[1], n2 = k.n[2];
  for (int i = threadIdx.x; i < tf; i += kTileThreads) tile[i] = 0ull;
  // Particles stream through the column in chunks of kTileThreads * PPT in index order (the
  // bricks' bins are contiguous).  The raw loads of the next chunk are issued before the
  // current one is accumulated, so they are in flight across the flush between bricks.
  constexpr int PPT = 4;
  constexpr int64_t CH = (int64_t)kTileThreads * PPT;
  const int64_t col_begin = bin_off[(int64_t)col * k.nb[2]], col_end = bin_off[(int64_t)(col + 1) * k.nb[2]];
  const bool any = col_end > col_begin;
  PRaw raw[PPT];
  if (any) load_raw<PPT>(raw, pmid, disp, col_begin, col_end);
  __syncthreads();
  for (int bz = 0; bz < k.nb[2]; ++bz) {
    const int o[3] = {cx * B.b0, cy * B.b1, bz * B.b2};
    const int64_t bid = (int64_t)col * k.nb[2] + bz;
    const int64_t j0 = bin_off[bid], j1 = bin_off[bid + 1];
    for (int64_t c0 = j0; c0 < j1; c0 += CH) {
      PRaw cur[PPT];
#pragma unroll
      for (int q = 0; q < PPT; ++q) cur[q] = raw[q];
      load_raw<PPT>(raw, pmid, disp, c0 + CH < j1 ? c0 + CH : j1, col_end);  // next chunk
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        int64_t j = c0 + threadIdx.x + (int64_t)q * kTileThreads;
        if (j >= j1) continue;
        int bs[3];
        float fr[3];
        raw_base_frac(cur[q], g.a1, k.n, bs, fr);
        const int c0i = local_cell(B, bs, o);
        const float v = vals ? vals[j] * vs : val;
#pragma unroll
        for (int c = 0; c < 8; ++c)
          atomicAdd(&tile[c0i + B.corner_off(c)], (unsigned long long)__float2ll_rn(v * frac_weight(fr, c) * kFix));
      }
    }
    __syncthreads();
    // z carry: the plane z = b2 is the next brick's z = 0 plane
    for (int r = threadIdx.x; r < rows; r += kTileThreads) {
      unsigned long long* row = tile + r * P;
      const unsigned long long cin = row[carry];
      row[carry] = row[B.b2];
      row[0] += cin;
      if (bz == 0) first[(int64_t)col * rows + r] = row[0];  // completed after the last brick
    }
    __syncthreads();
    // owned cells x < b0, y < b1, z < b2, then the x == b0 and y == b1 planes (z < b2)
    if constexpr (BrickShape<BX, BY, BZ>::kVec) {
      constexpr int C4 = BZ / 4;
      constexpr int OWN = BX * BY * C4, XP = (BY + 1) * C4, YP = BX * C4;
      for (int it = threadIdx.x; it < OWN + XP + YP; it += kTileThreads) {
        int x, y, c4;
        float* dst;
        if (it < OWN) {
          int r = it / C4;
          c4 = it - r * C4;
          x = r / BY;
          y = r - x * BY;
          dst = &mesh[((int64_t)(o[0] + x) * n1 + (o[1] + y)) * n2 + o[2] + 4 * c4];
        } else if (it < OWN + XP) {
          int r = (it - OWN) / C4;
          c4 = (it - OWN) - r * C4;
          x = BX;
          y = r;
          dst = &xs[((int64_t)col * (BY + 1) + y) * n2 + o[2] + 4 * c4];
        } else {
          int r = (it - OWN - XP) / C4;
          c4 = (it - OWN - XP) - r * C4;
          x = r;
          y = BY;
          dst = &ys[((int64_t)col * BX + x) * n2 + o[2] + 4 * c4];
        }
        const unsigned long long* src = &tile[B.cell(x, y, 4 * c4)];
        longlong2 lo = *reinterpret_cast<const longlong2*>(src), hi = *reinterpret_cast<const longlong2*>(src + 2);
        *reinterpret_cast<float4*>(dst) =
            make_float4((float)lo.x * us, (float)lo.y * us, (float)hi.x * us, (float)hi.y * us);
      }
    } else {
      const int own = B.b0 * B.b1 * B.b2, xp = B.t1() * B.b2, yp = B.b0 * B.b2;
      for (int i = threadIdx.x; i < own + xp + yp; i += kTileThreads) {
        if (i < own) {
          int r = i / B.b2, z = i - r * B.b2;
          int x = r / B.b1, y = r - x * B.b1;
          mesh[((int64_t)(o[0] + x) * n1 + (o[1] + y)) * n2 + (o[2] + z)] = unfix_s(tile[B.cell(x, y, z)]);
        } else if (i < own + xp) {
          int r = (i - own) / B.b2, z = (i - own) - r * B.b2;
          xs[((int64_t)col * B.t1() + r) * n2 + o[2] + z] = unfix_s(tile[B.cell(B.b0, r, z)]);
        } else {
          int r = (i - own - xp) / B.b2, z = (i - own - xp) - r * B.b2;
          ys[((int64_t)col * B.b0 + r) * n2 + o[2] + z] = unfix_s(tile[B.cell(r, B.b1, z)]);
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < tf; i += kTileThreads)
      if (i % P != carry) tile[i] = 0ull;
    __syncthreads();
  }
  // the first brick's z = 0 plane: its own part plus the carry of the last brick (periodic)
  for (int r = threadIdx.x; r < rows; r += kTileThreads) {
    const int x = r / B.t1(), y = r - x * B.t1();
    const float v = unfix_s(first[(int64_t)col * rows + r] + tile[r * P + carry]);
    if (x < B.b0 && y < B.b1)
      mesh[((int64_t)(cx * B.b0 + x) * n1 + (cy * B.b1 + y)) * n2] = v;
    else if (x == B.b0)
      xs[((int64_t)col * B.t1() + y) * n2] = v;
    else
      ys[((int64_t)col * B.b0 + x) * n2] = v;
  }
}

// Add the upper x/y planes of the neighbouring columns into the lower x/y planes of each
// column.  A cell with lx = X % b0 and ly = Y % b1 receives
//   lx == 0: XS of column (cx-1, cy) row ly; and when ly == 0 also XS of (cx-1, cy-1) row b1
//   ly == 0: YS of column (cx, cy-1) row lx
// (columns periodic).  One thread per z element, z fastest: every access is coalesced.
__global__ void __launch_bounds__(kBlock) k_fold_xy(Bricks k, const float* __restrict__ xs,
                                                    const float* __restrict__ ys, float* __restrict__ mesh) {
  const int b0 = k.b[0], b1 = k.b[1];
  const int nb0 = k.nb[0], nb1 = k.nb[1];
  const int64_t nx = k.n[0], ny = k.n[1], nz = k.n[2];
  const int64_t nxp = (int64_t)nb0 * ny * nz;        // x-plane cells
  const int64_t nyp = (nx - nb0) * nb1 * nz;         // y-plane cells with lx != 0
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nxp + nyp;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t X, Y, Z;
    float add;
    if (i < nxp) {
      Z = i % nz;
      int64_t t = i / nz;
      Y = t % ny;
      const int cxp = (int)(t / ny);
      X = (int64_t)cxp * b0;
      const int cy = (int)(Y / b1), ly = (int)(Y - (int64_t)cy * b1);
      const int sx = cxp == 0 ? nb0 - 1 : cxp - 1;
      add = xs[(((int64_t)sx * nb1 + cy) * (b1 + 1) + ly) * nz + Z];
      if (ly == 0) {
        const int sy = cy == 0 ? nb1 - 1 : cy - 1;
        add += xs[(((int64_t)sx * nb1 + sy) * (b1 + 1) + b1) * nz + Z];
        add += ys[(((int64_t)cxp * nb1 + sy) * b0) * nz + Z];
      }
    } else {
      const int64_t j = i - nxp;
      Z = j % nz;
      int64_t t = j / nz;
      const int64_t xr = t % (nx - nb0);
      const int cyp = (int)(t / (nx - nb0));
      const int cx = (int)(xr / (b0 - 1)), lx = 1 + (int)(xr - (int64_t)cx * (b0 - 1));
      X = (int64_t)cx * b0 + lx;
      Y = (int64_t)cyp * b1;
      const int sy = cyp == 0 ? nb1 - 1 : cyp - 1;
      add = ys[(((int64_t)cx * nb1 + sy) * b0 + lx) * nz + Z];
    }
    const int64_t flat = (X * ny + Y) * nz + Z;
    mesh[flat] = mesh[flat] + add;
  }
}

// Particles of one brick processed per tile staging: their three interpolated components
// wait in LDS, so each component tile is read from HBM once per brick unless the brick holds
// more than kGatherCap particles (rare: ~2x the mean occupancy at 1/8 particle per cell).
constexpr int kGatherCap = 2048;
constexpr int kGatherPPT = kGatherCap / kTileThreads;

// Stage rows z in [0, b2) of one component's tile for brick (cx, cy, bz) (rows wrap in x, y).
template <int BX, int BY, int BZ>
__device__ __forceinline__ void stage_rows(float* tile, const float* __restrict__ F, const BrickShape<BX, BY, BZ>& B,
                                           const Bricks& k, int cx, int cy, int bz) {
  const int ox = cx * B.b0, oy = cy * B.b1, oz = bz * B.b2;
  if constexpr (BrickShape<BX, BY, BZ>::kVec) {
    constexpr int C4 = BZ / 4;
    constexpr int ITEMS = (BX + 1) * (BY + 1) * C4;
    constexpr int ONE = (ITEMS + kTileThreads - 1) / kTileThreads;
    constexpr int ROUND = ONE <= 6 ? ONE : 4;
#pragma unroll 1
    for (int base = 0; base < ITEMS; base += ROUND * kTileThreads) {
      float4 v[ROUND];
#pragma unroll
      for (int u = 0; u < ROUND; ++u) {  // unconditional (clamped) loads: the round stays in flight
        int it = min(base + (int)threadIdx.x + u * kTileThreads, ITEMS - 1);
        int r = it / C4, c4 = it - r * C4;
        int x = r / (BY + 1), y = r - x * (BY + 1);
        int gx = ox + x, gy = oy + y;
        gx -= gx >= k.n[0] ? k.n[0] : 0;
        gy -= gy >= k.n[1] ? k.n[1] : 0;
        v[u] = *reinterpret_cast<const float4*>(&F[((int64_t)gx * k.n[1] + gy) * k.n[2] + oz + 4 * c4]);
      }
#pragma unroll
      for (int u = 0; u < ROUND; ++u) {
        int it = base + threadIdx.x + u * kTileThreads;
        if (it < ITEMS) {
          int r = it / C4, c4 = it - r * C4;
          float* dst = &tile[r * (BZ + 2) + 4 * c4];
          *reinterpret_cast<float2*>(dst) = make_float2(v[u].x, v[u].y);
          *reinterpret_cast<float2*>(dst + 2) = make_float2(v[u].z, v[u].w);
        }
      }
    }
  } else {
    const int t1 = B.t1(), b2 = B.b2, P = B.pitch();
    const int items = B.rows() * b2;
    for (int i0 = 0; i0 < items; i0 += kTileThreads * kBatch) {
      float v[kBatch];
      int dst[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        int iu = i0 + u * kTileThreads + (int)threadIdx.x;
        int i = min(iu, items - 1);
        int r = i / b2, z = i - r * b2;
        int x = r / t1, y = r - x * t1;
        int gx = ox + x, gy = oy + y;
        gx -= gx >= k.n[0] ? k.n[0] : 0;
        gy -= gy >= k.n[1] ? k.n[1] : 0;
        v[u] = F[((int64_t)gx * k.n[1] + gy) * k.n[2] + oz + z];
        dst[u] = iu < items ? r * P + z : -1;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (dst[u] >= 0) tile[dst[u]] = v[u];
    }
  }
}

// The z = 0 plane of brick (cx, cy, bz) of one component, read directly (rows wrap in x, y).
template <int BX, int BY, int BZ>
__device__ __forceinline__ void load_plane(float* plane, const float* __restrict__ F, const BrickShape<BX, BY, BZ>& B,
                                           const Bricks& k, int cx, int cy, int bz) {
  const int t1 = B.t1();
  for (int r = threadIdx.x; r < B.rows(); r += kTileThreads) {
    int x = r / t1, y = r - x * t1;
    int gx = cx * B.b0 + x, gy = cy * B.b1 + y;
    gx -= gx >= k.n[0] ? k.n[0] : 0;
    gy -= gy >= k.n[1] ? k.n[1] : 0;
    plane[r] = F[((int64_t)gx * k.n[1] + gy) * k.n[2] + (int64_t)bz * B.b2];
  }
}

template <int BX, int BY, int BZ>
__global__ void __launch_bounds__(kTileThreads) k_gather_columns(
    const int16_t* __restrict__ pmid, const float* __restrict__ disp, const int32_t* __restrict__ bin_off,
    const int32_t* __restrict__ perm, Geom g, Bricks k, const float* __restrict__ f0,
    const float* __restrict__ f1, const float* __restrict__ f2, float* __restrict__ acc) {
  constexpr int kLds = BX ? (BX + 1) * (BY + 1) * (BZ + 2) : kTileMaxFloats;
  constexpr int kPl = BX ? (BX + 1) * (BY + 1) : kPlaneMax;
  __shared__ __align__(16) float tile[kLds];
  __shared__ float res[3][kGatherCap];
  __shared__ float plane[2][3][kPl];  // [parity of bz][component]: lower z plane of brick bz
  const BrickShape<BX, BY, BZ> B(k);
  const int col = patch_column_of(blockIdx.x, gridDim.x, k.nb[0], k.nb[1]);
  const int cx = col / k.nb[1], cy = col - cx * k.nb[1];
  const int rows = B.rows(), P = B.pitch(), nb2 = k.nb[2];
  // the brick above the top one is brick 0 (periodic)
  load_plane(plane[nb2 & 1][0], f0, B, k, cx, cy, 0);
  load_plane(plane[nb2 & 1][1], f1, B, k, cx, cy, 0);
  load_plane(plane[nb2 & 1][2], f2, B, k, cx, cy, 0);
  for (int bz = nb2 - 1; bz >= 0; --bz) {
    const int o[3] = {cx * B.b0, cy * B.b1, bz * B.b2};
    const int64_t bid = (int64_t)col * nb2 + bz;
    const int64_t j0 = bin_off[bid], j1 = bin_off[bid + 1];
    const int up = (bz + 1) & 1, me = bz & 1;
    if (j0 == j1) {  // no particles: only this brick's lower planes are needed (by the brick below)
      __syncthreads();
      load_plane(plane[me][0], f0, B, k, cx, cy, bz);
      load_plane(plane[me][1], f1, B, k, cx, cy, bz);
      load_plane(plane[me][2], f2, B, k, cx, cy, bz);
      continue;
    }
    for (int64_t s0 = j0; s0 < j1; s0 += kGatherCap) {
      const int64_t s1 = s0 + kGatherCap < j1 ? s0 + kGatherCap : j1;
      // each particle's tile cell and fractions, once for the three components (clamped
      // loads; their latency overlaps the first tile's staging)
      int c0[kGatherPPT];
      float fr[kGatherPPT][3];
#pragma unroll
      for (int q = 0; q < kGatherPPT; ++q) {
        int64_t j = s0 + threadIdx.x + (int64_t)q * kTileThreads;
        int bs[3];
        base_frac(pmid, disp, j < s1 ? j : s1 - 1, g.a1, k.n, bs, fr[q]);
        c0[q] = local_cell(B, bs, o);
      }
#pragma unroll 1
      for (int comp = 0; comp < 3; ++comp) {
        const float* __restrict__ F = comp == 0 ? f0 : (comp == 1 ? f1 : f2);
        __syncthreads();  // the previous component's interpolation is done with the tile
        stage_rows<BX, BY, BZ>(tile, F, B, k, cx, cy, bz);
        for (int r = threadIdx.x; r < rows; r += kTileThreads) tile[r * P + B.b2] = plane[up][comp][r];
        __syncthreads();
        for (int r = threadIdx.x; r < rows; r += kTileThreads) plane[me][comp][r] = tile[r * P];
#pragma unroll
        for (int q = 0; q < kGatherPPT; ++q) {
          const int jl = threadIdx.x + q * kTileThreads;
          if (s0 + jl >= s1) continue;
          float sum = 0.f;
#pragma unroll
          for (int c = 0; c < 8; ++c) sum = sum + tile[c0[q] + B.corner_off(c)] * frac_weight(fr[q], c);
          res[comp][jl] = sum;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kGatherPPT; ++q) {
        const int jl = threadIdx.x + q * kTileThreads;
        if (s0 + jl < s1) {
          int64_t i = perm[s0 + jl];
          acc[i * 3 + 0] = res[0][jl];
          acc[i * 3 + 1] = res[1][jl];
          acc[i * 3 + 2] = res[2][jl];
        }
      }
    }
  }
}

// The displacement cotangent of the force step, tiled like k_gather_columns: for sorted
// particle j, dc = sum_m coef_m(j) sum_c M_m(c) grad W_c(j) with the meshes M = (F_x, F_y, F_z,
// dens_cot) and coef = (acc_cot of j, x, y, z, from the sorted component-major copies; the
// particle value), written as dc / a1 through perm: the reverse of the three gathers
// (pallas.py:321-355) and the scatter adjoint's displacement term (pallas.py:358-409) in one
// pass.  grad W_c = (s_x w_y w_z, w_x s_y w_z, w_x w_y s_z) with s = sign(-(f - b)), sign(0) =
// 0 (the corner_grad convention).  Each mesh's tile is staged in turn; a particle's partial sums
// wait in its thread's own LDS slots.  With a zero particle value the density term is skipped.
template <int BX, int BY, int BZ>
__global__ void __launch_bounds__(kTileThreads) k_disp_cot_columns(
    const int16_t* __restrict__ pmid, const float* __restrict__ disp, const int32_t* __restrict__ bin_off,
    const int32_t* __restrict__ perm, Geom g, Bricks k, const float* __restrict__ f0, const float* __restrict__ f1,
    const float* __restrict__ f2, const float* __restrict__ dcot, const float* __restrict__ cot_s, int64_t n,
    float val, float* __restrict__ disp_cot, int accum) {
  constexpr int kLds = BX ? (BX + 1) * (BY + 1) * (BZ + 2) : kTileMaxFloats;
  constexpr int kPl = BX ? (BX + 1) * (BY + 1) : kPlaneMax;
  __shared__ __align__(16) float tile[kLds];
  __shared__ float res[3][kGatherCap];
  __shared__ float plane[2][4][kPl];  // [parity of bz][mesh]: lower z plane of brick bz
  const BrickShape<BX, BY, BZ> B(k);
  const int col = patch_column_of(blockIdx.x, gridDim.x, k.nb[0], k.nb[1]);
  const int cx = col / k.nb[1], cy = col - cx * k.nb[1];
  const int rows = B.rows(), P = B.pitch(), nb2 = k.nb[2];
  const int nm = val != 0.f ? 4 : 3;
  const int m0 = f0 ? 0 : 3;  // the first mesh: the force components, or the density cotangent alone
  // the brick above the top one is brick 0 (periodic)
  for (int m = m0; m < nm; ++m)
    load_plane(plane[nb2 & 1][m], m == 0 ? f0 : (m == 1 ? f1 : (m == 2 ? f2 : dcot)), B, k, cx, cy, 0);
  for (int bz = nb2 - 1; bz >= 0; --bz) {
    const int o[3] = {cx * B.b0, cy * B.b1, bz * B.b2};
    const int64_t bid = (int64_t)col * nb2 + bz;
    const int64_t j0 = bin_off[bid], j1 = bin_off[bid + 1];
    const int up = (bz + 1) & 1, me = bz & 1;
    if (j0 == j1) {  // no particles: only this brick's lower planes are needed (by the brick below)
      __syncthreads();
      for (int m = m0; m < nm; ++m)
        load_plane(plane[me][m], m == 0 ? f0 : (m == 1 ? f1 : (m == 2 ? f2 : dcot)), B, k, cx, cy, bz);
      continue;
    }
    for (int64_t s0 = j0; s0 < j1; s0 += kGatherCap) {
      const int64_t s1 = s0 + kGatherCap < j1 ? s0 + kGatherCap : j1;
      // each particle's tile cell and fractions, once for every mesh
      int c0[kGatherPPT];
      float fr[kGatherPPT][3];
#pragma unroll
      for (int q = 0; q < kGatherPPT; ++q) {
        int64_t j = s0 + threadIdx.x + (int64_t)q * kTileThreads;
        int bs[3];
        base_frac(pmid, disp, j < s1 ? j : s1 - 1, g.a1, k.n, bs, fr[q]);
        c0[q] = local_cell(B, bs, o);
      }
#pragma unroll 1
      for (int m = m0; m < nm; ++m) {
        const float* __restrict__ F = m == 0 ? f0 : (m == 1 ? f1 : (m == 2 ? f2 : dcot));
        __syncthreads();  // the previous mesh's pass is done with the tile
        stage_rows<BX, BY, BZ>(tile, F, B, k, cx, cy, bz);
        for (int r = threadIdx.x; r < rows; r += kTileThreads) tile[r * P + B.b2] = plane[up][m][r];
        __syncthreads();
        for (int r = threadIdx.x; r < rows; r += kTileThreads) plane[me][m][r] = tile[r * P];
        float cf[kGatherPPT];
#pragma unroll
        for (int q = 0; q < kGatherPPT; ++q) {  // this mesh's coefficients (clamped loads)
          const int64_t j = s0 + threadIdx.x + (int64_t)q * kTileThreads;
          cf[q] = m < 3 ? cot_s[(int64_t)m * n + (j < s1 ? j : s1 - 1)] : val;
        }
#pragma unroll
        for (int q = 0; q < kGatherPPT; ++q) {
          const int jl = threadIdx.x + q * kTileThreads;
          if (s0 + jl >= s1) continue;
          float w[3][2], sg[3][2];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            float f = fr[q][a];
            asm volatile("" : "+v"(f));  // weights per mesh pass: kept out of registers between passes
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) {
              const float delta = f - (float)bb;
              w[a][bb] = 1.f - fabsf(delta);
              sg[a][bb] = cic_sign(delta);
            }
          }
          float gs[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const int bx = c & 1, by = (c >> 1) & 1, bzz = (c >> 2) & 1;
            const float v = tile[c0[q] + B.corner_off(c)];
            gs[0] = gs[0] + v * (sg[0][bx] * w[1][by] * w[2][bzz]);
            gs[1] = gs[1] + v * (w[0][bx] * sg[1][by] * w[2][bzz]);
            gs[2] = gs[2] + v * (w[0][bx] * w[1][by] * sg[2][bzz]);
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) res[a][jl] = (m == m0 ? 0.f : res[a][jl]) + cf[q] * gs[a];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kGatherPPT; ++q) {
        const int jl = threadIdx.x + q * kTileThreads;
        if (s0 + jl < s1) {
          const int64_t i = perm[s0 + jl];
#pragma unroll
          for (int a = 0; a < 3; ++a) disp_cot[i * 3 + a] = (accum ? disp_cot[i * 3 + a] : 0.f) + res[a][jl] / g.a1;
        }
      }
    }
  }
}

// invalid particles (sorted last, past the last brick) get a zero acceleration
__global__ void k_zero_invalid(const int32_t* __restrict__ perm, const int32_t* __restrict__ bin_off,
                               int64_t nbins, int64_t n, float* __restrict__ acc) {
  for (int64_t j = bin_off[nbins] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = perm[j];
    acc[i * 3 + 0] = 0.f;
    acc[i * 3 + 1] = 0.f;
    acc[i * 3 + 2] = 0.f;
  }
}

// Reverse of the three gathers without brick tiling (gather^T = float-atomic scatter of
// acc_cot, plus the displacement cotangent from the mesh values), pallas.py:321-355 per
// component.
__global__ void __launch_bounds__(kBlock) k_gather3_bwd_sorted(
    const int16_t* __restrict__ pmid, const float* __restrict__ disp, const uint8_t* __restrict__ valid,
    const int32_t* __restrict__ perm, int64_t n, Geom g, const float* __restrict__ fx,
    const float* __restrict__ fy, const float* __restrict__ fz, const float* __restrict__ acc_cot,
    float* __restrict__ gx, float* __restrict__ gy, float* __restrict__ gz, float* __restrict__ disp_cot_s) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float dc[3] = {0.f, 0.f, 0.f};
  if (valid[j]) {
    int64_t i = perm[j];
    float ac[3] = {acc_cot[i * 3 + 0], acc_cot[i * 3 + 1], acc_cot[i * 3 + 2]};
    AxisCic ax[3];
    particle_cic(g, pmid, disp, j, ax);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      corner_index(g, ax, c, flat);
      float w = corner_weight(ax, c);
      float gr[3];
      corner_grad(ax, c, gr);
      atomic_add_f32(gx + flat, ac[0] * w);
      atomic_add_f32(gy + flat, ac[1] * w);
      atomic_add_f32(gz + flat, ac[2] * w);
      float inner = ac[0] * fx[flat] + ac[1] * fy[flat] + ac[2] * fz[flat];
#pragma unroll
      for (int a = 0; a < 3; ++a) dc[a] = dc[a] + inner * gr[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) disp_cot_s[j * 3 + a] = dc[a] / g.a1;
}

// Sorted, component-major copies of the acceleration cotangent (cot_s[c n + j] = acc_cot of
// sorted particle j) for the tiled gather-adjoint scatter, and per-block maxima of |acc_cot|.
__global__ void __launch_bounds__(kBlock) k_cot_sorted(const int32_t* __restrict__ perm,
                                                       const float* __restrict__ acc_cot, int64_t n,
                                                       float* __restrict__ cot_s, float* __restrict__ pmax) {
  float mx = 0.f;  // +inf once a non-finite value is seen (fmaxf alone would skip a NaN)
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    const int64_t i = perm[j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = acc_cot[i * 3 + c];
      cot_s[c * n + j] = v;
      mx = isfinite(v) ? fmaxf(mx, fabsf(v)) : INFINITY;
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  __shared__ float part[kBlock / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int w = 0; w < kBlock / 64; ++w) m = fmaxf(m, part[w]);
    pmax[blockIdx.x] = m;
  }
}

// scale[0] = 1 / max |acc_cot| (1 when all are zero; capped at 2^127 when the maximum is so
// small that its reciprocal overflows): the values enter the fixed point at magnitude <= 1.
// A non-finite cotangent makes the scale NaN, and the tiled scatter's output scale (the
// reciprocal) with it, so every mesh cotangent and with them every particle's displacement
// cotangent and the Omega_m cotangent come out NaN, as the unfused float chain would give.
__global__ void k_cot_scale(const float* __restrict__ pmax, int nb, float* __restrict__ scale) {
  if (threadIdx.x != 0) return;
  float m = 0.f;
  for (int i = 0; i < nb; ++i) m = fmaxf(m, pmax[i]);
  scale[0] = !(m <= 3.402823466e38f) ? __int_as_float(0x7fc00000) : (m > 0.f ? fminf(1.f / m, 1.7014118e38f) : 1.f);
}

// Scatter adjoint for the density: disp_cot += v * sum_c dens_cot[c] grad_c / cell, then the
// total is written back in caller order (scatter.py:341-375, pallas.py:358-409).
__global__ void __launch_bounds__(kBlock) k_scatter_bwd_sorted(
    const int16_t* __restrict__ pmid, const float* __restrict__ disp, const uint8_t* __restrict__ valid,
    const int32_t* __restrict__ perm, int64_t n, Geom g, float val, const float* __restrict__ dens_cot,
    const float* __restrict__ disp_cot_s, float* __restrict__ disp_cot) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float dc[3] = {0.f, 0.f, 0.f};
  if (valid[j] && val != 0.f) {
    AxisCic ax[3];
    particle_cic(g, pmid, disp, j, ax);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      corner_index(g, ax, c, flat);
      float gr[3];
      corner_grad(ax, c, gr);
      float mv = dens_cot[flat] * val;
#pragma unroll
      for (int a = 0; a < 3; ++a) dc[a] = dc[a] + mv * gr[a];
    }
  }
  int64_t i = perm[j];
#pragma unroll
  for (int a = 0; a < 3; ++a)  // added to the gather's terms (disp_cot_s), or to disp_cot itself when null
    disp_cot[i * 3 + a] = (disp_cot_s ? disp_cot_s[j * 3 + a] : disp_cot[i * 3 + a]) + dc[a] / g.a1;
}

// omega_cot += 1.5 * sum_x src_cot(x) * (rho(x) [+ rho_add(x)] - 1); dens_cot = src_cot * scale,
// in place, and copied to rho_cot when given (the cotangent of an added density)
__global__ void __launch_bounds__(kBlock) k_source_bwd(float* __restrict__ src_cot,
                                                       const float* __restrict__ rho,
                                                       const float* __restrict__ rho_add,
                                                       float* __restrict__ rho_cot, int64_t n,
                                                       float scale, double* __restrict__ part) {
  // sum_x src_cot (rho - 1): float products, float64 sums (per thread, then per block in a
  // fixed order), one partial per block; float4 accesses when every array is 16-byte aligned
  const bool vec = ((reinterpret_cast<uintptr_t>(src_cot) | reinterpret_cast<uintptr_t>(rho) |
                     reinterpret_cast<uintptr_t>(rho_add) | reinterpret_cast<uintptr_t>(rho_cot)) & 15) == 0;
  const int64_t n4 = vec ? n / 4 : 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (int64_t q = g0; q < n4; q += stride) {
    const float4 sc = reinterpret_cast<const float4*>(src_cot)[q];
    float4 r = reinterpret_cast<const float4*>(rho)[q];
    if (rho_add) {
      const float4 ra = reinterpret_cast<const float4*>(rho_add)[q];
      r = make_float4(r.x + ra.x, r.y + ra.y, r.z + ra.z, r.w + ra.w);
    }
    acc += (double)(sc.x * (r.x - 1.f)) + (double)(sc.y * (r.y - 1.f)) + (double)(sc.z * (r.z - 1.f)) +
           (double)(sc.w * (r.w - 1.f));
    const float4 o = make_float4(sc.x * scale, sc.y * scale, sc.z * scale, sc.w * scale);
    reinterpret_cast<float4*>(src_cot)[q] = o;
    if (rho_cot) reinterpret_cast<float4*>(rho_cot)[q] = o;
  }
  for (int64_t i = 4 * n4 + g0; i < n; i += stride) {
    const float sc = src_cot[i];
    float r = rho[i];
    if (rho_add) r += rho_add[i];
    acc += (double)(sc * (r - 1.f));
    src_cot[i] = sc * scale;
    if (rho_cot) rho_cot[i] = sc * scale;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double wsum[kBlock / 64];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

// omega_cot[0] += 1.5 * (the block partials of k_source_bwd, summed in block order)
__global__ void k_source_bwd_sum(const double* __restrict__ part, int nb, float* __restrict__ omega_cot) {
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += part[b];
  omega_cot[0] += 1.5f * (float)t;
}

constexpr int kSrcBlocks = 2048;  // blocks of k_source_bwd (and its partials)

struct Layout {
  int nx, ny, nz, m;
  int64_t pitch;  // complex row pitch (float2) of a full half spectrum, >= m+1, multiple of 16
  int64_t nreal;  // nx*ny*nz
  int64_t ncplx;  // nx*ny*pitch
  // Band of the force pipeline.  With the particle-Nyquist filter active every mode with
  // |k_i| > pi/l_p (1 + 8 eps) is zero after the Green multiply, so the passes after the z
  // r2c keep only kz < nkz and, after the y forward pass, only the 2*ky_keep+1 rows with
  // |ky| <= ky_keep (packed).  Unfiltered axes keep everything (nkz = m+1, ky_keep = -1).
  int nkz;         // stored spectrum columns (kz = 0..nkz-1); nkz = m+1 includes Nyquist
  int ky_keep;     // packed y band half-width, -1 = full rows
  int nky;         // rows after the y forward pass (2*ky_keep+1 or ny)
  int64_t pitch1;  // row pitch of the banded spectra, multiple of 16
  int64_t s1;      // nx*ny*pitch1:  full-y banded spectrum
  int64_t s2;      // nx*nky*pitch1: packed-y banded spectrum
};

// Largest |m| on an axis of n points whose wavenumber 2 pi m / (n cell) can pass the
// particle-Nyquist mask |k| <= pi/l_p (1 + 8 eps_f32) (core/utils.py:33-37).  The slack of
// 1e-4 makes the band a superset of the device-side float32 mask, which still decides
// every mode inside the band.
int band_half_width(int n, double cell, double spacing) {
  double kmax = 3.14159265358979323846 / spacing * (1.0 + 8.0 * 1.1920928955078125e-07) * (1.0 + 1e-4);
  double dk = 2.0 * 3.14159265358979323846 / ((double)n * cell);
  return (int)std::floor(kmax / dk);
}

Layout make_layout(const pm_gravity_desc* d) {
  const int32_t* s = d->mesh_shape;
  Layout L;
  L.nx = s[0];
  L.ny = s[1];
  L.nz = s[2];
  L.m = s[2] / 2;
  L.pitch = (int64_t)align_up((size_t)L.m + 1, 16);
  L.nreal = (int64_t)L.nx * L.ny * L.nz;
  L.ncplx = (int64_t)L.nx * L.ny * L.pitch;
  L.nkz = L.m + 1;
  L.ky_keep = -1;
  L.nky = L.ny;
  if (d->mask_active) {
    int kz = band_half_width(L.nz, d->cell_size, d->ptcl_spacing);
    if (kz < L.m) L.nkz = kz + 1;
    int ky = band_half_width(L.ny, d->cell_size, d->ptcl_spacing);
    if (2 * ky + 1 < L.ny) {
      L.ky_keep = ky;
      L.nky = 2 * ky + 1;
    }
  }
  L.pitch1 = (int64_t)align_up((size_t)L.nkz, 16);
  L.s1 = (int64_t)L.nx * L.ny * L.pitch1;
  L.s2 = (int64_t)L.nx * L.nky * L.pitch1;
  return L;
}

int check_desc(const pm_gravity_desc* d, int64_t n) {
  PM_RETURN_IF(d == nullptr, PM_ERROR_INVALID_ARGUMENT, "desc is NULL");
  PM_RETURN_IF(n < 0 || n > INT32_MAX, PM_ERROR_UNSUPPORTED, "particle count must be in [0, 2^31)");
  PM_RETURN_IF(!fft::supported_shape(d->mesh_shape), PM_ERROR_UNSUPPORTED,
               "mesh axes must be powers of two in [2, 2048]");
  PM_RETURN_IF(!(d->cell_size > 0.f), PM_ERROR_INVALID_ARGUMENT, "cell_size must be positive");
  PM_RETURN_IF(d->mask_active && !(d->ptcl_spacing > 0.f), PM_ERROR_INVALID_ARGUMENT,
               "ptcl_spacing must be positive when the particle-Nyquist mask is active");
  return sort::check_sort_args(n, d->mesh_shape, d->brick);
}

constexpr int kCotBlocks = 1024;  // blocks of k_cot_sorted (and slots of their maxima)

struct Buffers {
  int32_t* perm;
  int16_t* pmid_s;
  float* disp_s;
  uint8_t* valid_s;
  float* tables;
  int32_t* bins;   // brick offsets of the bucketed particles
  bool tiled;      // brick-tiled scatter/gather (the brick divides the mesh and fits LDS)
  float* xs;       // upper x / y planes of every brick column (tiled scatter)
  float* ys;
  unsigned long long* first;  // first brick's lower z plane per column (tiled scatter)
  float* r0;  // density, then F_x
  float* r1;  // F_y
  float* r2;  // F_z
  float2* ca;
  float2* cb;
  float2* cc;
  float2* cd;  // packed-y spectra (only when the y band is packed)
  float2* ce;
  // reverse only: F_x (the forward keeps rho in r0), three mesh cotangents
  float* g0;
  float* g1;
  float* g2;
  float* g3;
  float* disp_cot_s;
  float* cot_s;      // sorted acc_cot, component-major (tiled reverse)
  float* cot_pmax;   // per-block maxima, then the fixed-point scale
  double* src_part;  // per-block partials of the Omega_m cotangent
};

bool carve(Arena& ws, const Layout& L, int64_t n, const int32_t* shape, const int32_t* brick, bool reverse,
           Buffers& b) {
  b.xs = b.ys = nullptr;
  b.first = nullptr;
  int64_t nbins = 1;
  for (int a = 0; a < 3; ++a) nbins *= ceil_div(shape[a], brick[a]);
  b.bins = ws.take<int32_t>(nbins + 2);
  b.tiled = tiled_ok(shape, brick);
  if (b.tiled) {
    Bricks k = make_bricks(shape, brick);
    int64_t ncols = (int64_t)k.nb[0] * k.nb[1];
    b.xs = ws.take<float>(xs_floats(k));
    b.ys = ws.take<float>(ys_floats(k));
    b.first = ws.take<unsigned long long>(ncols * (k.b[0] + 1) * (k.b[1] + 1));
  }
  b.perm = ws.take<int32_t>(n);
  b.pmid_s = ws.take<int16_t>(3 * n);
  b.disp_s = ws.take<float>(3 * n);
  b.valid_s = ws.take<uint8_t>(n);
  b.tables = ws.take<float>(fft::spec_table_floats(shape));
  b.r0 = ws.take<float>(L.nreal);
  b.r1 = ws.take<float>(L.nreal);
  b.r2 = ws.take<float>(L.nreal);
  // banded force pipeline (see force_mesh_stage and adjoint_spectral_stage): s1-sized ca,
  // cc, ce and s2-sized cb, cd
  b.ca = ws.take<float2>(L.s1);
  b.cb = ws.take<float2>(L.s2);
  b.cc = ws.take<float2>(L.s1);
  b.cd = ws.take<float2>(L.s2);
  b.ce = ws.take<float2>(L.s1);
  if (reverse) {
    b.g0 = ws.take<float>(L.nreal);
    b.g1 = ws.take<float>(L.nreal);
    b.g2 = ws.take<float>(L.nreal);
    b.g3 = ws.take<float>(L.nreal);
    b.disp_cot_s = b.tiled ? nullptr : ws.take<float>(3 * n);
    b.cot_s = b.tiled ? ws.take<float>(3 * n) : nullptr;
    b.cot_pmax = b.tiled ? ws.take<float>(kCotBlocks + 1) : nullptr;
    b.src_part = ws.take<double>(kSrcBlocks);
  }
  return ws.ok();
}

size_t workspace_bytes(const pm_gravity_desc* d, int64_t n, bool reverse) {
  Layout L = make_layout(d);
  Arena probe(reinterpret_cast<void*>(uintptr_t(256)), SIZE_MAX / 2);
  Buffers b;
  carve(probe, L, n, d->mesh_shape, d->brick, reverse, b);
  // the sort's temporaries are carved after the step buffers
  return probe.used + 256 + sort::bucket_workspace_bytes(d->mesh_shape, d->brick) + 256;
}

Geom step_geom(const pm_gravity_desc* d) {
  Geom g;
  for (int a = 0; a < 3; ++a) {
    g.ls[a] = d->mesh_shape[a];
    g.gs[a] = d->mesh_shape[a];
    g.off[a] = 0.f;
  }
  g.a1 = d->cell_size;
  g.a2 = 0.f;
  return g;
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kBlock)); }

int fill(float* p, int64_t n, hipStream_t s) {
  if (n == 0) return PM_SUCCESS;
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, kBlock * 4), 1 << 16);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(kBlock), 0, s, p, n, 0.f);
  PM_LAUNCHED("k_fill", s);
  return PM_SUCCESS;
}

// Bucketing by brick (with the payload gathered into bucket order), density.  Leaves rho in b.r0.
// Tiled CIC scatter of the bucketed particles into mesh (whose every cell it writes): the
// constant particle value, or per-particle values vals (sorted order) with vscale[0] * |vals|
// <= 1; then the fold of the neighbouring columns' upper x/y planes.
int scatter_tiled(const pm_gravity_desc* d, const Buffers& b, float* mesh, const float* vals, const float* vscale,
                  hipStream_t s) {
  const Geom g = step_geom(d);
  Bricks k = make_bricks(d->mesh_shape, d->brick);
  const unsigned ncols = (unsigned)((int64_t)k.nb[0] * k.nb[1]);
  if (is_brick(k, 16, 16, 32))
    hipLaunchKernelGGL((k_scatter_columns<16, 16, 32>), dim3(ncols), dim3(kTileThreads), 0, s, b.pmid_s, b.disp_s,
                       b.bins, d->particle_value, g, k, mesh, b.xs, b.ys, b.first, vals, vscale);
  else
    hipLaunchKernelGGL((k_scatter_columns<0, 0, 0>), dim3(ncols), dim3(kTileThreads), 0, s, b.pmid_s, b.disp_s,
                       b.bins, d->particle_value, g, k, mesh, b.xs, b.ys, b.first, vals, vscale);
  PM_LAUNCHED("k_scatter_columns", s);
  const int64_t plane_cells = ((int64_t)k.nb[0] * k.n[1] + (int64_t)(k.n[0] - k.nb[0]) * k.nb[1]) * k.n[2];
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(plane_cells, kBlock), 8192));
  hipLaunchKernelGGL(k_fold_xy, dim3(blocks), dim3(kBlock), 0, s, k, b.xs, b.ys, mesh);
  PM_LAUNCHED("k_fold_xy", s);
  return PM_SUCCESS;
}

// The particles binned by brick into the workspace's sorted copies (b.perm, b.bins, b.pmid_s,
// b.disp_s, b.valid_s).
int bucket_stage(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                 const pm_gravity_desc* d, Buffers& b, Arena& ws, hipStream_t s) {
  // the Lagrangian grid the particle count implies (mesh * cell / spacing per axis)
  int32_t grid[3] = {0, 0, 0};
  if (d->ptcl_spacing > 0.f)
    for (int a = 0; a < 3; ++a) {
      const double v = (double)d->mesh_shape[a] * d->cell_size / d->ptcl_spacing;
      grid[a] = (v >= 1.0 && v < 1e9) ? (int32_t)lrint(v) : 0;
    }
  return sort::bucket_cells(pmid, disp, valid, n, d->mesh_shape, d->cell_size, d->brick, grid, b.perm, b.bins,
                            b.pmid_s, b.disp_s, b.valid_s, ws, s);
}

int density_stage(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                  const pm_gravity_desc* d, const Layout& L, Buffers& b, Arena& ws, hipStream_t s) {
  const Geom g = step_geom(d);
  PM_TRY(bucket_stage(pmid, disp, valid, n, d, b, ws, s));
  if (b.tiled) return scatter_tiled(d, b, b.r0, nullptr, nullptr, s);
  PM_TRY(fill(b.r0, L.nreal, s));
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter_sorted, dim3(grid_for(n)), dim3(kBlock), 0, s, b.pmid_s, b.disp_s, b.valid_s,
                       n, d->particle_value, g, b.r0);
    PM_LAUNCHED("k_scatter_sorted", s);
  }
  return PM_SUCCESS;
}

int gather_stage(const int16_t* pmid_s, const float* disp_s, const uint8_t* valid_s, const uint8_t* valid,
                 int64_t n, const pm_gravity_desc* d, const Buffers& b, const float* fx, const float* fy,
                 const float* fz, float* acc, hipStream_t s) {
  if (n == 0) return PM_SUCCESS;
  const Geom g = step_geom(d);
  if (b.tiled) {
    Bricks k = make_bricks(d->mesh_shape, d->brick);
    const int64_t nbins = (int64_t)k.nb[0] * k.nb[1] * k.nb[2];
    const unsigned ncols = (unsigned)((int64_t)k.nb[0] * k.nb[1]);
    if (is_brick(k, 16, 16, 32))
      hipLaunchKernelGGL((k_gather_columns<16, 16, 32>), dim3(ncols), dim3(kTileThreads), 0, s, pmid_s, disp_s,
                         b.bins, b.perm, g, k, fx, fy, fz, acc);
    else
      hipLaunchKernelGGL((k_gather_columns<0, 0, 0>), dim3(ncols), dim3(kTileThreads), 0, s, pmid_s, disp_s,
                         b.bins, b.perm, g, k, fx, fy, fz, acc);
    PM_LAUNCHED("k_gather_columns", s);
    if (valid) {
      unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, kBlock), 4096);
      hipLaunchKernelGGL(k_zero_invalid, dim3(blocks), dim3(kBlock), 0, s, b.perm, b.bins, nbins, n, acc);
      PM_LAUNCHED("k_zero_invalid", s);
    }
    return PM_SUCCESS;
  }
  hipLaunchKernelGGL(k_gather3_sorted, dim3(grid_for(n)), dim3(kBlock), 0, s, pmid_s, disp_s, valid_s, b.perm, n, g,
                     fx, fy, fz, acc);
  PM_LAUNCHED("k_gather3_sorted", s);
  return PM_SUCCESS;
}

// Tiled displacement cotangent of the reverse pass: F = (b.g0, b.r1, b.r2), the density
// cotangent in b.g1, the sorted acc_cot copies in b.cot_s; particles outside every brick (the
// invalid ones) get zero.
int disp_cot_tiled(const pm_gravity_desc* d, const Buffers& b, int64_t n, const uint8_t* valid, float* disp_cot,
                   hipStream_t s) {
  const Geom g = step_geom(d);
  Bricks k = make_bricks(d->mesh_shape, d->brick);
  const int64_t nbins = (int64_t)k.nb[0] * k.nb[1] * k.nb[2];
  const unsigned ncols = (unsigned)((int64_t)k.nb[0] * k.nb[1]);
  const float val = d->particle_value;
  if (is_brick(k, 16, 16, 32))
    hipLaunchKernelGGL((k_disp_cot_columns<16, 16, 32>), dim3(ncols), dim3(kTileThreads), 0, s, b.pmid_s, b.disp_s,
                       b.bins, b.perm, g, k, b.g0, b.r1, b.r2, b.g1, b.cot_s, n, val, disp_cot, 0);
  else
    hipLaunchKernelGGL((k_disp_cot_columns<0, 0, 0>), dim3(ncols), dim3(kTileThreads), 0, s, b.pmid_s, b.disp_s,
                       b.bins, b.perm, g, k, b.g0, b.r1, b.r2, b.g1, b.cot_s, n, val, disp_cot, 0);
  PM_LAUNCHED("k_disp_cot_columns", s);
  if (valid) {
    unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, kBlock), 4096);
    hipLaunchKernelGGL(k_zero_invalid, dim3(blocks), dim3(kBlock), 0, s, b.perm, b.bins, nbins, n, disp_cot);
    PM_LAUNCHED("k_zero_invalid", s);
  }
  return PM_SUCCESS;
}

// Density (b.r0) -> force meshes F_x (b.r0), F_y (b.r1), F_z (b.r2).
int force_mesh_stage(const pm_gravity_desc* d, const Layout& L, Buffers& b, const fft::SpecTables& tb,
                     hipStream_t s, bool keep_density, const float* rho_add) {
  using namespace fft;
  const float coef = 1.5f * d->omega_m;  // 1.5 * Omega_m in float32 (gravity.py:630)
  const bool zband = L.nkz <= L.m;      // Nyquist column not stored
  const bool ypack = L.ky_keep >= 0;
  // every pass writes a buffer other than the one it reads (the two-output passes need it):
  //   A (s1) z r2c output, later chi = y inverse of psi (D_z applied by the last z c2r)
  //   B (s2) y forward output (packed |ky| band)
  //   G (s1) psi = x inverse of the Green-multiplied spectrum, later F_x's y inverse
  //   C (s2) D_x psi
  //   E (s1) D_y chi
  float2* A = b.ca;
  float2* B = b.cd;
  float2* G = b.cc;
  float2* C = b.cb;
  float2* E = b.ce;
  const int64_t p1 = L.pitch1;
  ZPass z1{};
  z1.rin = b.r0;
  z1.cmain = A;
  z1.cnyq = zband ? A : A + L.m;
  z1.pitch_r = L.nz;
  z1.pitch_c = p1;
  z1.nyq_stride = p1;
  z1.nlines = (int64_t)L.nx * L.ny;
  z1.src_shift = -1.f;
  z1.src_scale = coef;
  z1.rin_add = rho_add;  // learned density correction (null: none)
  z1.kz_lim = L.nkz;
  z1.nyq = zband ? 0 : 1;
  PM_TRY(launch_z(L.nz, false, z1, tb, s));
  const int ncols = L.nkz;
  // y forward, packing the |ky| band when the filter allows it
  ColPass yf{A, B, nullptr, p1, 1, (int64_t)L.ny * p1, ncols, 0, 0, 0, 0, 0};
  yf.o_row = p1;
  yf.o_batch = (int64_t)L.nky * p1;
  yf.keep = L.ky_keep;
  PM_TRY(launch_cols(L.ny, kColFwd, yf, L.nx, tb, 1, s));
  // x: forward, Green, two inverses; columns are kz, batches are the (packed) ky rows
  ColPass xg{B, G, C, (int64_t)L.nky * p1, 1, p1, ncols, 1, 0, 0, 1, 0};
  if (ypack) {
    xg.ky_split = L.ky_keep;
    xg.ky_shift = L.ny - L.nky;
  }
  PM_TRY(launch_cols(L.nx, kColGreen, xg, L.nky, tb, 0, s));
  // y inverses expand the band back to full rows
  ColPass yi2{G, A, E, p1, 1, (int64_t)L.nky * p1, ncols, 0, 0, 0, 0, 0};
  yi2.o_row = p1;
  yi2.o_batch = (int64_t)L.ny * p1;
  yi2.keep = L.ky_keep;
  yi2.compact_in = 1;
  PM_TRY(launch_cols(L.ny, kColInv2, yi2, L.nx, tb, 1, s));
  ColPass yi{C, G, nullptr, p1, 1, (int64_t)L.nky * p1, ncols, 0, 0, 0, 0, 0};
  yi.o_row = p1;
  yi.o_batch = (int64_t)L.ny * p1;
  yi.keep = L.ky_keep;
  yi.compact_in = 1;
  PM_TRY(launch_cols(L.ny, kColInv, yi, L.nx, tb, 1, s));
  const float norm = (float)(2.0 / ((double)L.nx * L.ny * L.nz));
  ZPass zc{};
  zc.pitch_r = L.nz;
  zc.pitch_c = p1;
  zc.nyq_stride = p1;
  zc.nlines = (int64_t)L.nx * L.ny;
  zc.out_scale = norm;
  zc.kz_lim = L.nkz;
  zc.nyq = zband ? 0 : 1;
  float* fx = keep_density ? b.g0 : b.r0;  // reverse keeps rho for the Omega_m cotangent
  const float2* srcs[3] = {G, E, A};
  float* dsts[3] = {fx, b.r1, b.r2};
  for (int c = 0; c < 3; ++c) {
    zc.cmain = const_cast<float2*>(srcs[c]);
    zc.cnyq = zband ? zc.cmain : zc.cmain + L.m;
    zc.rout = dsts[c];
    zc.deriv = c == 2;  // D_z on the last component
    PM_TRY(launch_z(L.nz, true, zc, tb, s));
  }
  return PM_SUCCESS;
}

// Spectral adjoint of the force meshes: src_cot (b.g1, real) = irfftn(G * sum_i conj(D_i) *
// rfftn(g_i)) for the three cotangent meshes g_i (b.g1, b.g2, b.g3; the symbol is Hermitian, so
// the transpose of irfftn(H rfftn(.)) is irfftn(conj(H) rfftn(.))).  G * D_i vanishes outside
// the particle-Nyquist band, so the passes are force_mesh_stage's in reverse on the same band:
// three z r2c passes storing the kz band, three y passes packing the |ky| band, one x pass
// that transforms the three spectra, combines them mode by mode and inverts (kColAdj), the y
// inverse expanding the band and the z c2r.  Buffers: ca, cc, ce (s1) and cb, cd (s2).
int adjoint_spectral_stage(const pm_gravity_desc* d, const Layout& L, Buffers& b, const fft::SpecTables& tb,
                           hipStream_t s) {
  using namespace fft;
  const bool zband = L.nkz <= L.m;
  const bool ypack = L.ky_keep >= 0;
  const int64_t p1 = L.pitch1;
  const int ncols = L.nkz;
  float* reals[3] = {b.g1, b.g2, b.g3};
  float2* s1buf[3] = {b.ca, b.cc, b.ce};
  ZPass zr{};
  zr.pitch_r = L.nz;
  zr.pitch_c = p1;
  zr.nyq_stride = p1;
  zr.nlines = (int64_t)L.nx * L.ny;
  zr.src_shift = 0.f;
  zr.src_scale = 1.f;
  zr.kz_lim = L.nkz;
  zr.nyq = zband ? 0 : 1;
  for (int c = 0; c < 3; ++c) {
    zr.rin = reals[c];
    zr.cmain = s1buf[c];
    zr.cnyq = zband ? s1buf[c] : s1buf[c] + L.m;
    PM_TRY(launch_z(L.nz, false, zr, tb, s));
  }
  // y forward, packed: ca -> cb, cc -> cd, then ce -> ca (ca is free again; s2 <= s1)
  float2* s2buf[3] = {b.cb, b.cd, b.ca};
  for (int c = 0; c < 3; ++c) {
    ColPass yf{s1buf[c], s2buf[c], nullptr, p1, 1, (int64_t)L.ny * p1, ncols, 0, 0, 0, 0, 0};
    yf.o_row = p1;
    yf.o_batch = (int64_t)L.nky * p1;
    yf.keep = L.ky_keep;
    PM_TRY(launch_cols(L.ny, kColFwd, yf, L.nx, tb, 1, s));
  }
  // x: three forward transforms, the combined symbol, one inverse -> cc (packed band)
  ColPass xa{s2buf[0], b.cc, nullptr, (int64_t)L.nky * p1, 1, p1, ncols, 1, 0, 0, 1, 0};
  xa.in1 = s2buf[1];
  xa.in2 = s2buf[2];
  if (ypack) {
    xa.ky_split = L.ky_keep;
    xa.ky_shift = L.ny - L.nky;
  }
  PM_TRY(launch_cols(L.nx, kColAdj, xa, L.nky, tb, 0, s));
  // y inverse expanding the band: cc -> ce
  ColPass yi{b.cc, b.ce, nullptr, p1, 1, (int64_t)L.nky * p1, ncols, 0, 0, 0, 0, 0};
  yi.o_row = p1;
  yi.o_batch = (int64_t)L.ny * p1;
  yi.keep = L.ky_keep;
  yi.compact_in = 1;
  PM_TRY(launch_cols(L.ny, kColInv, yi, L.nx, tb, 1, s));
  ZPass zc{};
  zc.pitch_r = L.nz;
  zc.pitch_c = p1;
  zc.nyq_stride = p1;
  zc.nlines = (int64_t)L.nx * L.ny;
  zc.out_scale = (float)(2.0 / ((double)L.nx * L.ny * L.nz));
  zc.kz_lim = L.nkz;
  zc.nyq = zband ? 0 : 1;
  zc.cmain = b.ce;
  zc.cnyq = zband ? b.ce : b.ce + L.m;
  zc.rout = b.g1;
  zc.deriv = 0;
  return launch_z(L.nz, true, zc, tb, s);
}

// The particle-Nyquist-filtered potential of a real mesh on the force step's band:
// out = add + irfftn(mask * (-1/k^2) * rfftn(scale * (in + shift))).  With shift = -1 and
// scale = 1.5 Omega_m on a density mesh it is the potential of the force's source
// (gravity.py:626-633 with the Poisson and filter symbols of :174-196, 263-321); the operator
// is real and symmetric, so with shift = 0 it is also its own adjoint (a cotangent's path back
// to the source).  The passes are force_mesh_stage's for one output: z r2c storing the kz band
// (A), y forward packing the |ky| band (A -> B), the x pass with the Green multiply and one
// inverse (kColPot, B -> G), the y inverse expanding the band (G -> A), z c2r (+ add).
int potential_stage(const Layout& L, float2* A, float2* B, float2* G, const fft::SpecTables& tb, const float* in,
                    float shift, float scale, const float* add, float* out, hipStream_t s) {
  using namespace fft;
  const bool zband = L.nkz <= L.m;
  const int64_t p1 = L.pitch1;
  const int ncols = L.nkz;
  ZPass zr{};
  zr.rin = in;
  zr.cmain = A;
  zr.cnyq = zband ? A : A + L.m;
  zr.pitch_r = L.nz;
  zr.pitch_c = p1;
  zr.nyq_stride = p1;
  zr.nlines = (int64_t)L.nx * L.ny;
  zr.src_shift = shift;
  zr.src_scale = scale;
  zr.kz_lim = L.nkz;
  zr.nyq = zband ? 0 : 1;
  PM_TRY(launch_z(L.nz, false, zr, tb, s));
  ColPass yf{A, B, nullptr, p1, 1, (int64_t)L.ny * p1, ncols, 0, 0, 0, 0, 0};
  yf.o_row = p1;
  yf.o_batch = (int64_t)L.nky * p1;
  yf.keep = L.ky_keep;
  PM_TRY(launch_cols(L.ny, kColFwd, yf, L.nx, tb, 1, s));
  ColPass xp{B, G, nullptr, (int64_t)L.nky * p1, 1, p1, ncols, 1, 0, 0, 1, 0};
  if (L.ky_keep >= 0) {
    xp.ky_split = L.ky_keep;
    xp.ky_shift = L.ny - L.nky;
  }
  PM_TRY(launch_cols(L.nx, kColPot, xp, L.nky, tb, 0, s));
  ColPass yi{G, A, nullptr, p1, 1, (int64_t)L.nky * p1, ncols, 0, 0, 0, 0, 0};
  yi.o_row = p1;
  yi.o_batch = (int64_t)L.ny * p1;
  yi.keep = L.ky_keep;
  yi.compact_in = 1;
  PM_TRY(launch_cols(L.ny, kColInv, yi, L.nx, tb, 1, s));
  ZPass zc{};
  zc.pitch_r = L.nz;
  zc.pitch_c = p1;
  zc.nyq_stride = p1;
  zc.nlines = (int64_t)L.nx * L.ny;
  zc.out_scale = (float)(2.0 / ((double)L.nx * L.ny * L.nz));
  zc.kz_lim = L.nkz;
  zc.nyq = zband ? 0 : 1;
  zc.cmain = A;
  zc.cnyq = zband ? A : A + L.m;
  zc.rout = out;
  zc.rout_add = add;
  zc.deriv = 0;
  return launch_z(L.nz, true, zc, tb, s);
}

size_t potential_workspace_bytes(const pm_gravity_desc* d) {
  const Layout L = make_layout(d);
  return align_up(sizeof(float) * fft::spec_table_floats(d->mesh_shape), 256) +
         align_up(sizeof(float2) * L.s1, 256) * 2 + align_up(sizeof(float2) * L.s2, 256) + 256;
}

int prepare(const pm_gravity_desc* d, int64_t n, void* workspace, size_t workspace_bytes_, bool reverse,
            Layout& L, Buffers& b, Arena& ws, fft::SpecTables& tb, hipStream_t s) {
  PM_TRY(check_desc(d, n));
  PM_TRY(require_device_tables());
  size_t need = workspace_bytes(d, n, reverse);
  PM_RETURN_IF(workspace == nullptr || workspace_bytes_ < need, PM_ERROR_WORKSPACE,
               "gravity workspace too small: need " + std::to_string(need) + " bytes");
  L = make_layout(d);
  PM_RETURN_IF(!carve(ws, L, n, d->mesh_shape, d->brick, reverse, b), PM_ERROR_WORKSPACE,
               "workspace carve failed");
  const fft::SpecTables sc = fft::spec_scalars(d->cell_size, d->ptcl_spacing, d->mask_active);
  PM_TRY(fft::launch_spec_tables(b.tables, d->mesh_shape, sc, s));
  tb = fft::spec_tables_view(b.tables, d->mesh_shape, sc);
  return PM_SUCCESS;
}

}  // namespace
}  // namespace grav
}  // namespace pm

using namespace pm;

extern "C" {

int pm_gravity_workspace(const pm_gravity_desc* desc, int64_t n, size_t* bytes) {
  int st = grav::check_desc(desc, n);
  if (st) return st;
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = grav::workspace_bytes(desc, n, false);
  return PM_SUCCESS;
}

int pm_gravity_vjp_workspace(const pm_gravity_desc* desc, int64_t n, size_t* bytes) {
  int st = grav::check_desc(desc, n);
  if (st) return st;
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = grav::workspace_bytes(desc, n, true);
  return PM_SUCCESS;
}

int pm_gravity(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
               const pm_gravity_desc* desc, float* acc, void* workspace, size_t workspace_bytes,
               void* stream) {
  return pm_gravity_corrected(pmid, disp, valid, n, desc, nullptr, acc, workspace, workspace_bytes, stream);
}

int pm_gravity_corrected(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                         const pm_gravity_desc* desc, const float* rho_add, float* acc, void* workspace,
                         size_t workspace_bytes, void* stream) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  int st = grav::prepare(desc, n, workspace, workspace_bytes, false, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || acc == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "pmid/disp/acc NULL");
  st = grav::density_stage(pmid, disp, valid, n, desc, L, b, ws, s);
  if (st) return st;
  st = grav::force_mesh_stage(desc, L, b, tb, s, false, rho_add);
  if (st) return st;
  return grav::gather_stage(b.pmid_s, b.disp_s, b.valid_s, valid, n, desc, b, b.r0, b.r1, b.r2, acc, s);
}

int pm_gravity_corrected_resume(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                                const pm_gravity_desc* desc, const float* rho_add, float* acc, void* workspace,
                                size_t workspace_bytes, void* stream) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  // the same carve as the pm_gravity_density call: the bucketed particles and the density
  // mesh it left are at the same places
  int st = grav::prepare(desc, n, workspace, workspace_bytes, false, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || acc == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "pmid/disp/acc NULL");
  st = grav::force_mesh_stage(desc, L, b, tb, s, false, rho_add);
  if (st) return st;
  return grav::gather_stage(b.pmid_s, b.disp_s, b.valid_s, valid, n, desc, b, b.r0, b.r1, b.r2, acc, s);
}

int pm_gravity_density_offset(const pm_gravity_desc* desc, int64_t n, size_t* offset) {
  PM_TRY(grav::check_desc(desc, n));
  PM_RETURN_IF(offset == nullptr, PM_ERROR_INVALID_ARGUMENT, "offset is NULL");
  // the forward carve on a stand-in base (never dereferenced): prepare() carves from the
  // workspace start in the same order, so the density's offset is the same there
  char* const base = reinterpret_cast<char*>(uintptr_t(1) << 12);
  Arena ws(base, SIZE_MAX / 2);
  grav::Buffers b{};
  const grav::Layout L = grav::make_layout(desc);
  PM_RETURN_IF(!grav::carve(ws, L, n, desc->mesh_shape, desc->brick, false, b), PM_ERROR_WORKSPACE,
               "workspace carve failed");
  *offset = (size_t)(reinterpret_cast<char*>(b.r0) - base);
  return PM_SUCCESS;
}

int pm_gravity_density(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                       const pm_gravity_desc* desc, float* rho, void* workspace, size_t workspace_bytes,
                       void* stream) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  int st = grav::prepare(desc, n, workspace, workspace_bytes, false, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(rho == nullptr || (n > 0 && (pmid == nullptr || disp == nullptr)), PM_ERROR_INVALID_ARGUMENT,
               "pmid/disp/rho NULL");
  st = grav::density_stage(pmid, disp, valid, n, desc, L, b, ws, s);
  if (st) return st;
  if (rho == b.r0) return PM_SUCCESS;  // the caller reads the density in place (pm_gravity_density_offset)
  if (((reinterpret_cast<uintptr_t>(rho) | reinterpret_cast<uintptr_t>(b.r0)) & 15) == 0) {
    hipLaunchKernelGGL(grav::k_copy_mesh, dim3(2048), dim3(grav::kBlock), 0, s, b.r0, rho, (int64_t)L.nreal);
    PM_LAUNCHED("k_copy_mesh", s);
  } else {
    PM_HIP_TRY(hipMemcpyAsync(rho, b.r0, sizeof(float) * L.nreal, hipMemcpyDeviceToDevice, s));
  }
  return PM_SUCCESS;
}

int pm_gravity_vjp(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                   const pm_gravity_desc* desc, const float* acc_cot, float* disp_cot,
                   float* omega_cot, void* workspace, size_t workspace_bytes, void* stream) {
  return pm_gravity_corrected_vjp(pmid, disp, valid, n, desc, nullptr, acc_cot, disp_cot, omega_cot, nullptr,
                                  workspace, workspace_bytes, stream);
}

// The force VJP; with resume, the forward meshes (and the bucketed particles) are those a
// pm_gravity_keep call left in the same workspace for the same particles.
static int vjp_impl(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                    const pm_gravity_desc* desc, const float* rho_add, const float* acc_cot, float* disp_cot,
                    float* omega_cot, float* rho_cot, void* workspace, size_t workspace_bytes, void* stream,
                    bool resume) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  int st = grav::prepare(desc, n, workspace, workspace_bytes, true, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || acc_cot == nullptr || disp_cot == nullptr),
               PM_ERROR_INVALID_ARGUMENT, "pmid/disp/acc_cot/disp_cot NULL");
  PM_RETURN_IF(omega_cot == nullptr, PM_ERROR_INVALID_ARGUMENT, "omega_cot NULL");
  const Geom g = grav::step_geom(desc);
  // recompute the forward meshes: rho stays in r0, F_x goes to g0, F_y r1, F_z r2
  if (!resume) {
    st = grav::density_stage(pmid, disp, valid, n, desc, L, b, ws, s);
    if (st) return st;
    st = grav::force_mesh_stage(desc, L, b, tb, s, true, rho_add);
    if (st) return st;
  }
  // gather^T: acc_cot scattered into three cotangent meshes, displacement cotangent from F.
  // With brick tiling each component goes through the tiled scatter (sorted, scaled copies of
  // acc_cot); otherwise through float atomics beside the displacement cotangent.
  // The tiled path takes the displacement cotangent after the spectral adjoint, in one pass
  // over F and the density cotangent.
  if (b.tiled) {
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, grav::kBlock), grav::kCotBlocks));
    hipLaunchKernelGGL(grav::k_cot_sorted, dim3(nb), dim3(grav::kBlock), 0, s, b.perm, acc_cot, n, b.cot_s,
                       b.cot_pmax);
    PM_LAUNCHED("k_cot_sorted", s);
    hipLaunchKernelGGL(grav::k_cot_scale, dim3(1), dim3(64), 0, s, b.cot_pmax, (int)nb, b.cot_pmax + grav::kCotBlocks);
    PM_LAUNCHED("k_cot_scale", s);
    float* gm[3] = {b.g1, b.g2, b.g3};
    for (int c = 0; c < 3; ++c) {
      st = grav::scatter_tiled(desc, b, gm[c], b.cot_s + (int64_t)c * n, b.cot_pmax + grav::kCotBlocks, s);
      if (st) return st;
    }
  } else {
    st = grav::fill(b.g1, L.nreal, s);
    if (st) return st;
    st = grav::fill(b.g2, L.nreal, s);
    if (st) return st;
    st = grav::fill(b.g3, L.nreal, s);
    if (st) return st;
    if (n > 0) {
      hipLaunchKernelGGL(grav::k_gather3_bwd_sorted, dim3(grav::grid_for(n)), dim3(grav::kBlock), 0, s,
                         b.pmid_s, b.disp_s, b.valid_s, b.perm, n, g, b.g0, b.r1, b.r2, acc_cot, b.g1, b.g2, b.g3,
                         b.disp_cot_s);
      PM_LAUNCHED("k_gather3_bwd_sorted", s);
    }
  }
  // spectral adjoint: src_cot (g1) = irfftn(G * sum_i conj(D_i) rfftn(g_i)) on the band
  st = grav::adjoint_spectral_stage(desc, L, b, tb, s);
  if (st) return st;
  {
    const float coef = 1.5f * desc->omega_m;
    const unsigned blocks =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(L.nreal, grav::kBlock * 16), grav::kSrcBlocks));
    hipLaunchKernelGGL(grav::k_source_bwd, dim3(blocks), dim3(grav::kBlock), 0, s, b.g1, b.r0, rho_add, rho_cot,
                       L.nreal, coef, b.src_part);
    PM_LAUNCHED("k_source_bwd", s);
    hipLaunchKernelGGL(grav::k_source_bwd_sum, dim3(1), dim3(64), 0, s, b.src_part, (int)blocks, omega_cot);
    PM_LAUNCHED("k_source_bwd_sum", s);
  }
  if (n > 0 && b.tiled) return grav::disp_cot_tiled(desc, b, n, valid, disp_cot, s);
  if (n > 0) {
    hipLaunchKernelGGL(grav::k_scatter_bwd_sorted, dim3(grav::grid_for(n)), dim3(grav::kBlock), 0, s, b.pmid_s,
                       b.disp_s, b.valid_s, b.perm, n, g, desc->particle_value, b.g1, b.disp_cot_s, disp_cot);
    PM_LAUNCHED("k_scatter_bwd_sorted", s);
  }
  return PM_SUCCESS;
}

int pm_gravity_corrected_vjp(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                             const pm_gravity_desc* desc, const float* rho_add, const float* acc_cot,
                             float* disp_cot, float* omega_cot, float* rho_cot, void* workspace,
                             size_t workspace_bytes, void* stream) {
  return vjp_impl(pmid, disp, valid, n, desc, rho_add, acc_cot, disp_cot, omega_cot, rho_cot, workspace,
                  workspace_bytes, stream, false);
}

int pm_gravity_keep(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                    const pm_gravity_desc* desc, float* acc, void* workspace, size_t workspace_bytes, void* stream) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  // the VJP's carve: the meshes stay where pm_gravity_vjp_resume finds them
  int st = grav::prepare(desc, n, workspace, workspace_bytes, true, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || acc == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "pmid/disp/acc NULL");
  st = grav::density_stage(pmid, disp, valid, n, desc, L, b, ws, s);
  if (st) return st;
  st = grav::force_mesh_stage(desc, L, b, tb, s, true, nullptr);
  if (st) return st;
  return grav::gather_stage(b.pmid_s, b.disp_s, b.valid_s, valid, n, desc, b, b.g0, b.r1, b.r2, acc, s);
}

int pm_gravity_vjp_resume(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                          const pm_gravity_desc* desc, const float* acc_cot, float* disp_cot, float* omega_cot,
                          void* workspace, size_t workspace_bytes, void* stream) {
  return vjp_impl(pmid, disp, valid, n, desc, nullptr, acc_cot, disp_cot, omega_cot, nullptr, workspace,
                  workspace_bytes, stream, true);
}

int pm_gravity_density_cot(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                           const pm_gravity_desc* desc, const float* dens_cot, float* disp_cot, void* workspace,
                           size_t workspace_bytes, void* stream) {
  hipStream_t s = as_stream(stream);
  grav::Layout L;
  grav::Buffers b{};
  Arena ws(workspace, workspace_bytes);
  fft::SpecTables tb;
  int st = grav::prepare(desc, n, workspace, workspace_bytes, false, L, b, ws, tb, s);
  if (st) return st;
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || dens_cot == nullptr || disp_cot == nullptr),
               PM_ERROR_INVALID_ARGUMENT, "pmid/disp/dens_cot/disp_cot NULL");
  if (n == 0) return PM_SUCCESS;
  PM_TRY(grav::bucket_stage(pmid, disp, valid, n, desc, b, ws, s));
  const Geom g = grav::step_geom(desc);
  const float val = desc->particle_value;
  if (b.tiled) {  // invalid particles are in no bin: their disp_cot stays as it is
    grav::Bricks k = grav::make_bricks(desc->mesh_shape, desc->brick);
    const unsigned ncols = (unsigned)((int64_t)k.nb[0] * k.nb[1]);
    if (grav::is_brick(k, 16, 16, 32))
      hipLaunchKernelGGL((grav::k_disp_cot_columns<16, 16, 32>), dim3(ncols), dim3(grav::kTileThreads), 0, s,
                         b.pmid_s, b.disp_s, b.bins, b.perm, g, k, nullptr, nullptr, nullptr, dens_cot, nullptr, n,
                         val, disp_cot, 1);
    else
      hipLaunchKernelGGL((grav::k_disp_cot_columns<0, 0, 0>), dim3(ncols), dim3(grav::kTileThreads), 0, s, b.pmid_s,
                         b.disp_s, b.bins, b.perm, g, k, nullptr, nullptr, nullptr, dens_cot, nullptr, n, val,
                         disp_cot, 1);
    PM_LAUNCHED("k_disp_cot_columns", s);
    return PM_SUCCESS;
  }
  hipLaunchKernelGGL(grav::k_scatter_bwd_sorted, dim3(grav::grid_for(n)), dim3(grav::kBlock), 0, s, b.pmid_s,
                     b.disp_s, b.valid_s, b.perm, n, g, val, dens_cot, nullptr, disp_cot);
  PM_LAUNCHED("k_scatter_bwd_sorted", s);
  return PM_SUCCESS;
}

int pm_potential_workspace(const pm_gravity_desc* desc, size_t* bytes) {
  PM_TRY(grav::check_desc(desc, 0));
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = grav::potential_workspace_bytes(desc);
  return PM_SUCCESS;
}

int pm_gravity_potential(const float* in, float shift, float scale, const float* add, const pm_gravity_desc* desc,
                         float* out, void* workspace, size_t workspace_bytes, void* stream) {
  PM_TRY(grav::check_desc(desc, 0));
  PM_TRY(require_device_tables());
  PM_RETURN_IF(in == nullptr || out == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL mesh");
  const size_t need = grav::potential_workspace_bytes(desc);
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < need, PM_ERROR_WORKSPACE,
               "potential workspace too small: need " + std::to_string(need) + " bytes");
  hipStream_t s = as_stream(stream);
  const grav::Layout L = grav::make_layout(desc);
  Arena ws(workspace, workspace_bytes);
  float* tables = ws.take<float>(fft::spec_table_floats(desc->mesh_shape));
  float2* A = ws.take<float2>(L.s1);
  float2* G = ws.take<float2>(L.s1);
  float2* B = ws.take<float2>(L.s2);
  PM_RETURN_IF(!ws.ok(), PM_ERROR_WORKSPACE, "potential workspace carve failed");
  const fft::SpecTables sc = fft::spec_scalars(desc->cell_size, desc->ptcl_spacing, desc->mask_active);
  PM_TRY(fft::launch_spec_tables(tables, desc->mesh_shape, sc, s));
  const fft::SpecTables tb = fft::spec_tables_view(tables, desc->mesh_shape, sc);
  return grav::potential_stage(L, A, B, G, tb, in, shift, scale, add, out, s);
}

int pm_spectral_workspace(const pm_gravity_desc* desc, size_t* bytes) {
  PM_RETURN_IF(desc == nullptr || bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(!fft::supported_shape(desc->mesh_shape), PM_ERROR_UNSUPPORTED,
               "mesh axes must be powers of two in [2, 2048]");
  *bytes = align_up(sizeof(float) * fft::spec_table_floats(desc->mesh_shape), 256);
  return PM_SUCCESS;
}

int pm_spectral_force_multiply(const float* in_complex, float* out_complex, const pm_gravity_desc* desc,
                               int32_t axis, int32_t adjoint, void* workspace, size_t workspace_bytes,
                               void* stream) {
  PM_RETURN_IF(desc == nullptr || in_complex == nullptr || out_complex == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(axis < -1 || axis > 2, PM_ERROR_INVALID_ARGUMENT, "axis must be -1, 0, 1 or 2");
  size_t need = 0;
  int st = pm_spectral_workspace(desc, &need);
  if (st) return st;
  PM_RETURN_IF(desc->mask_active && !(desc->ptcl_spacing > 0.f), PM_ERROR_INVALID_ARGUMENT,
               "ptcl_spacing must be positive when the particle-Nyquist mask is active");
  PM_RETURN_IF(!(desc->cell_size > 0.f), PM_ERROR_INVALID_ARGUMENT, "cell_size must be positive");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < need, PM_ERROR_WORKSPACE,
               "spectral workspace too small: need " + std::to_string(need) + " bytes");
  hipStream_t s = as_stream(stream);
  float* tables = static_cast<float*>(workspace);
  const fft::SpecTables sc = fft::spec_scalars(desc->cell_size, desc->ptcl_spacing, desc->mask_active);
  st = fft::launch_spec_tables(tables, desc->mesh_shape, sc, s);
  if (st) return st;
  fft::SpecTables tb = fft::spec_tables_view(tables, desc->mesh_shape, sc);
  return fft::launch_spectral_multiply(reinterpret_cast<const float2*>(in_complex),
                                       reinterpret_cast<float2*>(out_complex), desc->mesh_shape, tb, axis,
                                       adjoint, s);
}

}  // extern "C"
