// Periodic 3x3x3 convolutions of channels-last meshes on the matrix cores (BF16 MFMA,
// float32 accumulation): the learned potential correction's layers and their gradients.
//
// Implicit GEMM with M = output voxels, N = 16 output channels, K = 27 taps x CIN input
// channels, on v_mfma_f32_16x16x32_bf16.  A workgroup owns an 8x8x8 block of output voxels;
// its input block with a one-voxel periodic halo (10^3 x CIN bf16) and the weights are staged
// in LDS, and each wave runs 16-voxel M tiles over the K steps (32 = 32/CIN taps x CIN).
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

__device__ __forceinline__ float gelu(float x) {
  // tanh form of GELU (jax.nn.gelu's default approximate=True)
  const float c = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float u = c * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

// Stage the (kH)^3 x CIN input block of output block (bx, by, bz) (periodic wrap) in LDS as
// [x][y][z][CIN] bf16, 16-byte vectors.
template <int CIN>
__device__ __forceinline__ void stage_input(__bf16* lds, const __bf16* __restrict__ x, int n0, int n1, int n2,
                                            int ox, int oy, int oz) {
  constexpr int V = CIN / 8;  // 16-byte vectors per voxel
  constexpr int ITEMS = kH * kH * kH * V;
  for (int it = threadIdx.x; it < ITEMS; it += kThreads) {
    const int v = it % V, vox = it / V;
    const int z = vox % kH, y = (vox / kH) % kH, xx = vox / (kH * kH);
    const int gx = wrap(ox + xx - 1, n0), gy = wrap(oy + y - 1, n1), gz = wrap(oz + z - 1, n2);
    const uint4 val = *reinterpret_cast<const uint4*>(x + (((int64_t)gx * n1 + gy) * n2 + gz) * CIN + 8 * v);
    *reinterpret_cast<uint4*>(lds + vox * CIN + 8 * v) = val;
  }
}

// Voxel m (0..15) of M tile `tile`: tiles are 2 y-rows x 8 z of one x plane.
__device__ __forceinline__ void tile_voxel(int tile, int m, int& vx, int& vy, int& vz) {
  vx = tile / 4;
  vy = (tile % 4) * 2 + m / 8;
  vz = m % 8;
}

// Forward: pre = conv(x) + bias (+ res), act = gelu(pre) (or pre when !GELU), written as
// float32 pre-activations (when pre != null; the first pch of the 16 channels per voxel) and
// bf16 activations (when act != null).
template <int CIN, bool GELU>
__global__ void __launch_bounds__(kThreads) k_conv_fwd(const __bf16* __restrict__ x, const __bf16* __restrict__ w,
                                                       const float* __restrict__ bias, const __bf16* __restrict__ res,
                                                       float* __restrict__ pre, __bf16* __restrict__ act, int n0,
                                                       int n1, int n2, int pch) {
  constexpr int TPS = 32 / CIN;                       // taps per K step
  constexpr int STEPS = (kTaps + TPS - 1) / TPS;      // K steps (the last may be padded)
  __shared__ __align__(16) __bf16 xin[kH * kH * kH * CIN];
  __shared__ __align__(16) __bf16 wl[STEPS * TPS * kCout * CIN];
  const int nbz = n2 / kB, nby = n1 / kB;
  const int blk = blockIdx.x;
  const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
  const int ox = bx * kB, oy = by * kB, oz = bz * kB;
  stage_input<CIN>(xin, x, n0, n1, n2, ox, oy, oz);
  for (int i = threadIdx.x; i < STEPS * TPS * kCout * CIN; i += kThreads)
    wl[i] = i < kTaps * kCout * CIN ? w[i] : (__bf16)0.f;
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  // K group g of a step: tap s*TPS + g / (CIN/8), channels 8*(g % (CIN/8)) .. +8
  const int gt = g / (CIN / 8), gc = 8 * (g % (CIN / 8));
  for (int tile = wave; tile < kVox / 16; tile += 4) {
    int vx, vy, vz;
    tile_voxel(tile, m, vx, vy, vz);
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int tap = s * TPS + gt;
      const int t = tap < kTaps ? tap : 0;  // padded taps read a valid voxel; their weights are 0
      const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
      const int lv = ((vx + 1 + dx) * kH + (vy + 1 + dy)) * kH + (vz + 1 + dz);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(xin + lv * CIN + gc);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(wl + (tap * kCout + m) * CIN + gc);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    // lane holds output channel m for voxels 4g .. 4g+3 of the tile
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int qx, qy, qz;
      tile_voxel(tile, 4 * g + i, qx, qy, qz);
      const int64_t vox = ((int64_t)(ox + qx) * n1 + (oy + qy)) * n2 + (oz + qz);
      float z = c[i] + bias[m];
      if (res) z += (float)res[vox * kCout + m];
      if (pre && m < pch) pre[vox * pch + m] = z;
      if (act) act[vox * kCout + m] = (__bf16)(GELU ? gelu(z) : z);
    }
  }
}

// Data gradient of a 16 -> 16 layer: dx[v][ci] = sum_tap sum_co dz[v - off(tap)][co] w[tap][co][ci]
// (the same implicit GEMM with the taps mirrored and the weight matrix transposed); dz is the
// float32 gradient of the pre-activations, rounded to bf16 for the matrix cores.  When gact
// is set, the result is multiplied by gelu'(gact) (the previous layer's pre-activation) so the
// kernel returns that layer's dz directly; `add` (optional) is added first (residual path).
__global__ void __launch_bounds__(kThreads) k_conv_bwd_data(const float* __restrict__ dz, const __bf16* __restrict__ w,
                                                            const float* __restrict__ add,
                                                            const float* __restrict__ gact, float* __restrict__ dx,
                                                            int n0, int n1, int n2) {
  constexpr int CIN = 16, TPS = 2, STEPS = 14;
  __shared__ __align__(16) __bf16 gin[kH * kH * kH * kCout];
  __shared__ __align__(16) __bf16 wt[STEPS * TPS * CIN * kCout];  // [tap][ci][co]
  const int nbz = n2 / kB, nby = n1 / kB;
  const int blk = blockIdx.x;
  const int bz = blk % nbz, by = (blk / nbz) % nby, bx = blk / (nbz * nby);
  const int ox = bx * kB, oy = by * kB, oz = bz * kB;
  for (int it = threadIdx.x; it < kH * kH * kH * kCout; it += kThreads) {
    const int c = it % kCout, vox = it / kCout;
    const int z = vox % kH, y = (vox / kH) % kH, xx = vox / (kH * kH);
    const int gx = wrap(ox + xx - 1, n0), gy = wrap(oy + y - 1, n1), gz = wrap(oz + z - 1, n2);
    gin[it] = (__bf16)dz[(((int64_t)gx * n1 + gy) * n2 + gz) * kCout + c];
  }
  for (int i = threadIdx.x; i < STEPS * TPS * CIN * kCout; i += kThreads) {
    const int co = i % kCout, ci = (i / kCout) % CIN, tap = i / (kCout * CIN);
    wt[i] = tap < kTaps ? w[(tap * kCout + co) * CIN + ci] : (__bf16)0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m = lane % 16, g = lane / 16;
  const int gt = g / 2, gc = 8 * (g % 2);
  for (int tile = wave; tile < kVox / 16; tile += 4) {
    int vx, vy, vz;
    tile_voxel(tile, m, vx, vy, vz);
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int tap = s * TPS + gt;
      const int t = tap < kTaps ? tap : 0;
      const int dx_ = t / 9 - 1, dy = (t / 3) % 3 - 1, dz_ = t % 3 - 1;
      // output voxel v receives from v - off(tap)
      const int lv = ((vx + 1 - dx_) * kH + (vy + 1 - dy)) * kH + (vz + 1 - dz_);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(gin + lv * kCout + gc);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(wt + (tap * CIN + m) * kCout + gc);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int qx, qy, qz;
      tile_voxel(tile, 4 * g + i, qx, qy, qz);
      const int64_t e = (((int64_t)(ox + qx) * n1 + (oy + qy)) * n2 + (oz + qz)) * CIN + m;
      float v = c[i];
      if (add) v += add[e];
      if (gact) v *= gelu_grad(gact[e]);
      dx[e] = v;
    }
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

unsigned nblocks(const int32_t* shape) {
  return (unsigned)((int64_t)(shape[0] / conv::kB) * (shape[1] / conv::kB) * (shape[2] / conv::kB));
}

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
  const dim3 grid(nblocks(shape)), block(conv::kThreads);
  if (cin == 8 && gelu)
    hipLaunchKernelGGL((conv::k_conv_fwd<8, true>), grid, block, 0, s, xb, wb, bias, rb, pre, ab, shape[0], shape[1],
                       shape[2], (int)pre_channels);
  else if (cin == 8)
    hipLaunchKernelGGL((conv::k_conv_fwd<8, false>), grid, block, 0, s, xb, wb, bias, rb, pre, ab, shape[0], shape[1],
                       shape[2], (int)pre_channels);
  else if (gelu)
    hipLaunchKernelGGL((conv::k_conv_fwd<16, true>), grid, block, 0, s, xb, wb, bias, rb, pre, ab, shape[0],
                       shape[1], shape[2], (int)pre_channels);
  else
    hipLaunchKernelGGL((conv::k_conv_fwd<16, false>), grid, block, 0, s, xb, wb, bias, rb, pre, ab, shape[0],
                       shape[1], shape[2], (int)pre_channels);
  PM_LAUNCHED("k_conv_fwd", s);
  return PM_SUCCESS;
}

int pm_conv3d_bwd_data(const float* dz, const void* w, const float* add, const float* gact, float* dx,
                       const int32_t shape[3], void* stream) {
  PM_TRY(check_shape(shape));
  PM_RETURN_IF(dz == nullptr || w == nullptr || dx == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(conv::k_conv_bwd_data, dim3(nblocks(shape)), dim3(conv::kThreads), 0, s, dz,
                     static_cast<const __bf16*>(w), add, gact, dx, shape[0], shape[1], shape[2]);
  PM_LAUNCHED("k_conv_bwd_data", s);
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
