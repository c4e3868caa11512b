// Kick / drift updates of the N-body integrator and their adjoints (elementwise over the
// (n, 3) float32 particle arrays; HBM-bound, float4 vectorised with a scalar tail).
//
// The adjoints return the scalar the cosmology cotangent of the step factor is scaled by
// (the dot products of nbody/integrator.py:572-573 and :472), reduced deterministically:
// fixed per-block partial sums in float64, then one block sums the partials in order.
#include <algorithm>

#include "pm_common.h"

namespace pm {
namespace nbody {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;

inline unsigned blocks_for(int64_t n4) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n4, kBlock), kMaxBlocks));
}

// y += f * x over m floats (y, x 16-byte aligned)
// out = y + f x (out may be y: in place; one pass either way)
__global__ void __launch_bounds__(kBlock) k_axpy(const float* y, const float* __restrict__ x, float* out, int64_t m,
                                                 float f) {
  const int64_t m4 = m / 4;
  const float4* y4 = reinterpret_cast<const float4*>(y);
  const float4* x4 = reinterpret_cast<const float4*>(x);
  float4* o4 = reinterpret_cast<float4*>(out);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m4; i += (int64_t)gridDim.x * blockDim.x) {
    float4 a = y4[i], b = x4[i];
    o4[i] = make_float4(a.x + b.x * f, a.y + b.y * f, a.z + b.z * f, a.w + b.w * f);
  }
  for (int64_t i = 4 * m4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = y[i] + x[i] * f;
}

__device__ __forceinline__ double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __shared__ double part[kBlock / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
  return t;
}

// Reverse of an update y_out = y_in + f * x with cotangent c of y_out:
//   y_in = y_out - f * x        (state reconstruction)
//   g   += f * c                (cotangent of x; g may be null)
//   partial[block] = sum c * x  (cotangent of f, before the cosmology chain rule)
__global__ void __launch_bounds__(kBlock) k_update_adj(float* __restrict__ y, const float* __restrict__ x,
                                                       const float* __restrict__ c, float* __restrict__ g,
                                                       int64_t m, float f, double* __restrict__ partial) {
  const int64_t m4 = m / 4;
  double acc = 0.0;
  float4* y4 = reinterpret_cast<float4*>(y);
  const float4* x4 = reinterpret_cast<const float4*>(x);
  const float4* c4 = reinterpret_cast<const float4*>(c);
  float4* g4 = reinterpret_cast<float4*>(g);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m4; i += (int64_t)gridDim.x * blockDim.x) {
    float4 a = y4[i], b = x4[i], cc = c4[i];
    y4[i] = make_float4(a.x - b.x * f, a.y - b.y * f, a.z - b.z * f, a.w - b.w * f);
    if (g) {
      float4 gg = g4[i];
      g4[i] = make_float4(gg.x + cc.x * f, gg.y + cc.y * f, gg.z + cc.z * f, gg.w + cc.w * f);
    }
    acc += (double)(cc.x * b.x) + (double)(cc.y * b.y) + (double)(cc.z * b.z) + (double)(cc.w * b.w);
  }
  for (int64_t i = 4 * m4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    y[i] = y[i] - x[i] * f;
    if (g) g[i] = g[i] + c[i] * f;
    acc += (double)(c[i] * x[i]);
  }
  double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock) k_sum_partials(const double* __restrict__ partial, int nb,
                                                         double* __restrict__ out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) v += partial[i];
  double t = block_sum(v);
  if (threadIdx.x == 0) *out = t;
}

int check_vec(const void* p, const char* what) {
  PM_RETURN_IF(p == nullptr, PM_ERROR_INVALID_ARGUMENT, std::string(what) + " is NULL");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(p) % 16 != 0, PM_ERROR_INVALID_ARGUMENT,
               std::string(what) + " must be 16-byte aligned");
  return PM_SUCCESS;
}

int update(const float* y, const float* x, float* out, int64_t n, float f, void* stream, const char* name) {
  PM_RETURN_IF(n < 0, PM_ERROR_INVALID_ARGUMENT, "n must be >= 0");
  if (n == 0) return PM_SUCCESS;
  PM_TRY(check_vec(y, "state"));
  PM_TRY(check_vec(x, "input"));
  PM_TRY(check_vec(out, "output"));
  hipStream_t s = as_stream(stream);
  const int64_t m = 3 * n;
  hipLaunchKernelGGL(k_axpy, dim3(blocks_for(m / 4 + 1)), dim3(kBlock), 0, s, y, x, out, m, f);
  PM_LAUNCHED(name, s);
  return PM_SUCCESS;
}

int update_adj(float* y, const float* x, const float* c, float* g, int64_t n, float f, double* dot, void* ws,
               size_t ws_bytes, void* stream, const char* name) {
  PM_RETURN_IF(n < 0, PM_ERROR_INVALID_ARGUMENT, "n must be >= 0");
  PM_RETURN_IF(dot == nullptr, PM_ERROR_INVALID_ARGUMENT, "dot is NULL");
  PM_RETURN_IF(ws == nullptr || ws_bytes < sizeof(double) * kMaxBlocks, PM_ERROR_WORKSPACE,
               "adjoint workspace too small");
  PM_TRY(check_vec(y, "state"));
  PM_TRY(check_vec(x, "input"));
  PM_TRY(check_vec(c, "cotangent"));
  if (g) PM_TRY(check_vec(g, "input cotangent"));
  hipStream_t s = as_stream(stream);
  const int64_t m = 3 * n;
  const unsigned nb = blocks_for(m / 4 + 1);
  double* partial = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_update_adj, dim3(nb), dim3(kBlock), 0, s, y, x, c, g, m, f, partial);
  PM_LAUNCHED(name, s);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partial, (int)nb, dot);
  PM_LAUNCHED("k_sum_partials", s);
  return PM_SUCCESS;
}

}  // namespace
}  // namespace nbody
}  // namespace pm

using namespace pm;

extern "C" {

int pm_kick(const float* vel, const float* acc, float* vel_out, int64_t n, float factor, void* stream) {
  return nbody::update(vel, acc, vel_out, n, factor, stream, "k_kick");
}

int pm_drift(const float* disp, const float* vel, float* disp_out, int64_t n, float factor, void* stream) {
  return nbody::update(disp, vel, disp_out, n, factor, stream, "k_drift");
}

int pm_nbody_adj_workspace(size_t* bytes) {
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = sizeof(double) * nbody::kMaxBlocks;
  return PM_SUCCESS;
}

int pm_kick_adj(float* vel, const float* acc, const float* vel_cot, float* acc_cot, int64_t n, float factor,
                double* dot, void* workspace, size_t workspace_bytes, void* stream) {
  return nbody::update_adj(vel, acc, vel_cot, acc_cot, n, factor, dot, workspace, workspace_bytes, stream,
                           "k_kick_adj");
}

int pm_drift_adj(float* disp, const float* vel, const float* disp_cot, float* vel_cot, int64_t n, float factor,
                 double* dot, void* workspace, size_t workspace_bytes, void* stream) {
  return nbody::update_adj(disp, vel, disp_cot, vel_cot, n, factor, dot, workspace, workspace_bytes, stream,
                           "k_drift_adj");
}

}  // extern "C"
