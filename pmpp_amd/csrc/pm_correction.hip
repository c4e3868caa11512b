// Learned mesh correction of the force step: the elementwise stages around the BF16 conv
// stack, the training loss and the optimizer.
//
// Forward, per force evaluation on a density mesh rho (float32, N cells):
//   S       = c (rho - 1),  c = 1.5 Omega_m                       (the source of the solve)
//   sigma   = sqrt(mean(S^2) + 1e-6)                              (mesh_cnn.py:79-80)
//   x0      = bf16 [S / sigma, a, Omega_m, sigma8, 0, 0, 0, 0]     (mesh_cnn.py:21-27)
// or, with the potential channel (phi the filtered potential of S, pm_gravity_potential):
//   sigma_p = sqrt(mean(phi^2) + 1e-6)                            (mesh_cnn.py:81-82)
//   x0      = bf16 [S / sigma, phi / sigma_p, a, Omega_m, sigma8, 0, 0, 0]
//   o       = conv stack(x0), channel 0 of the output layer (stored alone, n floats)
//   g       = r tanh(w . [a, Omega_m, sigma8] + b)                (the direct gain, :98-103)
//   rho_add = g (rho - 1) + r tanh(o) sigma / c                   (:97, :104 moved to the source)
// so that the solve sees S' = (1 + g) S + r tanh(o) sigma: the reference's potential residual
// g phi + r tanh(o) scale applied to the source instead (phi is linear in S), which keeps the
// correction inside the banded spectral step with no extra FFT pair.
//
// Reductions are deterministic: one float64 partial per block, summed in block order by
// the consumer.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

#include "pm_common.h"

namespace pm {
namespace corr {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;  // partial-sum slots per reduction

unsigned grid_for(int64_t n, int per_thread = 4) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, (int64_t)kBlock * per_thread), kMaxBlocks));
}

__device__ __forceinline__ double block_sum(double v, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
  __syncthreads();
  return t;  // valid on thread 0
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int nb) {
  double t = 0.0;
  for (int i = 0; i < nb; ++i) t += p[i];
  return t;
}

// ----------------------------------------------------------- features ----
// sum ((rho + shift) c)^2: float4 loads with four independent float64 chains per thread (rho
// is 16-byte aligned); the partition of the mesh over threads is fixed by the grid, so the sum
// is deterministic
__global__ void __launch_bounds__(kBlock) k_source_sumsq(const float* __restrict__ rho, int64_t n, float shift,
                                                         float c, double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const int64_t n4 = n / 4;
  const float4* r4 = reinterpret_cast<const float4*>(rho);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += (int64_t)gridDim.x * kBlock) {
    const float4 r = r4[i];
    const float s0 = (r.x + shift) * c, s1 = (r.y + shift) * c, s2 = (r.z + shift) * c, s3 = (r.w + shift) * c;
    a0 += (double)(s0 * s0);
    a1 += (double)(s1 * s1);
    a2 += (double)(s2 * s2);
    a3 += (double)(s3 * s3);
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float s = (rho[i] + shift) * c;
    a0 += (double)(s * s);
  }
  const double t = block_sum((a0 + a1) + (a2 + a3), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// stats[0] = sigma, stats[1] = gain g, stats[2] = tanh of the gain pre-activation
__global__ void k_source_scale(const double* __restrict__ partials, int nb, int64_t n, float r,
                               const float* __restrict__ lin, float a, float om, float s8,
                               float* __restrict__ stats) {
  if (threadIdx.x != 0) return;
  const double ms = sum_partials(partials, nb) / (double)n;
  const float sigma = sqrtf((float)ms + 1e-6f);
  const float th = tanhf(lin[0] * a + lin[1] * om + lin[2] * s8 + lin[3]);
  stats[0] = sigma;
  stats[1] = r * th;
  stats[2] = th;
}

// stats[3] = sigma_p = sqrt(mean phi^2 + 1e-6)
__global__ void k_potential_scale(const double* __restrict__ partials, int nb, int64_t n, float* __restrict__ stats) {
  if (threadIdx.x != 0) return;
  stats[3] = sqrtf((float)(sum_partials(partials, nb) / (double)n) + 1e-6f);
}

__global__ void __launch_bounds__(kBlock) k_features(const float* __restrict__ rho, const float* __restrict__ phi,
                                                     int64_t n, float c, const float* __restrict__ stats, float a,
                                                     float om, float s8, uint4* __restrict__ x0) {
  const float inv = 1.f / stats[0];
  const __bf16 ba = (__bf16)a, bo = (__bf16)om, bs = (__bf16)s8, z = (__bf16)0.f;
  if (phi) {  // uniform
    const float pinv = 1.f / stats[3];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
      const float s = (rho[i] - 1.f) * c;
      __bf16 v[8] = {(__bf16)(s * inv), (__bf16)(phi[i] * pinv), ba, bo, bs, z, z, z};
      x0[i] = *reinterpret_cast<const uint4*>(v);
    }
    return;
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float s = (rho[i] - 1.f) * c;
    __bf16 v[8] = {(__bf16)(s * inv), ba, bo, bs, z, z, z, z};
    x0[i] = *reinterpret_cast<const uint4*>(v);
  }
}

// ----------------------------------------------------------- residual ----
__global__ void __launch_bounds__(kBlock) k_residual(const float* __restrict__ out, const float* __restrict__ rho,
                                                     int64_t n, float c, float r, const float* __restrict__ stats,
                                                     float* __restrict__ rho_add) {
  const float g = stats[1];
  const float amp = r * stats[0] / c;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    rho_add[i] = g * (rho[i] - 1.f) + amp * tanhf(out[i]);
}

// dz (the output layer's pre-activation gradient: 16 channels with channel 0 live, or that
// channel alone when one_channel) from rho_cot, and per-block partials of
// sum(rho_cot (rho - 1)) for the gain.
__global__ void __launch_bounds__(kBlock) k_residual_bwd(const float* __restrict__ rho_cot,
                                                         const float* __restrict__ out,
                                                         const float* __restrict__ rho, int64_t n, float c, float r,
                                                         const float* __restrict__ stats, float* __restrict__ dzf,
                                                         int one_channel, double* __restrict__ partials) {
  float4* __restrict__ dz = reinterpret_cast<float4*>(dzf);
  __shared__ double lds[kBlock / 64];
  const float amp = r * stats[0] / c;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float rc = rho_cot[i];
    const float t = tanhf(out[i]);
    acc += (double)(rc * (rho[i] - 1.f));
    const float d0 = rc * amp * (1.f - t * t);
    if (one_channel) {  // uniform
      dzf[i] = d0;
      continue;
    }
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    dz[i * 4 + 0] = make_float4(d0, 0.f, 0.f, 0.f);
    dz[i * 4 + 1] = zero;
    dz[i * 4 + 2] = zero;
    dz[i * 4 + 3] = zero;
  }
  const double tot = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// d/d[w0, w1, w2, b] of the gain layer
__global__ void k_gain_bwd(const double* __restrict__ partials, int nb, float r, const float* __restrict__ stats,
                           float a, float om, float s8, float* __restrict__ dlin) {
  if (threadIdx.x != 0) return;
  const float dg = (float)sum_partials(partials, nb);
  const float th = stats[2];
  const float dp = dg * r * (1.f - th * th);
  dlin[0] = dp * a;
  dlin[1] = dp * om;
  dlin[2] = dp * s8;
  dlin[3] = dp;
}

// ------------------------------------------------------ input-path reverse ----
// The density and Omega_m cotangents of rho_add through the network's inputs.  With
// T = tanh(o), S = c (rho - 1), x0_0 = S / sigma, x0_w = Omega_m (w = 2, or 3 with the
// potential channel) and dx0 the gradient of the network input (channels 0..3):
//   sigma_bar = (r / c) sum rho_cot T - sum dx0_0 S / sigma^2
//   S_bar     = dx0_0 / sigma + sigma_bar S / (N sigma)
//   rho_bar   = g rho_cot + c S_bar
//   Om_bar    = 1.5 sum S_bar (rho - 1) + sum dx0_w - 1.5 (r sigma / c^2) sum rho_cot T
//               + r (1 - tanh_g^2) w_1 sum rho_cot (rho - 1)
// With the potential channel x0_1 = phi / sigma_p (phi linear in S, so d phi / d Omega_m =
// phi / Omega_m) also
//   phi_bar   = dx0_1 / sigma_p - phi sum dx0_1 phi / (N sigma_p^3)
//   Om_bar   += sum phi_bar phi / Omega_m
// and phi_bar's path to the density (c times the potential operator, which is symmetric)
// is added by the caller.  Pass 1 reduces the sums, pass 2 writes rho_bar (and phi_bar) and
// reduces sum S_bar (rho - 1) (and sum phi_bar phi).
__global__ void __launch_bounds__(kBlock) k_input_sums(const float* __restrict__ rho_cot,
                                                       const float* __restrict__ out,
                                                       const float* __restrict__ rho,
                                                       const float* __restrict__ phi,
                                                       const float4* __restrict__ dx0, int64_t n, float c,
                                                       double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float rc = rho_cot[i];
    const float d = rho[i] - 1.f;
    const float4 g = dx0[i];  // channels 0..3 of the input gradient
    s1 += (double)(rc * tanhf(out[i]));
    s2 += (double)(g.x * (d * c));
    s3 += (double)(phi ? g.w : g.z);
    s4 += (double)(rc * d);
    if (phi) s5 += (double)(g.y * phi[i]);
  }
  const double t1 = block_sum(s1, lds);
  const double t2 = block_sum(s2, lds);
  const double t3 = block_sum(s3, lds);
  const double t4 = block_sum(s4, lds);
  const double t5 = block_sum(s5, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t1;
    partials[kMaxBlocks + blockIdx.x] = t2;
    partials[2 * kMaxBlocks + blockIdx.x] = t3;
    partials[3 * kMaxBlocks + blockIdx.x] = t4;
    partials[4 * kMaxBlocks + blockIdx.x] = t5;
  }
}

// scal[0] = sigma_bar, scal[1] = the Omega_m terms known after pass 1, scal[2] = phi_bar's
// coefficient of phi (potential channel)
__global__ void k_input_scalars(const double* __restrict__ partials, int nb, int64_t n, float c, float r,
                                const float* __restrict__ stats, const float* __restrict__ lin, int potential,
                                float* __restrict__ scal) {
  if (threadIdx.x != 0) return;
  const double p1 = sum_partials(partials, nb);
  const double p2 = sum_partials(partials + kMaxBlocks, nb);
  const double p3 = sum_partials(partials + 2 * kMaxBlocks, nb);
  const double p4 = sum_partials(partials + 3 * kMaxBlocks, nb);
  const double sigma = stats[0];
  const double th = stats[2];
  scal[0] = (float)((double)r / c * p1 - p2 / (sigma * sigma));
  scal[1] = (float)(p3 - 1.5 * (double)r * sigma / ((double)c * c) * p1 + (double)r * (1.0 - th * th) * lin[1] * p4);
  if (potential) {
    const double sp = stats[3];
    scal[2] = (float)(sum_partials(partials + 4 * kMaxBlocks, nb) / ((double)n * sp * sp * sp));
  }
}

__global__ void __launch_bounds__(kBlock) k_input_bwd(const float* __restrict__ rho_cot,
                                                      const float* __restrict__ rho,
                                                      const float* __restrict__ phi,
                                                      const float4* __restrict__ dx0, int64_t n, float c,
                                                      const float* __restrict__ stats,
                                                      const float* __restrict__ scal, float* __restrict__ rho_bar,
                                                      float* __restrict__ phi_bar, double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  const float sigma = stats[0], g = stats[1];
  const float sb = scal[0] / ((float)n * sigma);
  const float inv = 1.f / sigma;
  const float pinv = phi ? 1.f / stats[3] : 0.f;
  const float pc = phi ? scal[2] : 0.f;
  double acc = 0.0, accp = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float d = rho[i] - 1.f;
    const float S = d * c;
    const float4 gx = dx0[i];
    const float Sbar = gx.x * inv + sb * S;
    rho_bar[i] = g * rho_cot[i] + c * Sbar;
    acc += (double)(Sbar * d);
    if (phi) {  // uniform
      const float p = phi[i];
      const float pb = gx.y * pinv - pc * p;
      phi_bar[i] = pb;
      accp += (double)(pb * p);
    }
  }
  const double t = block_sum(acc, lds);
  const double tp = block_sum(accp, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t;
    partials[kMaxBlocks + blockIdx.x] = tp;
  }
}

__global__ void k_input_omega(const double* __restrict__ partials, int nb, const float* __restrict__ scal, float om,
                              int potential, float* __restrict__ omega_cot) {
  if (threadIdx.x != 0) return;
  double t = 1.5 * sum_partials(partials, nb) + (double)scal[1];
  if (potential) t += sum_partials(partials + kMaxBlocks, nb) / (double)om;
  omega_cot[0] += (float)t;
}

// ---------------------------------------------------------------- loss ----
// loss = sum |acc - target|^2 / sum |target|^2 over 3n floats; acc_cot = d loss / d acc
__global__ void __launch_bounds__(kBlock) k_loss_partials(const float* __restrict__ acc,
                                                          const float* __restrict__ target, int64_t m,
                                                          double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  double num = 0.0, den = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBlock) {
    const float t = target[i];
    const float d = acc[i] - t;
    num += (double)d * d;
    den += (double)t * t;
  }
  const double tn = block_sum(num, lds);
  const double td = block_sum(den, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = tn;
    partials[kMaxBlocks + blockIdx.x] = td;
  }
}

__global__ void k_loss_finalize(const double* __restrict__ partials, int nb, double* __restrict__ loss) {
  if (threadIdx.x != 0) return;
  const double num = sum_partials(partials, nb);
  const double den = sum_partials(partials + kMaxBlocks, nb);
  loss[0] = den > 0.0 ? num / den : 0.0;
  loss[1] = den > 0.0 ? 1.0 / den : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_loss_grad(const float* __restrict__ acc,
                                                      const float* __restrict__ target, int64_t m,
                                                      const double* __restrict__ loss, float* __restrict__ acc_cot) {
  const float w = (float)(2.0 * loss[1]);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBlock)
    acc_cot[i] = w * (acc[i] - target[i]);
}

// ------------------------------------------------------------ optimizer ----
__global__ void __launch_bounds__(kBlock) k_sumsq(const float* __restrict__ g, int64_t n,
                                                  double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    acc += (double)g[i] * g[i];
  const double t = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Global-norm clip then Adamax (Kingma & Ba 2015, sec. 7.1, with optax's placement of eps:
// u = max(b2 u, |g| + eps), p -= lr m / (1 - b1^t) / u); the whole step is skipped when
// the gradient norm is not finite.  count (the step counter) advances in k_count.
__global__ void __launch_bounds__(kBlock) k_adamax(float* __restrict__ p, const float* __restrict__ grad,
                                                   float* __restrict__ mu, float* __restrict__ nu,
                                                   const int32_t* __restrict__ count, int64_t n,
                                                   const double* __restrict__ partials, int nb, float lr, float b1,
                                                   float b2, float eps, float clip, __bf16* __restrict__ pbf) {
  __shared__ float s_scale, s_bc;
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    const double norm = sqrt(sum_partials(partials, nb));
    s_ok = isfinite(norm) ? 1 : 0;
    s_scale = (clip > 0.f && norm >= (double)clip) ? (float)(clip / norm) : 1.f;
    const int t = count[0] + 1;
    s_bc = 1.f - powf(b1, (float)t);
  }
  __syncthreads();
  if (!s_ok) return;
  const float scale = s_scale, bc = s_bc;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float g = grad[i] * scale;
    const float m = b1 * mu[i] + (1.f - b1) * g;
    const float u = fmaxf(fabsf(g) + eps, b2 * nu[i]);
    const float v = p[i] - lr * ((m / bc) / u);
    mu[i] = m;
    nu[i] = u;
    p[i] = v;
    if (pbf) pbf[i] = (__bf16)v;
  }
}

__global__ void k_count(int32_t* __restrict__ count, const double* __restrict__ partials, int nb) {
  if (threadIdx.x != 0) return;
  if (isfinite(sum_partials(partials, nb))) count[0] = count[0] + 1;
}

__global__ void __launch_bounds__(kBlock) k_to_bf16(const float* __restrict__ src, int64_t n,
                                                    __bf16* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[i] = (__bf16)src[i];
}

}  // namespace
}  // namespace corr
}  // namespace pm

using namespace pm;

namespace {
constexpr size_t kPartialBytes = sizeof(double) * 5 * corr::kMaxBlocks + 256;

int check_mesh(int64_t n) {
  PM_RETURN_IF(n < 1, PM_ERROR_INVALID_ARGUMENT, "mesh must have at least one cell");
  return PM_SUCCESS;
}
}  // namespace

extern "C" {

int pm_correction_workspace(size_t* bytes) {
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = kPartialBytes;
  return PM_SUCCESS;
}

int pm_correction_features(const float* rho, const float* phi, int64_t n, float omega_m, float a, float sigma8,
                           float max_residual, const float* gain_params, void* x0, float* stats, void* workspace,
                           size_t workspace_bytes, void* stream) {
  PM_TRY(check_mesh(n));
  PM_RETURN_IF(rho == nullptr || stats == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(gain_params == nullptr && x0 == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "gain_params NULL (stats given) needs x0");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(x0) % 16 != 0, PM_ERROR_INVALID_ARGUMENT, "x0 must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  const float c = 1.5f * omega_m;
  if (gain_params == nullptr) {  // stats given (a slab of a mesh whose statistics are known)
    hipLaunchKernelGGL(corr::k_features, dim3(corr::grid_for(n, 8)), dim3(corr::kBlock), 0, s, rho, phi, n, c,
                       stats, a, omega_m, sigma8, static_cast<uint4*>(x0));
    PM_LAUNCHED("k_features", s);
    return PM_SUCCESS;
  }
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < kPartialBytes, PM_ERROR_WORKSPACE,
               "correction workspace too small");
  double* part = static_cast<double*>(workspace);
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(rho) % 16 != 0 || reinterpret_cast<uintptr_t>(phi) % 16 != 0,
               PM_ERROR_INVALID_ARGUMENT, "rho and phi must be 16-byte aligned");
  const unsigned nb = corr::grid_for(n, 16);
  hipLaunchKernelGGL(corr::k_source_sumsq, dim3(nb), dim3(corr::kBlock), 0, s, rho, n, -1.f, c, part);
  PM_LAUNCHED("k_source_sumsq", s);
  hipLaunchKernelGGL(corr::k_source_scale, dim3(1), dim3(64), 0, s, part, (int)nb, n, max_residual, gain_params, a,
                     omega_m, sigma8, stats);
  PM_LAUNCHED("k_source_scale", s);
  if (phi) {
    hipLaunchKernelGGL(corr::k_source_sumsq, dim3(nb), dim3(corr::kBlock), 0, s, phi, n, 0.f, 1.f, part);
    PM_LAUNCHED("k_source_sumsq", s);
    hipLaunchKernelGGL(corr::k_potential_scale, dim3(1), dim3(64), 0, s, part, (int)nb, n, stats);
    PM_LAUNCHED("k_potential_scale", s);
  }
  if (x0 == nullptr) return PM_SUCCESS;  // stats only (the fused inference stem reads rho)
  hipLaunchKernelGGL(corr::k_features, dim3(corr::grid_for(n, 8)), dim3(corr::kBlock), 0, s, rho, phi, n, c, stats,
                     a, omega_m, sigma8, static_cast<uint4*>(x0));
  PM_LAUNCHED("k_features", s);
  return PM_SUCCESS;
}

int pm_correction_residual(const float* out, const float* rho, int64_t n, float omega_m, float max_residual,
                           const float* stats, float* rho_add, void* stream) {
  PM_TRY(check_mesh(n));
  PM_RETURN_IF(out == nullptr || rho == nullptr || stats == nullptr || rho_add == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(corr::k_residual, dim3(corr::grid_for(n, 8)), dim3(corr::kBlock), 0, s, out, rho, n,
                     1.5f * omega_m, max_residual, stats, rho_add);
  PM_LAUNCHED("k_residual", s);
  return PM_SUCCESS;
}

int pm_correction_residual_bwd(const float* rho_cot, const float* out, const float* rho, int64_t n, float omega_m,
                               float a, float sigma8, float max_residual, const float* stats, float* dz_out,
                               int32_t dz_channels, float* gain_grad, void* workspace, size_t workspace_bytes,
                               void* stream) {
  PM_TRY(check_mesh(n));
  PM_RETURN_IF(dz_channels != 1 && dz_channels != 16, PM_ERROR_INVALID_ARGUMENT, "dz_channels must be 1 or 16");
  PM_RETURN_IF(rho_cot == nullptr || out == nullptr || rho == nullptr || stats == nullptr || dz_out == nullptr ||
                   gain_grad == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(dz_out) % 16 != 0, PM_ERROR_INVALID_ARGUMENT,
               "dz_out must be 16-byte aligned");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < kPartialBytes, PM_ERROR_WORKSPACE,
               "correction workspace too small");
  hipStream_t s = as_stream(stream);
  double* part = static_cast<double*>(workspace);
  const unsigned nb = corr::grid_for(n);
  hipLaunchKernelGGL(corr::k_residual_bwd, dim3(nb), dim3(corr::kBlock), 0, s, rho_cot, out, rho, n,
                     1.5f * omega_m, max_residual, stats, dz_out, dz_channels == 1 ? 1 : 0, part);
  PM_LAUNCHED("k_residual_bwd", s);
  hipLaunchKernelGGL(corr::k_gain_bwd, dim3(1), dim3(64), 0, s, part, (int)nb, max_residual, stats, a, omega_m,
                     sigma8, gain_grad);
  PM_LAUNCHED("k_gain_bwd", s);
  return PM_SUCCESS;
}

int pm_correction_input_bwd(const float* rho_cot, const float* out, const float* rho, const float* phi,
                            const float* dx0, int64_t n, float omega_m, float max_residual, const float* stats,
                            const float* gain_params, float* rho_bar, float* phi_bar, float* omega_cot,
                            void* workspace, size_t workspace_bytes, void* stream) {
  PM_TRY(check_mesh(n));
  PM_RETURN_IF(rho_cot == nullptr || out == nullptr || rho == nullptr || dx0 == nullptr || stats == nullptr ||
                   gain_params == nullptr || rho_bar == nullptr || omega_cot == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(phi != nullptr && phi_bar == nullptr, PM_ERROR_INVALID_ARGUMENT, "phi given without phi_bar");
  PM_RETURN_IF(reinterpret_cast<uintptr_t>(dx0) % 16 != 0, PM_ERROR_INVALID_ARGUMENT, "dx0 must be 16-byte aligned");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < kPartialBytes, PM_ERROR_WORKSPACE,
               "correction workspace too small");
  hipStream_t s = as_stream(stream);
  const float c = 1.5f * omega_m;
  double* part = static_cast<double*>(workspace);
  float* scal = reinterpret_cast<float*>(static_cast<char*>(workspace) + sizeof(double) * 5 * corr::kMaxBlocks);
  const unsigned nb = corr::grid_for(n);
  const float4* dx4 = reinterpret_cast<const float4*>(dx0);
  const int pot = phi != nullptr;
  hipLaunchKernelGGL(corr::k_input_sums, dim3(nb), dim3(corr::kBlock), 0, s, rho_cot, out, rho, phi, dx4, n, c,
                     part);
  PM_LAUNCHED("k_input_sums", s);
  hipLaunchKernelGGL(corr::k_input_scalars, dim3(1), dim3(64), 0, s, part, (int)nb, n, c, max_residual, stats,
                     gain_params, pot, scal);
  PM_LAUNCHED("k_input_scalars", s);
  hipLaunchKernelGGL(corr::k_input_bwd, dim3(nb), dim3(corr::kBlock), 0, s, rho_cot, rho, phi, dx4, n, c, stats, scal,
                     rho_bar, phi_bar, part);
  PM_LAUNCHED("k_input_bwd", s);
  hipLaunchKernelGGL(corr::k_input_omega, dim3(1), dim3(64), 0, s, part, (int)nb, scal, omega_m, pot, omega_cot);
  PM_LAUNCHED("k_input_omega", s);
  return PM_SUCCESS;
}

int pm_force_loss(const float* acc, const float* target, int64_t n, double* loss, float* acc_cot, void* workspace,
                  size_t workspace_bytes, void* stream) {
  PM_RETURN_IF(n < 1, PM_ERROR_INVALID_ARGUMENT, "no particles");
  PM_RETURN_IF(acc == nullptr || target == nullptr || loss == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < kPartialBytes, PM_ERROR_WORKSPACE,
               "correction workspace too small");
  hipStream_t s = as_stream(stream);
  double* part = static_cast<double*>(workspace);
  const int64_t m = 3 * n;
  const unsigned nb = corr::grid_for(m);
  hipLaunchKernelGGL(corr::k_loss_partials, dim3(nb), dim3(corr::kBlock), 0, s, acc, target, m, part);
  PM_LAUNCHED("k_loss_partials", s);
  hipLaunchKernelGGL(corr::k_loss_finalize, dim3(1), dim3(64), 0, s, part, (int)nb, loss);
  PM_LAUNCHED("k_loss_finalize", s);
  if (acc_cot) {
    hipLaunchKernelGGL(corr::k_loss_grad, dim3(corr::grid_for(m, 8)), dim3(corr::kBlock), 0, s, acc, target, m, loss,
                       acc_cot);
    PM_LAUNCHED("k_loss_grad", s);
  }
  return PM_SUCCESS;
}

int pm_adamax_step(float* params, const float* grads, float* mu, float* nu, int32_t* count, int64_t n,
                   float learning_rate, float b1, float b2, float eps, float clip_norm, void* params_bf16,
                   void* workspace, size_t workspace_bytes, void* stream) {
  PM_RETURN_IF(n < 1, PM_ERROR_INVALID_ARGUMENT, "no parameters");
  PM_RETURN_IF(params == nullptr || grads == nullptr || mu == nullptr || nu == nullptr || count == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < kPartialBytes, PM_ERROR_WORKSPACE,
               "correction workspace too small");
  hipStream_t s = as_stream(stream);
  double* part = static_cast<double*>(workspace);
  const unsigned nb = corr::grid_for(n);
  hipLaunchKernelGGL(corr::k_sumsq, dim3(nb), dim3(corr::kBlock), 0, s, grads, n, part);
  PM_LAUNCHED("k_sumsq", s);
  hipLaunchKernelGGL(corr::k_adamax, dim3(corr::grid_for(n)), dim3(corr::kBlock), 0, s, params, grads, mu, nu, count,
                     n, part, (int)nb, learning_rate, b1, b2, eps, clip_norm, static_cast<__bf16*>(params_bf16));
  PM_LAUNCHED("k_adamax", s);
  hipLaunchKernelGGL(corr::k_count, dim3(1), dim3(64), 0, s, count, part, (int)nb);
  PM_LAUNCHED("k_count", s);
  return PM_SUCCESS;
}

int pm_to_bf16(const float* src, int64_t n, void* dst, void* stream) {
  PM_RETURN_IF(n < 0 || (n > 0 && (src == nullptr || dst == nullptr)), PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  if (n == 0) return PM_SUCCESS;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(corr::k_to_bf16, dim3(corr::grid_for(n)), dim3(corr::kBlock), 0, s, src, n,
                     static_cast<__bf16*>(dst));
  PM_LAUNCHED("k_to_bf16", s);
  return PM_SUCCESS;
}

}  // extern "C"
