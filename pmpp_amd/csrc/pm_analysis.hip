// Integer-shell power spectra of half-spectrum meshes (the P(k) metric).
//
// For every stored mode (kx, ky, kz <= Nz/2) of an rfft spectrum: shell = floor(|n| + 1e-9)
// of the integer mode vector, Hermitian multiplicity 2 (1 on the kz = 0 and even-Nz Nyquist
// planes), power |a|^2 (and |b|^2, Re a b*) times the mass-assignment deconvolution
// (sinc(nx/Nx) sinc(ny/Ny) sinc(nz/Nz))^(-2p), accumulated per shell.  Sums are float64:
// per-block LDS accumulators, then one global atomic per shell and block, so the summation
// order (not the value beyond float64 rounding) varies between runs.
#include <algorithm>

#include "pm_common.h"

namespace pm {
namespace analysis {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxShells = 2048;

__device__ __forceinline__ float sinc_f(float x) {
  if (x == 0.f) return 1.f;
  const float px = 3.14159265358979323846f * x;
  return sinpif(x) / px;
}

__global__ void k_zero_f64(double* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

template <bool CROSS>
__global__ void __launch_bounds__(kBlock) k_shell_sums(const float2* __restrict__ a, const float2* __restrict__ b,
                                                       int nx, int ny, int nz, int mas_power, int num_shells,
                                                       int density, double* __restrict__ sums) {
  __shared__ double acc[3][kMaxShells];
  const int nw = CROSS ? 3 : 1;
  for (int i = threadIdx.x; i < nw * num_shells; i += kBlock) acc[i / num_shells][i % num_shells] = 0.0;
  __syncthreads();
  const int nzh = nz / 2 + 1;
  const int64_t total = (int64_t)nx * ny * nzh;
  const int64_t nm = (int64_t)nx * ny * nz;
  // density inputs: delta_k = rho_k / mean(rho) away from k = 0, mean = Re rho_0 / N_m
  float sa = 1.f, sb = 1.f;
  if (density) {
    sa = (float)((double)nm / (double)a[0].x);
    if (CROSS) sb = (float)((double)nm / (double)b[0].x);
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int kz = (int)(i % nzh);
    const int ky = (int)((i / nzh) % ny);
    const int kx = (int)(i / ((int64_t)nzh * ny));
    const int mx = kx < (nx + 1) / 2 ? kx : kx - nx;
    const int my = ky < (ny + 1) / 2 ? ky : ky - ny;
    const int mz = kz;
    const int64_t msq = (int64_t)mx * mx + (int64_t)my * my + (int64_t)mz * mz;
    int shell = (int)floor(sqrt((double)msq) + 1e-9);
    shell = shell < num_shells ? shell : num_shells - 1;
    const float mult = (kz == 0 || ((nz % 2) == 0 && kz == nz / 2)) ? 1.f : 2.f;
    float w = 1.f;
    if (mas_power > 0) {
      const float win = sinc_f((float)mx / (float)nx) * sinc_f((float)my / (float)ny) * sinc_f((float)mz / (float)nz);
      const float inv = 1.f / win;
      float inv2 = inv * inv;
      w = 1.f;
      for (int p = 0; p < mas_power; ++p) w *= inv2;
    }
    float2 va = a[i];
    va = make_float2(va.x * sa, va.y * sa);
    atomicAdd(&acc[0][shell], (double)((va.x * va.x + va.y * va.y) * w * mult));
    if (CROSS) {
      float2 vb = b[i];
      vb = make_float2(vb.x * sb, vb.y * sb);
      atomicAdd(&acc[1][shell], (double)((vb.x * vb.x + vb.y * vb.y) * w * mult));
      atomicAdd(&acc[2][shell], (double)((va.x * vb.x + va.y * vb.y) * w * mult));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nw * num_shells; i += kBlock) {
    const double v = acc[i / num_shells][i % num_shells];
    if (v != 0.0) atomicAdd(&sums[i], v);
  }
}

// Reverse of the shell sums: grad_k = coef[shell(k)] * window(k) * s^2 * a_k, s the density
// normalisation (1 for an overdensity), shell and window exactly as the forward.  With
// coef[s] = 2 cell^3 dL/dP_s / nmodes_s, irfftn(grad) is dL/d(mesh) of L(P) (the constant
// term from the mean of a density input is left out: it carries no particle gradient).
__global__ void __launch_bounds__(kBlock) k_shell_grad(const float2* __restrict__ a, float2* __restrict__ g, int nx,
                                                       int ny, int nz, int mas_power, int num_shells,
                                                       const float* __restrict__ coef, int density) {
  const int nzh = nz / 2 + 1;
  const int64_t total = (int64_t)nx * ny * nzh;
  const int64_t nm = (int64_t)nx * ny * nz;
  const float s = density ? (float)((double)nm / (double)a[0].x) : 1.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int kz = (int)(i % nzh);
    const int ky = (int)((i / nzh) % ny);
    const int kx = (int)(i / ((int64_t)nzh * ny));
    const int mx = kx < (nx + 1) / 2 ? kx : kx - nx;
    const int my = ky < (ny + 1) / 2 ? ky : ky - ny;
    const int mz = kz;
    const int64_t msq = (int64_t)mx * mx + (int64_t)my * my + (int64_t)mz * mz;
    int shell = (int)floor(sqrt((double)msq) + 1e-9);
    shell = shell < num_shells ? shell : num_shells - 1;
    float w = 1.f;
    if (mas_power > 0) {
      const float win = sinc_f((float)mx / (float)nx) * sinc_f((float)my / (float)ny) * sinc_f((float)mz / (float)nz);
      const float inv = 1.f / win;
      const float inv2 = inv * inv;
      for (int p = 0; p < mas_power; ++p) w *= inv2;
    }
    const float f = coef[shell] * w * s * s;
    const float2 v = a[i];
    g[i] = make_float2(f * v.x, f * v.y);
  }
}

}  // namespace
}  // namespace analysis
}  // namespace pm

using namespace pm;

extern "C" {

int pm_power_spectrum(const float* spec_a, const float* spec_b, const int32_t shape[3], int32_t mas_power,
                      int32_t num_shells, int32_t density, double* sums, void* stream) {
  PM_RETURN_IF(spec_a == nullptr || shape == nullptr || sums == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "NULL argument");
  for (int a = 0; a < 3; ++a) PM_RETURN_IF(shape[a] < 1, PM_ERROR_INVALID_ARGUMENT, "mesh axes must be >= 1");
  PM_RETURN_IF(num_shells < 1 || num_shells > analysis::kMaxShells, PM_ERROR_UNSUPPORTED,
               "num_shells must be in [1, 2048]");
  PM_RETURN_IF(mas_power < 0 || mas_power > 4, PM_ERROR_INVALID_ARGUMENT, "mas_power must be in [0, 4]");
  hipStream_t s = as_stream(stream);
  const bool cross = spec_b != nullptr;
  const int nw = cross ? 3 : 1;
  hipLaunchKernelGGL(analysis::k_zero_f64, dim3(1), dim3(256), 0, s, sums, nw * num_shells);
  PM_LAUNCHED("k_zero_f64", s);
  const int64_t total = (int64_t)shape[0] * shape[1] * (shape[2] / 2 + 1);
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, analysis::kBlock * 16), 2048));
  const float2* a = reinterpret_cast<const float2*>(spec_a);
  const float2* b = reinterpret_cast<const float2*>(spec_b);
  if (cross)
    hipLaunchKernelGGL(analysis::k_shell_sums<true>, dim3(blocks), dim3(analysis::kBlock), 0, s, a, b, shape[0],
                       shape[1], shape[2], mas_power, num_shells, density, sums);
  else
    hipLaunchKernelGGL(analysis::k_shell_sums<false>, dim3(blocks), dim3(analysis::kBlock), 0, s, a, b, shape[0],
                       shape[1], shape[2], mas_power, num_shells, density, sums);
  PM_LAUNCHED("k_shell_sums", s);
  return PM_SUCCESS;
}

int pm_power_spectrum_bwd(const float* spec, const int32_t shape[3], int32_t mas_power, int32_t num_shells,
                          int32_t density, const float* shell_coef, float* grad_spec, void* stream) {
  PM_RETURN_IF(spec == nullptr || shape == nullptr || shell_coef == nullptr || grad_spec == nullptr,
               PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  for (int a = 0; a < 3; ++a) PM_RETURN_IF(shape[a] < 1, PM_ERROR_INVALID_ARGUMENT, "mesh axes must be >= 1");
  PM_RETURN_IF(num_shells < 1 || num_shells > analysis::kMaxShells, PM_ERROR_UNSUPPORTED,
               "num_shells must be in [1, 2048]");
  PM_RETURN_IF(mas_power < 0 || mas_power > 4, PM_ERROR_INVALID_ARGUMENT, "mas_power must be in [0, 4]");
  hipStream_t s = as_stream(stream);
  const int64_t total = (int64_t)shape[0] * shape[1] * (shape[2] / 2 + 1);
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, analysis::kBlock * 4), 65536));
  hipLaunchKernelGGL(analysis::k_shell_grad, dim3(blocks), dim3(analysis::kBlock), 0, s,
                     reinterpret_cast<const float2*>(spec), reinterpret_cast<float2*>(grad_spec), shape[0], shape[1],
                     shape[2], mas_power, num_shells, shell_coef, density);
  PM_LAUNCHED("k_shell_grad", s);
  return PM_SUCCESS;
}

}  // extern "C"
