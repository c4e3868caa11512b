// Initial conditions on the particle grid: the linear-mode filter and the spectral and
// real-space steps of 2LPT.  Spectra are the standard rfftn layout (n0, n1, n2/2 + 1) complex
// of pm_rfft3d; wavenumbers come from per-axis float32 tables (the angular fftfreq of the
// particle grid), so every mode sees exactly the same k as the reference's float32 kvec.
#include <algorithm>
#include <cmath>

#include "pm_common.h"

namespace pm {
namespace ic {
namespace {

constexpr int kBlock = 256;

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, (int64_t)kBlock * 4), 65536));
}

struct KTabs {
  const float* k0;
  const float* k1;
  const float* k2;
  int n0, n1, nzh;
};

__device__ __forceinline__ void mode_k(const KTabs& t, int64_t i, float& a, float& b, float& c) {
  const int kz = (int)(i % t.nzh);
  const int64_t r = i / t.nzh;
  const int ky = (int)(r % t.n1);
  const int kx = (int)(r / t.n1);
  a = t.k0[kx];
  b = t.k1[ky];
  c = t.k2[kz];
}

// linear interpolation on an ascending table, clamped at both ends (numpy/jax interp)
__device__ __forceinline__ float interp(float x, const float* __restrict__ xs, const float* __restrict__ ys, int n) {
  if (x <= xs[0]) return ys[0];
  if (x >= xs[n - 1]) return ys[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xs[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  const float t = (x - xs[lo]) / (xs[hi] - xs[lo]);
  return ys[lo] + t * (ys[hi] - ys[lo]);
}

// spec *= scale * sqrt(P(k)),  P(k) = coef (k / k_pivot)^n_s T(k)^2, P(0) = 0
__global__ void __launch_bounds__(kBlock) k_linear_modes(float2* __restrict__ spec, KTabs t,
                                                         const float* __restrict__ tk,
                                                         const float* __restrict__ tT, int nt, float coef,
                                                         float k_pivot, float n_s, float scale) {
  const int64_t total = (int64_t)t.n0 * t.n1 * t.nzh;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
    float a, b, c;
    mode_k(t, i, a, b, c);
    const float k = sqrtf(a * a + b * b + c * c);
    float f = 0.f;
    if (k > 0.f) {
      const float T = interp(k, tk, tT, nt);
      const float p = coef * powf(k / k_pivot, n_s) * T * T;
      f = scale * sqrtf(fmaxf(p, 0.f));
    }
    const float2 v = spec[i];
    spec[i] = make_float2(v.x * f, v.y * f);
  }
}

enum Op : int { kScale = 0, kLaplace = 1, kStrain = 2, kNegGrad = 3 };

__device__ __forceinline__ float pick(int axis, float a, float b, float c) {
  return axis == 0 ? a : (axis == 1 ? b : c);
}

// out = symbol(k) * in for one of: s (kScale), -1/k^2 with k = 0 -> 0 (kLaplace),
// -k_i k_j with the Nyquist k zeroed when i != j (kStrain), -i k_i with the Nyquist k
// zeroed (kNegGrad).  nyq[a] = pi / spacing_a, |(|k| - nyq)| <= tol[a] marks the Nyquist.
__global__ void __launch_bounds__(kBlock) k_spectral_op(const float2* __restrict__ in, float2* __restrict__ out,
                                                        KTabs t, int op, int ai, int aj, float s, float3 nyq,
                                                        float3 tol) {
  const int64_t total = (int64_t)t.n0 * t.n1 * t.nzh;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
    float a, b, c;
    mode_k(t, i, a, b, c);
    const float2 v = in[i];
    float2 r;
    if (op == kScale) {
      r = make_float2(v.x * s, v.y * s);
    } else if (op == kLaplace) {
      const float k2 = a * a + b * b + c * c;
      r = k2 != 0.f ? make_float2(-v.x / k2, -v.y / k2) : make_float2(0.f, 0.f);
    } else {
      float ki = pick(ai, a, b, c);
      const float nqi = pick(ai, nyq.x, nyq.y, nyq.z), tli = pick(ai, tol.x, tol.y, tol.z);
      if (op == kNegGrad) {
        if (fabsf(fabsf(ki) - nqi) <= tli) ki = 0.f;
        r = make_float2(ki * v.y, -ki * v.x);  // -i k v
      } else {
        float kj = pick(aj, a, b, c);
        if (ai != aj) {
          const float nqj = pick(aj, nyq.x, nyq.y, nyq.z), tlj = pick(aj, tol.x, tol.y, tol.z);
          if (fabsf(fabsf(ki) - nqi) <= tli) ki = 0.f;
          if (fabsf(fabsf(kj) - nqj) <= tlj) kj = 0.f;
        }
        const float m = -ki * kj;
        r = make_float2(m * v.x, m * v.y);
      }
    }
    out[i] = r;
  }
}

// acc = (init ? 0 : acc) + coef * a * b  (b == null: + coef * a)
__global__ void __launch_bounds__(kBlock) k_product_acc(float* __restrict__ acc, const float* __restrict__ a,
                                                        const float* __restrict__ b, int64_t n, float coef,
                                                        int init) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float p = b ? a[i] * b[i] : a[i];
    acc[i] = init ? coef * p : acc[i] + coef * p;
  }
}

// disp[:, axis] += D g, vel[:, axis] += V g for particles in grid order
__global__ void __launch_bounds__(kBlock) k_lpt_acc(float* __restrict__ disp, float* __restrict__ vel,
                                                    const float* __restrict__ g, int64_t n, int axis, float D,
                                                    float V) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float v = g[i];
    disp[3 * i + axis] += D * v;
    if (vel) vel[3 * i + axis] += V * v;
  }
}

}  // namespace
}  // namespace ic
}  // namespace pm

using namespace pm;

namespace {

int check_ktabs(const int32_t* shape, const float* const* ktab, ic::KTabs& t) {
  PM_RETURN_IF(shape == nullptr || ktab == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  for (int a = 0; a < 3; ++a) {
    PM_RETURN_IF(shape[a] < 1, PM_ERROR_INVALID_ARGUMENT, "grid axes must be >= 1");
    PM_RETURN_IF(ktab[a] == nullptr, PM_ERROR_INVALID_ARGUMENT, "wavenumber table NULL");
  }
  t.k0 = ktab[0];
  t.k1 = ktab[1];
  t.k2 = ktab[2];
  t.n0 = shape[0];
  t.n1 = shape[1];
  t.nzh = shape[2] / 2 + 1;
  return PM_SUCCESS;
}

}  // namespace

extern "C" {

int pm_ic_linear_modes(float* spec, const int32_t shape[3], const float* const ktab[3], const float* table_k,
                       const float* table_T, int32_t table_n, float coef, float k_pivot, float n_s, float scale,
                       void* stream) {
  ic::KTabs t{};
  PM_TRY(check_ktabs(shape, ktab, t));
  PM_RETURN_IF(spec == nullptr || table_k == nullptr || table_T == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "NULL argument");
  PM_RETURN_IF(table_n < 2, PM_ERROR_INVALID_ARGUMENT, "transfer table needs >= 2 entries");
  const int64_t total = (int64_t)t.n0 * t.n1 * t.nzh;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(ic::k_linear_modes, dim3(ic::grid_for(total)), dim3(ic::kBlock), 0, s,
                     reinterpret_cast<float2*>(spec), t, table_k, table_T, table_n, coef, k_pivot, n_s, scale);
  PM_LAUNCHED("k_linear_modes", s);
  return PM_SUCCESS;
}

int pm_ic_spectral_op(const float* in, float* out, const int32_t shape[3], const float* const ktab[3], int32_t op,
                      int32_t axis_i, int32_t axis_j, float scale, const float nyquist[3], const float tol[3],
                      void* stream) {
  ic::KTabs t{};
  PM_TRY(check_ktabs(shape, ktab, t));
  PM_RETURN_IF(in == nullptr || out == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(op < 0 || op > 3, PM_ERROR_INVALID_ARGUMENT, "unknown spectral op");
  PM_RETURN_IF((op >= 2) && (axis_i < 0 || axis_i > 2 || axis_j < 0 || axis_j > 2), PM_ERROR_INVALID_ARGUMENT,
               "axis out of range");
  PM_RETURN_IF(op >= 2 && (nyquist == nullptr || tol == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "Nyquist tables NULL");
  float3 nq = make_float3(0.f, 0.f, 0.f), tl = make_float3(-1.f, -1.f, -1.f);
  if (op >= 2) {
    nq = make_float3(nyquist[0], nyquist[1], nyquist[2]);
    tl = make_float3(tol[0], tol[1], tol[2]);
  }
  const int64_t total = (int64_t)t.n0 * t.n1 * t.nzh;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(ic::k_spectral_op, dim3(ic::grid_for(total)), dim3(ic::kBlock), 0, s,
                     reinterpret_cast<const float2*>(in), reinterpret_cast<float2*>(out), t, (int)op, (int)axis_i,
                     (int)axis_j, scale, nq, tl);
  PM_LAUNCHED("k_spectral_op", s);
  return PM_SUCCESS;
}

int pm_ic_product_acc(float* acc, const float* a, const float* b, int64_t n, float coef, int32_t init,
                      void* stream) {
  PM_RETURN_IF(n < 0 || (n > 0 && (acc == nullptr || a == nullptr)), PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  if (n == 0) return PM_SUCCESS;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(ic::k_product_acc, dim3(ic::grid_for(n)), dim3(ic::kBlock), 0, s, acc, a, b, n, coef, (int)init);
  PM_LAUNCHED("k_product_acc", s);
  return PM_SUCCESS;
}

int pm_ic_lpt_accumulate(float* disp, float* vel, const float* grad, int64_t n, int32_t axis, float D, float V,
                         void* stream) {
  PM_RETURN_IF(n < 0 || (n > 0 && (disp == nullptr || grad == nullptr)), PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(axis < 0 || axis > 2, PM_ERROR_INVALID_ARGUMENT, "axis out of range");
  if (n == 0) return PM_SUCCESS;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(ic::k_lpt_acc, dim3(ic::grid_for(n)), dim3(ic::kBlock), 0, s, disp, vel, grad, n, (int)axis, D,
                     V);
  PM_LAUNCHED("k_lpt_acc", s);
  return PM_SUCCESS;
}

}  // extern "C"
