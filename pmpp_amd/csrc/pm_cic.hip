// CIC mass assignment (scatter), interpolation (gather) and their hand-written adjoints.
//
// One thread per particle in caller order, all float32.  These are the general kernel-seam
// entry points (the reference's pallas_scatter/pallas_gather/*_bwd contract, with offset,
// explicit mesh cell size, validity mask, channels and local/global shapes).  The gravity
// step uses its own sorted-order kernels (pm_gravity.hip) built on the same device geometry.
#include <mutex>
#include <vector>

#include "pm_common.h"
#include "pm_sort.h"

namespace pm {

static thread_local std::string g_last_error;

struct Timing {
  std::vector<hipEvent_t> ev;
  std::vector<const char*> names;
  size_t used = 0;
  bool on = false;
};
static thread_local Timing g_timing;

void timing_mark(const char* name, hipStream_t s) {
  Timing& t = g_timing;
  if (!t.on || t.used >= t.ev.size()) return;
  if (hipEventRecord(t.ev[t.used], s) == hipSuccess) t.names[t.used++] = name;
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_scatter(const int16_t* __restrict__ pmid,
                                                    const float* __restrict__ disp,
                                                    const uint8_t* __restrict__ valid,
                                                    const float* __restrict__ val, float val_scalar,
                                                    int64_t n, int C, Geom g, float* __restrict__ mesh) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (valid && !valid[p]) return;
  AxisCic ax[3];
  particle_cic(g, pmid, disp, p, ax);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int64_t flat;
    if (!corner_index(g, ax, c, flat)) continue;
    float w = corner_weight(ax, c);
    for (int ch = 0; ch < C; ++ch) {
      float v = val ? val[p * C + ch] : val_scalar;
      atomic_add_f32(mesh + flat * C + ch, v * w);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_gather(const int16_t* __restrict__ pmid,
                                                   const float* __restrict__ disp,
                                                   const uint8_t* __restrict__ valid, int64_t n,
                                                   int C, Geom g, const float* __restrict__ mesh,
                                                   float* __restrict__ out) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  bool ok_p = !valid || valid[p];
  AxisCic ax[3];
  particle_cic(g, pmid, disp, p, ax);
  for (int ch = 0; ch < C; ++ch) {
    float acc = 0.f;
    if (ok_p) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        int64_t flat;
        if (!corner_index(g, ax, c, flat)) continue;
        acc = acc + mesh[flat * C + ch] * corner_weight(ax, c);
      }
    }
    out[p * C + ch] = acc;
  }
}

__device__ __forceinline__ float cell_scale(const Geom& g) { return g.a2 > 0.f ? g.a2 : g.a1; }

__global__ void __launch_bounds__(kBlock) k_gather_bwd(const int16_t* __restrict__ pmid,
                                                       const float* __restrict__ disp,
                                                       const uint8_t* __restrict__ valid, int64_t n,
                                                       int C, Geom g, const float* __restrict__ mesh,
                                                       const float* __restrict__ val_cot,
                                                       float* __restrict__ disp_cot,
                                                       float* __restrict__ mesh_cot) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float dc[3] = {0.f, 0.f, 0.f};
  if (!valid || valid[p]) {
    AxisCic ax[3];
    particle_cic(g, pmid, disp, p, ax);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      if (!corner_index(g, ax, c, flat)) continue;
      float w = corner_weight(ax, c);
      float gr[3];
      corner_grad(ax, c, gr);
      for (int ch = 0; ch < C; ++ch) {
        float vc = val_cot[p * C + ch];
        float mv = mesh[flat * C + ch];
        atomic_add_f32(mesh_cot + flat * C + ch, vc * w);
#pragma unroll
        for (int i = 0; i < 3; ++i) dc[i] = dc[i] + vc * mv * gr[i];
      }
    }
  }
  float s = cell_scale(g);
#pragma unroll
  for (int i = 0; i < 3; ++i) disp_cot[p * 3 + i] = dc[i] / s;
}

__global__ void __launch_bounds__(kBlock) k_scatter_bwd(const int16_t* __restrict__ pmid,
                                                        const float* __restrict__ disp,
                                                        const uint8_t* __restrict__ valid,
                                                        const float* __restrict__ val,
                                                        float val_scalar, int64_t n, int C, Geom g,
                                                        const float* __restrict__ mesh_cot,
                                                        float* __restrict__ disp_cot,
                                                        float* __restrict__ val_cot) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool in_range = p < n;
  bool ok_p = in_range && (!valid || valid[p]);
  float dc[3] = {0.f, 0.f, 0.f};
  float scalar_sum = 0.f;
  if (ok_p) {
    AxisCic ax[3];
    particle_cic(g, pmid, disp, p, ax);
    for (int ch = 0; ch < (val ? C : 0); ++ch) val_cot[p * C + ch] = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      if (!corner_index(g, ax, c, flat)) continue;
      float w = corner_weight(ax, c);
      float gr[3];
      corner_grad(ax, c, gr);
      float chsum = 0.f;
      for (int ch = 0; ch < C; ++ch) {
        float mv = mesh_cot[flat * C + ch];
        float pv = val ? val[p * C + ch] : val_scalar;
        float weighted = mv * w;
        if (val)
          val_cot[p * C + ch] += weighted;
        else
          scalar_sum += weighted;
        chsum = chsum + mv * pv;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) dc[i] = dc[i] + chsum * gr[i];
    }
  } else if (in_range && val) {
    for (int ch = 0; ch < C; ++ch) val_cot[p * C + ch] = 0.f;
  }
  if (in_range) {
    float s = cell_scale(g);
#pragma unroll
    for (int i = 0; i < 3; ++i) disp_cot[p * 3 + i] = dc[i] / s;
  }
  if (!val) {
    // one atomic per wave for the scalar-value cotangent
    for (int off = 32; off > 0; off >>= 1) scalar_sum += __shfl_xor(scalar_sum, off);
    if ((threadIdx.x & 63) == 0 && scalar_sum != 0.f) atomic_add_f32(val_cot, scalar_sum);
  }
}

int check_cic_args(const int16_t* pmid, const float* disp, int64_t n, int32_t C,
                   const pm_cic_geom* geom) {
  PM_RETURN_IF(geom == nullptr, PM_ERROR_INVALID_ARGUMENT, "geom is NULL");
  PM_RETURN_IF(n < 0, PM_ERROR_INVALID_ARGUMENT, "negative particle count");
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "pmid/disp NULL");
  PM_RETURN_IF(C < 1, PM_ERROR_INVALID_ARGUMENT, "channels must be >= 1");
  for (int i = 0; i < 3; ++i) {
    PM_RETURN_IF(geom->local_shape[i] < 1 || geom->global_shape[i] < 1,
                 PM_ERROR_INVALID_ARGUMENT, "mesh shapes must be positive");
  }
  PM_RETURN_IF(!(geom->particle_cell > 0.f), PM_ERROR_INVALID_ARGUMENT,
               "particle_cell must be positive");
  return PM_SUCCESS;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

const char* pm_last_error(void) { return g_last_error.c_str(); }
int pm_abi_version(void) { return PM_ABI_VERSION; }
int pm_init(void) {
  int st = init_device_tables();
  if (st) return st;
  return sort::init_noalloc_scratch();
}

int pm_timing_enable(int32_t max_events) {
  Timing& t = g_timing;
  for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  t.ev.clear();
  t.names.clear();
  t.used = 0;
  t.on = false;
  if (max_events <= 0) return PM_SUCCESS;
  t.ev.resize(max_events);
  t.names.resize(max_events, nullptr);
  for (int i = 0; i < max_events; ++i) PM_HIP_TRY(hipEventCreate(&t.ev[i]));
  t.on = true;
  return PM_SUCCESS;
}

int pm_timing_mark(const char* name, void* stream) {
  PM_RETURN_IF(!g_timing.on, PM_ERROR_INVALID_ARGUMENT, "timing is not enabled");
  timing_mark(name, as_stream(stream));
  return PM_SUCCESS;
}

int pm_timing_read(float* ms, const char** names, int32_t capacity, int32_t* count) {
  Timing& t = g_timing;
  PM_RETURN_IF(count == nullptr, PM_ERROR_INVALID_ARGUMENT, "count is NULL");
  *count = 0;
  if (!t.on || t.used == 0) return PM_SUCCESS;
  PM_HIP_TRY(hipEventSynchronize(t.ev[t.used - 1]));
  int n = 0;
  for (size_t i = 1; i < t.used && n < capacity; ++i, ++n) {
    float e = 0.f;
    PM_HIP_TRY(hipEventElapsedTime(&e, t.ev[i - 1], t.ev[i]));
    ms[n] = e;
    names[n] = t.names[i];
  }
  *count = n;
  t.used = 0;
  return PM_SUCCESS;
}

int pm_cic_scatter(const int16_t* pmid, const float* disp, const uint8_t* valid, const float* val,
                   float val_scalar, int64_t n, int32_t channels, const pm_cic_geom* geom, float* mesh,
                   void* stream) {
  int st = check_cic_args(pmid, disp, n, channels, geom);
  if (st) return st;
  PM_RETURN_IF(mesh == nullptr, PM_ERROR_INVALID_ARGUMENT, "mesh is NULL");
  if (n == 0) return PM_SUCCESS;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), pmid, disp, valid, val, val_scalar, n, channels,
                     make_geom(*geom), mesh);
  PM_LAUNCHED("k_scatter", as_stream(stream));
  return PM_SUCCESS;
}

int pm_cic_gather(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                  int32_t channels, const pm_cic_geom* geom, const float* mesh, float* out,
                  void* stream) {
  int st = check_cic_args(pmid, disp, n, channels, geom);
  if (st) return st;
  PM_RETURN_IF(mesh == nullptr || (n > 0 && out == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "mesh/out NULL");
  if (n == 0) return PM_SUCCESS;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), pmid, disp, valid, n, channels, make_geom(*geom), mesh, out);
  PM_LAUNCHED("k_gather", as_stream(stream));
  return PM_SUCCESS;
}

int pm_cic_gather_bwd(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                      int32_t channels, const pm_cic_geom* geom, const float* mesh,
                      const float* val_cot, float* disp_cot, float* mesh_cot, void* stream) {
  int st = check_cic_args(pmid, disp, n, channels, geom);
  if (st) return st;
  PM_RETURN_IF(mesh == nullptr || mesh_cot == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "mesh/mesh_cot NULL");
  PM_RETURN_IF(n > 0 && (val_cot == nullptr || disp_cot == nullptr), PM_ERROR_INVALID_ARGUMENT,
               "val_cot/disp_cot NULL");
  if (n == 0) return PM_SUCCESS;
  hipLaunchKernelGGL(k_gather_bwd, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), pmid, disp, valid, n, channels, make_geom(*geom), mesh,
                     val_cot, disp_cot, mesh_cot);
  PM_LAUNCHED("k_gather_bwd", as_stream(stream));
  return PM_SUCCESS;
}

int pm_cic_scatter_bwd(const int16_t* pmid, const float* disp, const uint8_t* valid,
                       const float* val, float val_scalar, int64_t n, int32_t channels,
                       const pm_cic_geom* geom, const float* mesh_cot, float* disp_cot,
                       float* val_cot, void* stream) {
  int st = check_cic_args(pmid, disp, n, channels, geom);
  if (st) return st;
  PM_RETURN_IF(mesh_cot == nullptr || val_cot == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "mesh_cot/val_cot NULL");
  PM_RETURN_IF(n > 0 && disp_cot == nullptr, PM_ERROR_INVALID_ARGUMENT, "disp_cot NULL");
  if (n == 0) return PM_SUCCESS;
  hipLaunchKernelGGL(k_scatter_bwd, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), pmid, disp, valid, val, val_scalar, n, channels,
                     make_geom(*geom), mesh_cot, disp_cot, val_cot);
  PM_LAUNCHED("k_scatter_bwd", as_stream(stream));
  return PM_SUCCESS;
}

}  // extern "C"
