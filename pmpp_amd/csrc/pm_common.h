// Shared host/device helpers of the pmpp_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pmpp_amd.h"

namespace pm {

// ------------------------------------------------------------------ errors --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PM_HIP_TRY(expr)                                                               \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return ::pm::fail(PM_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PM_LAUNCH_CHECK(what)                                                          \
  do {                                                                                 \
    hipError_t _e = hipGetLastError();                                                 \
    if (_e != hipSuccess)                                                              \
      return ::pm::fail(PM_ERROR_HIP, std::string("launch ") + (what) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PM_RETURN_IF(cond, code, msg)                                                  \
  do {                                                                                 \
    if (cond) return ::pm::fail((code), (msg));                                        \
  } while (0)

#define PM_TRY(expr)     \
  do {                   \
    int _st = (expr);    \
    if (_st) return _st; \
  } while (0)

int init_device_tables();     // pm_init: upload the twiddle table of the current device (synchronises)
int require_device_tables();  // error unless pm_init has run on the current device

// Optional per-kernel timing (pm_timing_* in the ABI): when enabled on the calling host
// thread, mark() records a HIP event on the stream after a kernel launch.
void timing_mark(const char* name, hipStream_t s);
#define PM_MARK(name, stream) ::pm::timing_mark((name), (stream))

// launch check + timing mark (name must have static storage: the timer keeps the pointer)
#define PM_LAUNCHED(what, stream) \
  do {                            \
    PM_LAUNCH_CHECK(what);        \
    PM_MARK(what, stream);        \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller workspace (256-byte aligned slices).
struct Arena {
  char* base;
  size_t cap;
  size_t used = 0;
  Arena(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <class T>
  T* take(size_t count) {
    size_t off = align_up(used, 256);
    used = off + count * sizeof(T);
    return (base && used <= cap) ? reinterpret_cast<T*>(base + off) : nullptr;
  }
  bool ok() const { return base != nullptr && used <= cap; }
};

// ------------------------------------------------------------- device math --
__device__ __forceinline__ int pos_mod(int v, int n) {
  int r = v % n;
  return r < 0 ? r + n : r;
}

// Per-axis CIC factors of one particle: for each axis and neighbour bit b in {0,1}
// the wrapped index, the 1-D weight 1-|delta| and the sign of d(weight)/d(delta-position)
// (the reference's sign(-delta), with sign(0) = 0).
struct AxisCic {
  int idx[2];
  float w[2];
  float s[2];
};

__device__ __forceinline__ float cic_sign(float delta) {
  return delta < 0.f ? 1.f : (delta > 0.f ? -1.f : 0.f);
}

// a2 = None branch (src/pmpp/cic/pallas.py:227-256, cic/enmesh.py:126-147):
// q = floor(offset/a1), r = offset - q a1, t = (disp - r)/a1, l = floor(t),
// base = pmid - q + l, f = t - l; corner b: idx = (base+b) mod N, w = 1-|f-b|.
__device__ __forceinline__ void axis_cic_same_cell(int pmid, float disp, float offset, float a1,
                                                   int global_n, AxisCic& out) {
  float qf = floorf(offset / a1);
  int q = (int)qf;
  float rem = offset - qf * a1;
  float t = (disp - rem) / a1;
  float lf = floorf(t);
  int base = pmid - q + (int)lf;
  float f = t - lf;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    out.idx[b] = pos_mod(base + b, global_n);
    float delta = f - (float)b;
    out.w[b] = 1.f - fabsf(delta);
    out.s[b] = cic_sign(delta);
  }
}

// Explicit mesh cell size a2 (src/pmpp/cic/pallas.py:192-223): position wrapped into the
// periodic box L = N a1, neighbour index floor(mod(pos + b a2, L)/a2), delta folded into
// (-L/2, L/2].
__device__ __forceinline__ float fmod_pos(float x, float m) {
  // jnp.mod for floats: result has the sign of the divisor
  float r = fmodf(x, m);
  return (r != 0.f && r < 0.f) ? r + m : r;
}

__device__ __forceinline__ void axis_cic_explicit(int pmid, float disp, float offset, float a1,
                                                  float a2, int global_n, AxisCic& out) {
  float period = (float)global_n * a1;
  float position = (float)pmid * a1 + disp - offset;
  position = fmod_pos(position, period);
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    float np = fmod_pos(position + (float)b * a2, period);
    int idx = (int)floorf(np / a2);
    float delta = position - (float)idx * a2;
    delta = delta > 0.5f * period ? delta - period : delta;
    delta = delta < -0.5f * period ? delta + period : delta;
    float dn = delta / a2;
    out.idx[b] = idx;
    out.w[b] = 1.f - fabsf(dn);
    out.s[b] = cic_sign(dn);
  }
}

struct Geom {
  int ls[3];
  int gs[3];
  float off[3];
  float a1;
  float a2;  // <= 0: same cell size as particles
};

inline Geom make_geom(const pm_cic_geom& g) {
  Geom r;
  for (int i = 0; i < 3; ++i) {
    r.ls[i] = g.local_shape[i];
    r.gs[i] = g.global_shape[i];
    r.off[i] = g.offset[i];
  }
  r.a1 = g.particle_cell;
  r.a2 = g.mesh_cell;
  return r;
}

__device__ __forceinline__ void particle_cic(const Geom& g, const int16_t* pmid, const float* disp,
                                             int64_t p, AxisCic ax[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int pm = (int)pmid[p * 3 + a];
    float d = disp[p * 3 + a];
    if (g.a2 > 0.f)
      axis_cic_explicit(pm, d, g.off[a], g.a1, g.a2, g.gs[a], ax[a]);
    else
      axis_cic_same_cell(pm, d, g.off[a], g.a1, g.gs[a], ax[a]);
  }
}

// Corner c of the reference order (bits (c>>0)&1, (c>>1)&1, (c>>2)&1 on x,y,z).
__device__ __forceinline__ float corner_weight(const AxisCic ax[3], int c) {
  int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
  return ax[0].w[bx] * ax[1].w[by] * ax[2].w[bz];
}

__device__ __forceinline__ bool corner_index(const Geom& g, const AxisCic ax[3], int c,
                                             int64_t& flat) {
  int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
  int ix = ax[0].idx[bx], iy = ax[1].idx[by], iz = ax[2].idx[bz];
  bool ok = ix >= 0 && ix < g.ls[0] && iy >= 0 && iy < g.ls[1] && iz >= 0 && iz < g.ls[2];
  flat = ((int64_t)ix * g.ls[1] + iy) * g.ls[2] + iz;
  return ok;
}

// d W_c / d(normalised position) per axis: sign_i * w_{i+1} * w_{i+2} (pallas.py:250-255).
__device__ __forceinline__ void corner_grad(const AxisCic ax[3], int c, float gr[3]) {
  int b[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    gr[i] = ax[i].s[b[i]] * ax[j].w[b[j]] * ax[k].w[b[k]];
  }
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
  // Hardware float atomic on coarse-grained device memory (global_atomic_add_f32).
  unsafeAtomicAdd(p, v);
}

}  // namespace pm
