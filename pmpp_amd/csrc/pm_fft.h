// Register/LDS Stockham-style FFT passes for gfx950 (wave64, 160 KB LDS per CU).
//
// A length-N line (N = 2^L <= 2048) is split N = N1 * N2 (N1 >= N2) and handled by
// P = N2 threads, each holding E = N1 complex values in registers:
//
//   forward:  x[N2 n1 + n2]  --DFT_N1 over n1 (registers)--> * W_N^(n2 k1)
//             --LDS exchange--> DFT_N2 over n2 (registers) --> X[k1 + N1 k2]
//   inverse:  X[k1 + N1 k2]  --DFT_N2 over k2--> * W_N^(-n2 k1) --LDS exchange-->
//             DFT_N1 over k1 --> x[n2 + N2 n1]
//
// so every pass reads and writes HBM exactly once and exchanges through LDS once; a
// forward pass followed by an inverse pass on the same axis needs no exchange between
// them (the forward output layout is the inverse input layout).  Butterfly twiddles
// inside the register DFTs are compile-time constants; the inter-stage twiddles
// W_N^j come from a 4096-entry float2 table computed in double precision on the host.
#pragma once

#include <type_traits>
#include <utility>

#include "pm_common.h"

namespace pm {
namespace fft {

constexpr int kTwN = 4096;  // twiddle table resolution: W_4096^i = exp(-2 pi i / 4096)

const float2* twiddle_table();  // device address of the current device's table

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr int bitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r = (r << 1) | ((x >> i) & 1);
  return r;
}

// ---------------------------------------------------------- compile-time trig
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
struct CS {
  double c, s;
};
// cos, sin of 2 pi k / n, reduced on integers to an angle in [0, pi/4].
constexpr CS cs2pi(long k, long n) {
  long m = ((k % n) + n) % n;
  long a = 8 * m;  // angle = 2 pi a / (8 n)
  double sc = 1.0, ss = 1.0;
  if (a > 4 * n) {  // (pi, 2pi): sin changes sign
    a = 8 * n - a;
    ss = -1.0;
  }
  if (a > 2 * n) {  // (pi/2, pi]: cos changes sign
    a = 4 * n - a;
    sc = -1.0;
  }
  bool swap = false;
  if (a > n) {  // (pi/4, pi/2]
    a = 2 * n - a;
    swap = true;
  }
  double x = 2.0 * kPi * (double)a / (8.0 * (double)n);
  double c = taylor_cos(x), s = taylor_sin(x);
  if (swap) {
    double t = c;
    c = s;
    s = t;
  }
  return CS{sc * c, ss * s};
}

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_mulconj(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
// (i * m) * a for real m
__device__ __forceinline__ float2 c_imul(float m, float2 a) { return make_float2(-m * a.y, m * a.x); }

// b * exp(-+ 2 pi i K / LEN) with compile-time constants (forward: minus, inverse: plus)
template <int K, int LEN, bool INV>
__device__ __forceinline__ float2 tw_mul(float2 b) {
  if constexpr (K == 0) {
    return b;
  } else if constexpr (4 * K == LEN) {
    return INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
  } else {
    constexpr CS w = cs2pi(K, LEN);
    constexpr float c = (float)w.c;
    constexpr float s = (float)w.s;
    if constexpr (INV)
      return make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
    else
      return make_float2(b.x * c + b.y * s, b.y * c - b.x * s);
  }
}

// In-register DFT of length N (power of two) on v[0..N), natural order in and out.
template <int N, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (N >= 2) {
    constexpr int L = ilog2(N);
    static_for<0, N>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int j = bitrev(i, L);
      if constexpr (i < j) {
        float2 t = v[i];
        v[i] = v[j];
        v[j] = t;
      }
    });
    static_for<1, L + 1>([&](auto S) {
      constexpr int len = 1 << decltype(S)::value;
      constexpr int half = len / 2;
      static_for<0, N / len>([&](auto Bk) {
        static_for<0, half>([&](auto Kk) {
          constexpr int k = decltype(Kk)::value;
          constexpr int i = decltype(Bk)::value * len + k;
          float2 a = v[i];
          float2 t = tw_mul<k, len, INV>(v[i + half]);
          v[i] = c_add(a, t);
          v[i + half] = c_sub(a, t);
        });
      });
    });
  }
}

template <int N>
struct Plan {
  static constexpr int L = ilog2(N);
  static constexpr int N1 = 1 << ((L + 1) / 2);
  static constexpr int N2 = N / N1;
  static constexpr int P = N2;       // threads per line
  static constexpr int E = N1;       // complex values per thread
  static constexpr int J3 = N1 / N2; // length-N2 DFTs per thread in the second stage (1 or 2)
  static_assert(N1 * N2 == N, "power of two");
};

// The twiddle table through a pointer the compiler cannot see through: table loads made
// through it are not merged with loads of the same entries made before it.  A kernel with a
// forward and an inverse transform otherwise keeps the forward's N1 - 1 twiddles live
// (2 (N1 - 1) VGPRs) until the inverse uses the same entries again.
__device__ __forceinline__ const float2* fresh_table(const float2* t) {
  uint64_t u = reinterpret_cast<uint64_t>(t);
  asm volatile("" : "+s"(u));
  return reinterpret_cast<const float2*>(u);
}

// 1/x from the hardware reciprocal and one Newton step: correctly rounded but for a few
// arguments per million (1 ulp off then), with a fraction of the registers of an IEEE
// division, whose temporaries the compiler interleaves across a column's unrolled slots
// (the x Green pass: 258 -> 200 VGPRs, the potential pass 226 -> 164).  x = 0 gives NaN; the
// callers select 0 for that mode.
__device__ __forceinline__ float recip_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, r, 1.f), r, r);
}

// W_N^j (forward sign) from the table; INV conjugates.
template <int N, bool INV>
__device__ __forceinline__ float2 table_tw(const float2* __restrict__ tw, int j) {
  float2 w = tw[(j & (N - 1)) * (kTwN / N)];
  return INV ? c_conj(w) : w;
}

// LDS addressing: the logical exchange matrix row = k1 * N2 + n2.  Idx maps a row to a
// float offset for the calling thread (its column / line slot is folded in by the functor).

// Forward core.  In: v[n1] = x[N2 n1 + p].  Out: v[j3 N2 + k2] = X[(p + P j3) + N1 k2].
template <int N, class Idx>
__device__ __forceinline__ void forward_core(float2* v, float* lds, int p, const Idx& idx,
                                             const float2* __restrict__ tw) {
  using PL = Plan<N>;
  constexpr int N1 = PL::N1, N2 = PL::N2, P = PL::P, E = PL::E;
  dft<N1, false>(v);
  if constexpr (N2 > 1) {
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = c_mul(v[k1], table_tw<N, false>(tw, p * k1));
  }
  float tmp[E];
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) lds[idx(k1 * N2 + p)] = v[k1].x;
  __syncthreads();
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3)
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) tmp[j3 * N2 + n2] = lds[idx((p + P * j3) * N2 + n2)];
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) lds[idx(k1 * N2 + p)] = v[k1].y;
  __syncthreads();
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3)
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2)
      v[j3 * N2 + n2] = make_float2(tmp[j3 * N2 + n2], lds[idx((p + P * j3) * N2 + n2)]);
  __syncthreads();
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3) dft<N2, false>(v + j3 * N2);
}

// Inverse core (unnormalised).  In: v[j3 N2 + k2] = X[(p + P j3) + N1 k2].
// Out: v[n1] = x[p + N2 n1].
template <int N, class Idx>
__device__ __forceinline__ void inverse_core(float2* v, float* lds, int p, const Idx& idx,
                                             const float2* __restrict__ tw) {
  using PL = Plan<N>;
  constexpr int N1 = PL::N1, N2 = PL::N2, P = PL::P, E = PL::E;
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3) dft<N2, true>(v + j3 * N2);
  if constexpr (N2 > 1) {
#pragma unroll
    for (int j3 = 0; j3 < PL::J3; ++j3)
#pragma unroll
      for (int n2 = 1; n2 < N2; ++n2)
        v[j3 * N2 + n2] = c_mul(v[j3 * N2 + n2], table_tw<N, true>(tw, n2 * (p + P * j3)));
  }
  float tmp[E];
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3)
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) lds[idx((p + P * j3) * N2 + n2)] = v[j3 * N2 + n2].x;
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) tmp[k1] = lds[idx(k1 * N2 + p)];
  __syncthreads();
#pragma unroll
  for (int j3 = 0; j3 < PL::J3; ++j3)
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) lds[idx((p + P * j3) * N2 + n2)] = v[j3 * N2 + n2].y;
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) v[k1] = make_float2(tmp[k1], lds[idx(k1 * N2 + p)]);
  __syncthreads();
  dft<N1, true>(v);
}

// Row (position along the FFT axis) held in register slot s after forward_core / before
// inverse_core ("X layout") and before forward_core / after inverse_core ("natural").
template <int N>
__device__ __forceinline__ int xrow(int p, int s) {
  using PL = Plan<N>;
  return (p + PL::P * (s / PL::N2)) + PL::N1 * (s % PL::N2);
}
template <int N>
__device__ __forceinline__ int nrow(int p, int s) {
  return p + Plan<N>::N2 * s;
}

// ------------------------------------------------------------ pass descriptors

// Spectral tables in the reference's float32 values (numerics/fft.py:4-44,
// core/utils.py:15-37, core/configuration.py:365-376).
struct SpecTables {
  const float* k[3];     // angular wavenumbers
  const float* mask[3];  // particle-Nyquist masks (0/1)
  const float* negk[3];  // -k with the Nyquist derivative zeroed
  // the scalars the tables are made from, for passes that evaluate them in registers
  double period;    // 2 pi / cell
  float klim;       // particle-Nyquist limit pi/l_p (1 + 8 eps_f32), rounded as the reference
  float nyq;        // mesh Nyquist pi / cell
  float nyq_eps;    // its zero-derivative tolerance (pi / cell) eps_f32
  int mask_active;
};

// Angular wavenumber of index i on an axis of n points (numerics/fft.py:30-41): the double
// product m * (1/n) * (2 pi / cell) rounded once to float; `half` is the rfft axis.
__device__ __forceinline__ float spec_k(int i, int n, bool half, double period) {
  const int m = half ? i : (i < (n + 1) / 2 ? i : i - n);
  return (float)((double)m * (1.0 / (double)n) * period);
}
__device__ __forceinline__ bool spec_keep(float k, const SpecTables& tb) {
  return !tb.mask_active || fabsf(k) <= tb.klim;
}
__device__ __forceinline__ float spec_negk(float k, const SpecTables& tb) {
  return (fabsf(fabsf(k) - tb.nyq) <= tb.nyq_eps) ? 0.f : -k;
}

// Column pass: FFT along a strided axis over tiles of T adjacent columns (unit stride).
// element (row r, column c, batch b) lives at base[b*s_batch + r*s_row + c*s_col].
struct ColPass {
  float2* in;
  float2* out0;
  float2* out1;
  int64_t s_row, s_col, s_batch;
  int ncols;
  // index of the other two spectral axes for the Green multiply: for column c of batch b,
  // ky = b*ky_b + c*ky_c (then + ky_shift when above ky_split), kz = b*kz_b + c*kz_c + kz0
  int ky_b, ky_c, kz_b, kz_c, kz0;
  // output strides (-1: same as the input's)
  int64_t o_row = -1, o_batch = -1;
  // band compaction of the transformed axis: only spectral rows |k| <= keep (k <= keep or
  // k >= N - keep) are stored, packed as 0..keep, then N-keep..N-1 -> keep+1..2*keep.
  // keep < 0: full rows.  compact_in selects the packed side: 1 = input (inverse passes
  // expand the band to full rows, the rows outside it are zero), 0 = output (forward passes).
  int keep = -1, compact_in = 0;
  int ky_split = 0x7fffffff, ky_shift = 0;
  // kColAdj: the y and z components' spectra (in is the x component's)
  const float2* in1 = nullptr;
  const float2* in2 = nullptr;
};

// kColAdj: the x pass of the force step's spectral adjoint: forward transforms of the three
// cotangent spectra, out = mask * (-1/k^2) * sum_i conj(D_i) in_i, inverse transform
// kColPot: x forward, mask * (-1/k^2), x inverse (the potential, one output)
enum ColMode { kColFwd = 0, kColInv = 1, kColInv2 = 2, kColGreen = 3, kColAdj = 4, kColPot = 5 };

// Line pass along the contiguous last axis (real <-> half spectrum, split main + Nyquist).
struct ZPass {
  const float* rin;    // R2C: real input rows (pitch_r floats)
  float* rout;         // C2R: real output rows
  float2* cmain;       // spectrum rows kz in [0, M) (pitch_c float2)
  float2* cnyq;        // Nyquist kz = M, one float2 per row at stride nyq_stride
  int64_t pitch_r, pitch_c, nyq_stride;
  int64_t nlines;
  float src_shift;  // R2C pre-op: (x + src_shift) * src_scale
  float src_scale;
  float out_scale;  // C2R: output scale
  int deriv;        // C2R pre-op: multiply by i*negk_z[k] when nonzero
  // band of stored spectrum columns: R2C stores kz < kz_lim (and the Nyquist column only when
  // nyq != 0); C2R reads them and takes every other column, the Nyquist one included when
  // nyq == 0, as zero
  int kz_lim = 0x7fffffff;
  int nyq = 1;
  // R2C: optional second real input (same layout) added before the pre-op
  const float* rin_add = nullptr;
  // C2R: optional real mesh (same layout as rout, may alias it) added to the scaled output
  const float* rout_add = nullptr;
};

}  // namespace fft
}  // namespace pm
