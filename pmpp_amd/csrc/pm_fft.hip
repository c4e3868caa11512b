// FFT passes (real<->half-spectrum along z, complex along y/x) and the fused spectral
// force passes of the PM step.  See pm_fft.h for the decomposition.
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "pm_fft.h"
#include "pm_fft_launch.h"

namespace pm {
namespace fft {

// ------------------------------------------------------------ twiddle table ----
static __device__ float2 g_twiddle[kTwN];

namespace {
std::mutex g_tw_mutex;
std::atomic<const float2*> g_tw_ptr[64];
}  // namespace

const float2* twiddle_table() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  return g_tw_ptr[dev].load(std::memory_order_acquire);
}

}  // namespace fft

// The one place that uploads the table and synchronises (pm_init): every other entry point
// only enqueues work on its stream, so it can run inside HIP graph capture.
int init_device_tables() {
  int dev = 0;
  PM_HIP_TRY(hipGetDevice(&dev));
  PM_RETURN_IF(dev < 0 || dev >= 64, PM_ERROR_UNSUPPORTED, "device ordinal >= 64");
  std::lock_guard<std::mutex> lock(fft::g_tw_mutex);
  if (fft::g_tw_ptr[dev].load(std::memory_order_acquire)) return PM_SUCCESS;
  std::vector<float2> host(fft::kTwN);
  for (int i = 0; i < fft::kTwN; ++i) {
    // exp(-2 pi i j / 4096) in double, rounded once to float
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)i / fft::kTwN;
    host[i] = make_float2((float)cosl(a), (float)(-sinl(a)));
  }
  host[0] = make_float2(1.f, 0.f);
  host[fft::kTwN / 4] = make_float2(0.f, -1.f);
  host[fft::kTwN / 2] = make_float2(-1.f, 0.f);
  host[3 * fft::kTwN / 4] = make_float2(0.f, 1.f);
  PM_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(fft::g_twiddle), host.data(), sizeof(float2) * fft::kTwN));
  void* ptr = nullptr;
  PM_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(fft::g_twiddle)));
  PM_HIP_TRY(hipDeviceSynchronize());
  fft::g_tw_ptr[dev].store(static_cast<const float2*>(ptr), std::memory_order_release);
  return PM_SUCCESS;
}

int require_device_tables() {
  PM_RETURN_IF(fft::twiddle_table() == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "pm_init has not run on the current device (it prepares the FFT twiddle table)");
  return PM_SUCCESS;
}

namespace fft {
namespace {

// ----------------------------------------------------------- column passes ----
// T adjacent columns per workgroup.  The y inverse with two outputs (kColInv2) runs two
// workgroups per tile (HALVES), each of which loads the same input tile (the second read is
// served by L2: the pair runs on one XCD at nearly the same time) and inverts one of the two
// outputs, so no thread keeps both spectra live.  The x Green pass (kColGreen) runs one
// workgroup per tile that parks its multiplied spectrum in its first output between the two
// inverses.  Outputs of two-output passes must not alias the input or each other.
template <int N, int T, int HALVES>
struct ColCfg {
  static constexpr int P = Plan<N>::P;
  static constexpr int THREADS = T * P;
  static constexpr int PADK = T;  // per-k1-block LDS pad: spreads the strided rows over banks
  static constexpr int LDS_FLOATS = N * T + Plan<N>::N1 * PADK;
};

// Column tiles are visited in an XCD-aware order: workgroups are dispatched round-robin
// over the 8 XCDs, so logical item q is given to workgroup (q % g8) * 8 + q / g8 ... i.e.
// consecutive logical items (the two halves of a tile, and neighbouring tiles, which share
// cache lines when T*8 < 128 bytes) run on the same XCD at nearly the same time.
template <int N, int T, int MODE, int HALVES>
__global__ void __launch_bounds__((ColCfg<N, T, HALVES>::THREADS))
    k_cols(ColPass io, SpecTables tb, int dax, const float2* __restrict__ tw, int gx, int gy, int g8) {
  using PL = Plan<N>;
  using CF = ColCfg<N, T, HALVES>;
  constexpr int E = PL::E, N2 = PL::N2;
  __shared__ float lds[CF::LDS_FLOATS];
  const int L = blockIdx.x;
  const int q = (L % 8) * g8 + L / 8;
  const int half = HALVES > 1 ? q % HALVES : 0;
  const int tile = HALVES > 1 ? q / HALVES : q;
  const int tilex = tile % gx;
  const int64_t b = tile / gx;
  if (b >= gy) return;  // padding workgroup of the XCD remap: uniform exit before any barrier
  const int lt = threadIdx.x;
  const int col = lt % T;
  const int p = lt / T;
  const int cg = tilex * T + col;
  const bool cok = cg < io.ncols;
  // loads use a clamped column so they are unconditional (branch-free batches stay in flight)
  const int cgl = cok ? cg : io.ncols - 1;
  const int64_t lbase = b * io.s_batch + (int64_t)cgl * io.s_col;
  const int64_t o_row = io.o_row >= 0 ? io.o_row : io.s_row;
  const int64_t obase = b * (io.o_batch >= 0 ? io.o_batch : io.s_batch) + (int64_t)cg * io.s_col;
  // row r = p + d of this thread's column: base + d * stride, with d a compile-time multiple of
  // the row step for every unrolled slot, so the per-slot offsets are uniform (scalar) and no
  // per-slot row index stays live in a VGPR across the transforms
  const int64_t ibase = lbase + (int64_t)p * io.s_row;
  const int64_t obp = obase + (int64_t)p * o_row;
  const int keep = io.keep;
  // band compaction of the transformed axis (see ColPass::keep)
  auto in_band = [&](int k) { return keep < 0 || k <= keep || k >= N - keep; };
  auto packed = [&](int k) { return (keep < 0 || k <= keep) ? k : k - (N - 2 * keep - 1); };
  float* my = lds;
  auto idx = [&](int row) { return row * T + col + (row / N2) * CF::PADK; };

  float2 v[E];
  if constexpr (MODE == kColAdj) {
    // Spectral adjoint of the force step, x pass (the columns are kz, the batches the packed ky
    // rows, as for kColGreen): sum_i conj(D_i) hat g_i with conj(i negk) = -i negk, then
    // mask * (-1/k^2), then the inverse x transform.  The y and z symbols are constant along
    // an x column, so by linearity those two components are combined before their transform:
    // U = X(negk_y g_y + negk_z g_z), then s = -i (negk_x X(g_x) + U); two lines live per
    // thread instead of three.
    int ky = (int)(b * io.ky_b) + cgl * io.ky_c;
    ky += ky > io.ky_split ? io.ky_shift : 0;
    int kz = (int)(b * io.kz_b) + cgl * io.kz_c + io.kz0;
    const float kyv = tb.k[1][ky], kzv = tb.k[2][kz];
    const float nky = tb.negk[1][ky], nkz = tb.negk[2][kz];
    const bool keep_yz = cok && tb.mask[1][ky] != 0.f && tb.mask[2][kz] != 0.f;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int64_t o = ibase + (int64_t)(nrow<N>(p, s) - p) * io.s_row;
      const float2 gy = io.in1[o], gz = io.in2[o];
      v[s] = make_float2(nky * gy.x + nkz * gz.x, nky * gy.y + nkz * gz.y);
    }
    forward_core<N>(v, my, p, idx, tw);
    asm volatile("" ::: "memory");  // the x component's loads after U's transform
    float2 w[E];
#pragma unroll
    for (int s = 0; s < E; ++s) w[s] = io.in[ibase + (int64_t)(nrow<N>(p, s) - p) * io.s_row];
    forward_core<N>(w, my, p, idx, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const float kxv = spec_k(xrow<N>(p, s), N, false, tb.period);
      const float nkx = spec_negk(kxv, tb);
      const float tx = nkx * w[s].x + v[s].x, ty = nkx * w[s].y + v[s].y;  // s = -i t
      const float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
      const bool keep = keep_yz && spec_keep(kxv, tb) && k2 != 0.f;
      const float r = recip_nr(k2);
      v[s] = keep ? make_float2(-ty * r, tx * r) : make_float2(0.f, 0.f);  // -s / k^2
    }
    inverse_core<N>(v, my, p, idx, tw);
    if (cok) {
#pragma unroll
      for (int s = 0; s < E; ++s) io.out0[obp + (int64_t)(nrow<N>(p, s) - p) * o_row] = v[s];
    }
    return;
  } else if constexpr (MODE == kColFwd || MODE == kColGreen || MODE == kColPot) {
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = io.in[ibase + (int64_t)(nrow<N>(p, s) - p) * io.s_row];
    forward_core<N>(v, my, p, idx, tw);
  } else {
    if (keep >= 0 && io.compact_in) {  // packed band in, zero outside it
#pragma unroll
      for (int s = 0; s < E; ++s) {
        int k = xrow<N>(p, s);
        v[s] = io.in[lbase + (int64_t)(in_band(k) ? packed(k) : 0) * io.s_row];
      }
#pragma unroll
      for (int s = 0; s < E; ++s)
        if (!in_band(xrow<N>(p, s))) v[s] = make_float2(0.f, 0.f);
    } else {
#pragma unroll
      for (int s = 0; s < E; ++s) v[s] = io.in[ibase + (int64_t)(xrow<N>(p, s) - p) * io.s_row];
    }
  }

  if constexpr (MODE == kColFwd) {
    if (cok) {
#pragma unroll
      for (int s = 0; s < E; ++s) {
        int k = xrow<N>(p, s);
        if (in_band(k)) io.out0[obase + (int64_t)packed(k) * o_row] = v[s];  // forward: packed output
      }
    }
    return;
  } else {
    if constexpr (MODE == kColGreen || MODE == kColPot) {
      // mask x Poisson symbol (gravity.py:174-196, 263-287); kColGreen's second half also
      // applies D_x = i*negk_x (configuration.py:365-376), kColPot has one half (the potential)
      int ky = (int)(b * io.ky_b) + cgl * io.ky_c;
      ky += ky > io.ky_split ? io.ky_shift : 0;
      int kz = (int)(b * io.kz_b) + cgl * io.kz_c + io.kz0;
      float kyv = tb.k[1][ky], kzv = tb.k[2][kz];
      bool keep_yz = cok && tb.mask[1][ky] != 0.f && tb.mask[2][kz] != 0.f;
      // k_x, its mask and -k_x are evaluated in registers (the same expressions as the
      // tables): 3*E table loads per thread would be hoisted above the transform and stay live
#pragma unroll
      for (int s = 0; s < E; ++s) {
        int kx = xrow<N>(p, s);
        float kxv = spec_k(kx, N, false, tb.period);
        float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
        bool keep = keep_yz && spec_keep(kxv, tb) && k2 != 0.f;
        const float r = recip_nr(k2);
        float g = keep ? -r : 0.f;  // one reciprocal per mode, shared by both parts
        float2 phi = make_float2(v[s].x * g, v[s].y * g);
        v[s] = (half == 1) ? c_imul(spec_negk(kxv, tb), phi) : phi;
      }
      if constexpr (MODE == kColGreen && HALVES == 1) {
        // one workgroup per tile: the multiplied spectrum phi is parked in out0 at this
        // thread's own output rows (where it writes its out0 values at the end, so no other
        // thread touches them), D_x phi is inverted into out1, then phi is read back and
        // inverted into out0 -- three transforms per tile instead of four (two workgroups that
        // each redo the forward); the parked tiles overflow L2, so the pass moves 1.8x its
        // algorithmic bytes, and is faster anyway: the transforms bound it, not HBM
        if (cok) {
#pragma unroll
          for (int s = 0; s < E; ++s) io.out0[obp + (int64_t)(nrow<N>(p, s) - p) * o_row] = v[s];
        }
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = c_imul(spec_negk(spec_k(xrow<N>(p, s), N, false, tb.period), tb), v[s]);
        inverse_core<N>(v, my, p, idx, fresh_table(tw));
        if (cok) {
#pragma unroll
          for (int s = 0; s < E; ++s) io.out1[obp + (int64_t)(nrow<N>(p, s) - p) * o_row] = v[s];
        }
        asm volatile("" ::: "memory");  // no twiddle of the first inverse kept live for the second
#pragma unroll
        for (int s = 0; s < E; ++s)
          v[s] = cok ? io.out0[obp + (int64_t)(nrow<N>(p, s) - p) * o_row] : make_float2(0.f, 0.f);
      }
    } else if constexpr (MODE == kColInv2) {
      if (half == 1) {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = c_imul(tb.negk[dax][xrow<N>(p, s)], v[s]);
      }
    }
    // after a forward transform in the same kernel (kColGreen, kColPot) the inverse's twiddles
    // are loaded afresh instead of kept live from the forward
    inverse_core<N>(v, my, p, idx, (MODE == kColGreen || MODE == kColPot) ? fresh_table(tw) : tw);
    float2* out = (half == 1) ? io.out1 : io.out0;
    if (cok) {
#pragma unroll
      for (int s = 0; s < E; ++s) out[obp + (int64_t)(nrow<N>(p, s) - p) * o_row] = v[s];
    }
  }
}

// --------------------------------------------------------------- z passes ----
template <int M, bool C2R>
struct ZCfg {
  static constexpr int P = Plan<M>::P;
  // long lines fill LDS (one slot per line): smaller workgroups let several of them share a CU
  // and interleave their load / transform / store phases.  From 1024-point lines on, the two
  // directions part ways (M = 512 against 128 threads for both: c2r 1.795 -> 1.76 ms with 256,
  // r2c 1.74 -> 1.63 ms with 64)
  static constexpr int THREADS = M >= 512 ? (C2R ? 256 : 64) : (M >= 256 ? 128 : 256);
  static constexpr int LINES = THREADS / P;
  // floats per line slot, with LSTR = P (mod 64) so the 64/P lines of a wave start on
  // disjoint groups of the 64 LDS banks (the exchange rows of one line are P apart).  R2C
  // pairs bins k and M - k of the complex transform through a natural-order copy of the line
  // (M+1 float2); C2R pairs them through a real-only copy in two rounds, so its slot only holds
  // the larger of that copy (M + 1 floats) and the transform's real/imaginary exchange
  // (M + M/N2 floats), and twice as many lines fit on a CU.
  static constexpr int LSTR_MIN =
      C2R ? (M + M / Plan<M>::N2 > M + 1 ? M + M / Plan<M>::N2 : M + 1) : 2 * (M + 2);
  static constexpr int LSTR = ((LSTR_MIN + 63) / 64) * 64 + (P % 64);
  static constexpr int LDS_FLOATS = LINES * LSTR;
};

// ADD (C2R only): the output is rout_add + the scaled line (a separate instantiation, so the
// force step's c2r keeps its register budget)
template <int M, bool C2R, bool ADD = false>
__global__ void __launch_bounds__((ZCfg<M, C2R>::THREADS)) k_zpass(ZPass zp, SpecTables tb,
                                                                 const float2* __restrict__ tw) {
  using PL = Plan<M>;
  using ZC = ZCfg<M, C2R>;
  constexpr int P = PL::P, E = PL::E, N2 = PL::N2;
  __shared__ float lds[ZC::LDS_FLOATS];
  const int p = threadIdx.x % P;
  const int line = threadIdx.x / P;
  const int64_t lg = (int64_t)blockIdx.x * ZC::LINES + line;
  const bool ok = lg < zp.nlines;
  float* slot = lds + line * ZC::LSTR;
  float2* nat = reinterpret_cast<float2*>(slot);
  auto idx = [&](int row) { return line * ZC::LSTR + row + row / N2; };
  float2 v[E];

  const int64_t lgl = ok ? lg : zp.nlines - 1;  // clamped line: unconditional, batched loads
  if constexpr (!C2R) {
    const float* src = zp.rin + lgl * zp.pitch_r;
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = *reinterpret_cast<const float2*>(src + 2 * nrow<M>(p, s));
    if (zp.rin_add) {  // uniform: the added field's loads join the same batch
      const float* src2 = zp.rin_add + lgl * zp.pitch_r;
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const float2 w = *reinterpret_cast<const float2*>(src2 + 2 * nrow<M>(p, s));
        v[s] = make_float2(v[s].x + w.x, v[s].y + w.y);
      }
    }
#pragma unroll
    for (int s = 0; s < E; ++s)
      v[s] = make_float2((v[s].x + zp.src_shift) * zp.src_scale, (v[s].y + zp.src_shift) * zp.src_scale);
    forward_core<M>(v, lds, p, idx, tw);
#pragma unroll
    for (int s = 0; s < E; ++s) nat[xrow<M>(p, s)] = v[s];
    __syncthreads();
    float2* dst = zp.cmain + lg * zp.pitch_c;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      int k = p + P * j;
      float2 a = nat[k];
      float2 bc = c_conj(nat[(M - k) & (M - 1)]);
      float2 d = c_sub(a, bc);
      float2 fe = c_scale(c_add(a, bc), 0.5f);
      float2 fo = make_float2(0.5f * d.y, -0.5f * d.x);  // -i (a - bc) / 2
      float2 x = c_add(fe, c_mul(table_tw<2 * M, false>(tw, k), fo));
      if (ok && k < zp.kz_lim) dst[k] = x;
    }
    if (ok && p == 0 && zp.nyq) {
      float2 z0 = nat[0];
      zp.cnyq[lg * zp.nyq_stride] = make_float2(z0.x - z0.y, 0.f);
    }
  } else {
    (void)nat;
    const float2* srcm = zp.cmain + lgl * zp.pitch_c;
    float2 xn = zp.cnyq[lgl * zp.nyq_stride];  // cnyq is a valid address even when nyq == 0
    const int klim = zp.kz_lim < M ? zp.kz_lim : M;
#pragma unroll
    for (int j = 0; j < E; ++j)  // P*j < klim is uniform: columns past the band cost no load
      v[j] = (P * j < klim) ? srcm[min(p + P * j, klim - 1)] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < E; ++j)
      if (p + P * j >= klim) v[j] = make_float2(0.f, 0.f);
    if (!zp.nyq) xn = make_float2(0.f, 0.f);
    if (zp.deriv) {
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = c_imul(tb.negk[2][p + P * j], v[j]);
      xn = c_imul(tb.negk[2][M], xn);
    }
    // pair bins k (the X-layout slot) and M - k through a natural-order real copy of the line,
    // real parts first, then imaginary parts (the slot holds M + 1 floats, not M + 1 float2):
    // e = (X[k] + conj X[M-k]) / 2, d = X[k] - conj X[M-k]
    float* rl = lds + line * ZC::LSTR;
    float ex[E], dx[E];
#pragma unroll
    for (int j = 0; j < E; ++j) rl[p + P * j] = v[j].x;
    if (p == 0) rl[M] = xn.x;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int k = xrow<M>(p, s);
      const float a = rl[k], b = rl[M - k];
      ex[s] = a + b;
      dx[s] = a - b;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) rl[p + P * j] = v[j].y;
    if (p == 0) rl[M] = xn.y;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int k = xrow<M>(p, s);
      // irfft uses only the real parts of the DC and Nyquist bins
      const float a = k == 0 ? 0.f : rl[k], b = k == 0 ? 0.f : rl[M - k];
      float2 e = c_scale(make_float2(ex[s], a - b), 0.5f);
      float2 o = c_scale(c_mulconj(make_float2(dx[s], a + b), table_tw<2 * M, false>(tw, k)), 0.5f);
      v[s] = make_float2(e.x - o.y, e.y + o.x);  // e + i o
    }
    __syncthreads();
    inverse_core<M>(v, lds, p, idx, tw);
    if (ok) {
      float* dst = zp.rout + lg * zp.pitch_r;
      if constexpr (ADD) {
        const float* add = zp.rout_add + lg * zp.pitch_r;
#pragma unroll
        for (int s = 0; s < E; ++s) {
          int m = nrow<M>(p, s);
          const float2 w = *reinterpret_cast<const float2*>(add + 2 * m);
          v[s] = make_float2(w.x + v[s].x * zp.out_scale, w.y + v[s].y * zp.out_scale);
        }
      } else {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = make_float2(v[s].x * zp.out_scale, v[s].y * zp.out_scale);
      }
#pragma unroll
      for (int s = 0; s < E; ++s) *reinterpret_cast<float2*>(dst + 2 * nrow<M>(p, s)) = v[s];
    }
  }
}

// ------------------------------------------------------------ spectral tables ----
// Angular wavenumbers k = m * (1/n) * (2 pi / cell) evaluated in double and rounded to
// float (numerics/fft.py:30-41), the particle-Nyquist masks (core/utils.py:33-37) and
// negk with the derivative zeroed at the mesh Nyquist (configuration.py:369-376), all
// decided in float32 like the reference.
__global__ void k_spec_tables(float* out, int n0, int n1, int n2, SpecTables sc) {
  int n[3] = {n0, n1, n2};
  int len[3] = {n0, n1, n2 / 2 + 1};
  int off = 0;
  for (int a = 0; a < 3; ++a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len[a]; i += gridDim.x * blockDim.x) {
      const float k = spec_k(i, n[a], a == 2, sc.period);
      out[off + i] = k;
      out[off + len[a] + i] = spec_keep(k, sc) ? 1.f : 0.f;
      out[off + 2 * len[a] + i] = spec_negk(k, sc);
    }
    off += 3 * len[a];
  }
}

// out = (adjoint ? conj(D) : D) * G * in on a standard-layout spectrum; axis < 0: D = 1.
__global__ void k_spectral_multiply(const float2* __restrict__ in, float2* __restrict__ out, int nx,
                                    int ny, int nzh, SpecTables tb, int axis, int adjoint) {
  int64_t total = (int64_t)nx * ny * nzh;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int kz = (int)(i % nzh);
    int ky = (int)((i / nzh) % ny);
    int kx = (int)(i / ((int64_t)nzh * ny));
    float kxv = tb.k[0][kx], kyv = tb.k[1][ky], kzv = tb.k[2][kz];
    float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    bool keep = tb.mask[0][kx] != 0.f && tb.mask[1][ky] != 0.f && tb.mask[2][kz] != 0.f && k2 != 0.f;
    float2 v = in[i];
    float2 phi = keep ? make_float2(-v.x / k2, -v.y / k2) : make_float2(0.f, 0.f);
    if (axis >= 0) {
      int ki = axis == 0 ? kx : (axis == 1 ? ky : kz);
      float nk = tb.negk[axis][ki];
      phi = c_imul(adjoint ? -nk : nk, phi);
    }
    out[i] = phi;
  }
}

template <int N, int T, int MODE, int HALVES>
int launch_cols_cfg(const ColPass& cp, int batches, const SpecTables& tb, int dax, hipStream_t s,
                    const char* name) {
  using CF = ColCfg<N, T, HALVES>;
  const float2* tw = twiddle_table();
  PM_RETURN_IF(tw == nullptr, PM_ERROR_HIP, "twiddle table not initialised");
  const int gx = (int)ceil_div(cp.ncols, T);
  const int64_t tiles = (int64_t)gx * batches;
  const int64_t g = ceil_div(tiles * HALVES, 8) * 8;
  PM_RETURN_IF((HALVES > 1 || MODE == kColGreen) && (cp.in == cp.out0 || cp.in == cp.out1 || cp.out0 == cp.out1),
               PM_ERROR_INVALID_ARGUMENT, "two-output FFT passes need outputs distinct from their input");
  PM_RETURN_IF(g > INT32_MAX, PM_ERROR_UNSUPPORTED, "too many FFT column tiles");
  hipLaunchKernelGGL((k_cols<N, T, MODE, HALVES>), dim3((unsigned)g), dim3(CF::THREADS), 0, s, cp, tb, dax, tw,
                     gx, batches, (int)(g / 8));
  PM_LAUNCH_CHECK(name);
  PM_MARK(name, s);
  return PM_SUCCESS;
}

template <int N>
int launch_cols_n(int mode, const ColPass& cp, int batches, const SpecTables& tb, int dax,
                  hipStream_t s) {
  constexpr int T1 = (N >= 512) ? 8 : 16;  // one-output passes: 64 or 128-byte row segments
  constexpr int T2 = (N >= 2048) ? 4 : 8;   // two-output passes: two workgroups of T2 columns per tile
  switch (mode) {
    case kColFwd: return launch_cols_cfg<N, T1, kColFwd, 1>(cp, batches, tb, dax, s, "fft_cols_fwd");
    case kColInv: return launch_cols_cfg<N, T1, kColInv, 1>(cp, batches, tb, dax, s, "fft_cols_inv");
    case kColInv2: return launch_cols_cfg<N, T2, kColInv2, 2>(cp, batches, tb, dax, s, "fft_cols_inv2");
    case kColGreen: return launch_cols_cfg<N, T2, kColGreen, 1>(cp, batches, tb, dax, s, "fft_cols_green");
    case kColAdj: return launch_cols_cfg<N, T1, kColAdj, 1>(cp, batches, tb, dax, s, "fft_cols_adjoint");
    case kColPot: return launch_cols_cfg<N, T2, kColPot, 1>(cp, batches, tb, dax, s, "fft_cols_potential");
    default: return fail(PM_ERROR_INVALID_ARGUMENT, "bad column mode");
  }
}

template <int M>
int launch_z_n(bool c2r, const ZPass& zp, const SpecTables& tb, hipStream_t s) {
  using ZI = ZCfg<M, true>;   // c2r workgroups
  using ZF = ZCfg<M, false>;  // r2c workgroups
  const float2* tw = twiddle_table();
  PM_RETURN_IF(tw == nullptr, PM_ERROR_HIP, "twiddle table not initialised");
  if (zp.nlines == 0) return PM_SUCCESS;
  if (c2r) {
    const dim3 grid((unsigned)ceil_div(zp.nlines, ZI::LINES));
    if (zp.rout_add) {
      hipLaunchKernelGGL((k_zpass<M, true, true>), grid, dim3(ZI::THREADS), 0, s, zp, tb, tw);
      PM_LAUNCHED("fft_z_c2r_add", s);
    } else {
      hipLaunchKernelGGL((k_zpass<M, true>), grid, dim3(ZI::THREADS), 0, s, zp, tb, tw);
      PM_LAUNCHED("fft_z_c2r", s);
    }
  } else {
    const dim3 grid((unsigned)ceil_div(zp.nlines, ZF::LINES));
    hipLaunchKernelGGL((k_zpass<M, false>), grid, dim3(ZF::THREADS), 0, s, zp, tb, tw);
    PM_LAUNCHED("fft_z_r2c", s);
  }
  return PM_SUCCESS;
}

}  // namespace

#define PM_POW2_CASES(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

int launch_cols(int n, int mode, const ColPass& cp, int batches, const SpecTables& tb, int dax,
                hipStream_t s) {
  if (batches == 0 || cp.ncols == 0) return PM_SUCCESS;
  switch (n) {
#define PM_CASE(V) \
  case V: return launch_cols_n<V>(mode, cp, batches, tb, dax, s);
    PM_POW2_CASES(PM_CASE)
#undef PM_CASE
    default: return fail(PM_ERROR_UNSUPPORTED, "FFT axis length must be a power of two in [2, 2048]");
  }
}

int launch_z(int nz, bool c2r, const ZPass& zp, const SpecTables& tb, hipStream_t s) {
  if (nz < 2 || nz > 2048 || !is_pow2(nz))
    return fail(PM_ERROR_UNSUPPORTED, "last FFT axis must be a power of two in [2, 2048]");
  switch (nz / 2) {
#define PM_CASE(V) \
  case V: return launch_z_n<V>(c2r, zp, tb, s);
    PM_CASE(1) PM_CASE(2) PM_CASE(4) PM_CASE(8) PM_CASE(16) PM_CASE(32) PM_CASE(64) PM_CASE(128)
    PM_CASE(256) PM_CASE(512) PM_CASE(1024)
#undef PM_CASE
    default: return fail(PM_ERROR_UNSUPPORTED, "last FFT axis must be a power of two in [2, 2048]");
  }
}

bool supported_shape(const int32_t* shape) {
  for (int a = 0; a < 3; ++a)
    if (shape[a] < 2 || shape[a] > 2048 || !is_pow2(shape[a])) return false;
  return true;
}

size_t spec_table_floats(const int32_t* shape) {
  return 3 * ((size_t)shape[0] + shape[1] + (shape[2] / 2 + 1));
}

SpecTables spec_scalars(double cell_size, double ptcl_spacing, int mask_active) {
  SpecTables tb{};
  const float eps32 = 1.1920928955078125e-07f;
  tb.period = 2.0 * 3.14159265358979323846 / cell_size;
  const float k_nyq_ptcl = (float)(3.14159265358979323846 / ptcl_spacing);
  // two float32 roundings as the reference (volatile: no contraction into an fma)
  volatile float tol = k_nyq_ptcl * (8.0f * eps32);
  tb.klim = k_nyq_ptcl + tol;
  tb.nyq = (float)(3.14159265358979323846 / cell_size);
  tb.nyq_eps = tb.nyq * eps32;
  tb.mask_active = mask_active;
  return tb;
}

SpecTables spec_tables_view(const float* base, const int32_t* shape, const SpecTables& scalars) {
  SpecTables tb = scalars;
  int len[3] = {shape[0], shape[1], shape[2] / 2 + 1};
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    tb.k[a] = base + off;
    tb.mask[a] = base + off + len[a];
    tb.negk[a] = base + off + 2 * len[a];
    off += 3 * (size_t)len[a];
  }
  return tb;
}

int launch_spec_tables(float* out, const int32_t* shape, const SpecTables& scalars, hipStream_t s) {
  hipLaunchKernelGGL(k_spec_tables, dim3(8), dim3(256), 0, s, out, shape[0], shape[1], shape[2], scalars);
  PM_LAUNCHED("k_spec_tables", s);
  return PM_SUCCESS;
}

int launch_spectral_multiply(const float2* in, float2* out, const int32_t* shape, const SpecTables& tb,
                             int axis, int adjoint, hipStream_t s) {
  int64_t total = (int64_t)shape[0] * shape[1] * (shape[2] / 2 + 1);
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
  hipLaunchKernelGGL(k_spectral_multiply, dim3(blocks), dim3(256), 0, s, in, out, shape[0], shape[1],
                     shape[2] / 2 + 1, tb, axis, adjoint);
  PM_LAUNCHED("k_spectral_multiply", s);
  return PM_SUCCESS;
}

// Forward 3D r2c on a half spectrum with row pitch `pitch` (>= nz/2+1), in place over
// `spec` (z pass writes it, y and x passes update it).
int rfft3d_pitched(const float* in, float2* spec, const int32_t* shape, int64_t pitch, float shift,
                   float scale, const SpecTables& tb, hipStream_t s) {
  const int nx = shape[0], ny = shape[1], nz = shape[2], m = nz / 2;
  ZPass zp{};
  zp.rin = in;
  zp.cmain = spec;
  zp.cnyq = spec + m;
  zp.pitch_r = nz;
  zp.pitch_c = pitch;
  zp.nyq_stride = pitch;
  zp.nlines = (int64_t)nx * ny;
  zp.src_shift = shift;
  zp.src_scale = scale;
  int st = launch_z(nz, false, zp, tb, s);
  if (st) return st;
  ColPass cy{spec, spec, nullptr, pitch, 1, (int64_t)ny * pitch, m + 1, 0, 0, 0, 0, 0};
  st = launch_cols(ny, kColFwd, cy, nx, tb, 0, s);
  if (st) return st;
  ColPass cx{spec, spec, nullptr, (int64_t)ny * pitch, 1, pitch, m + 1, 0, 0, 0, 0, 0};
  return launch_cols(nx, kColFwd, cx, ny, tb, 0, s);
}

// Inverse 3D c2r (1/N normalised); the spectrum is used as scratch.
int irfft3d_pitched(float2* spec, float* out, const int32_t* shape, int64_t pitch, const SpecTables& tb,
                    hipStream_t s) {
  const int nx = shape[0], ny = shape[1], nz = shape[2], m = nz / 2;
  ColPass cx{spec, spec, nullptr, (int64_t)ny * pitch, 1, pitch, m + 1, 0, 0, 0, 0, 0};
  int st = launch_cols(nx, kColInv, cx, ny, tb, 0, s);
  if (st) return st;
  ColPass cy{spec, spec, nullptr, pitch, 1, (int64_t)ny * pitch, m + 1, 0, 0, 0, 0, 0};
  st = launch_cols(ny, kColInv, cy, nx, tb, 0, s);
  if (st) return st;
  ZPass zp{};
  zp.rout = out;
  zp.cmain = spec;
  zp.cnyq = spec + m;
  zp.pitch_r = nz;
  zp.pitch_c = pitch;
  zp.nyq_stride = pitch;
  zp.nlines = (int64_t)nx * ny;
  zp.out_scale = (float)(2.0 / ((double)nx * ny * nz));
  return launch_z(nz, true, zp, tb, s);
}

}  // namespace fft
}  // namespace pm

using namespace pm;

extern "C" {

int pm_fft_workspace(const int32_t shape[3], size_t* bytes) {
  PM_RETURN_IF(shape == nullptr || bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL argument");
  PM_RETURN_IF(!fft::supported_shape(shape), PM_ERROR_UNSUPPORTED,
               "FFT axes must be powers of two in [2, 2048]");
  *bytes = 0;
  return PM_SUCCESS;
}

int pm_rfft3d(const float* in, float* out_complex, const int32_t shape[3], void* workspace,
              size_t workspace_bytes, void* stream) {
  (void)workspace;
  (void)workspace_bytes;
  PM_RETURN_IF(shape == nullptr || in == nullptr || out_complex == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "NULL argument");
  PM_RETURN_IF(!fft::supported_shape(shape), PM_ERROR_UNSUPPORTED,
               "FFT axes must be powers of two in [2, 2048]");
  int st = require_device_tables();
  if (st) return st;
  PM_RETURN_IF((const void*)in == (const void*)out_complex, PM_ERROR_INVALID_ARGUMENT,
               "pm_rfft3d needs distinct input and output buffers");
  fft::SpecTables tb{};
  return fft::rfft3d_pitched(in, reinterpret_cast<float2*>(out_complex), shape, shape[2] / 2 + 1, 0.f,
                             1.f, tb, as_stream(stream));
}

int pm_irfft3d(float* in_complex, float* out, const int32_t shape[3], void* workspace,
               size_t workspace_bytes, void* stream) {
  (void)workspace;
  (void)workspace_bytes;
  PM_RETURN_IF(shape == nullptr || in_complex == nullptr || out == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "NULL argument");
  PM_RETURN_IF(!fft::supported_shape(shape), PM_ERROR_UNSUPPORTED,
               "FFT axes must be powers of two in [2, 2048]");
  int st = require_device_tables();
  if (st) return st;
  PM_RETURN_IF((const void*)in_complex == (const void*)out, PM_ERROR_INVALID_ARGUMENT,
               "pm_irfft3d needs distinct input and output buffers");
  fft::SpecTables tb{};
  return fft::irfft3d_pitched(reinterpret_cast<float2*>(in_complex), out, shape, shape[2] / 2 + 1, tb,
                              as_stream(stream));
}

}  // extern "C"
