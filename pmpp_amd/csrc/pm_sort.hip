// Stable, deterministic cell sort (least-significant-digit radix, digits of up to 10 bits).
//
// Each pass is a stable counting sort of one digit:
//   1. k_digit_hist   one wave per contiguous segment of kSeg items counts its digits;
//   2. exclusive scan of the digit-major (digit, segment) count matrix;
//   3. k_digit_scatter the same waves walk their segment in order, rank each item among
//      equal digits of its 64-item group with ballot masks, and place it at
//      offset[digit][segment] + running count -- input order is preserved inside every
//      key, so the result equals numpy.argsort(keys, kind="stable") bit for bit.
// The key of the first pass is computed on the fly from (pmid, disp): the raveled brick
// index of the containing cell, with invalid particles mapped to the sentinel.
#include <algorithm>
#include <atomic>

#include "pm_common.h"
#include "pm_sort.h"

namespace pm {
namespace sort {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kIters = 128;               // 64-item groups per segment
constexpr int kSeg = kWave * kIters;      // items per wave segment
constexpr int kMaxDigitBits = 10;
constexpr int kMaxRadix = 1 << kMaxDigitBits;

struct KeyGeom {
  int n[3];
  int b[3];
  int nb[3];
  int nshift[3];  // log2(n) when n is a power of two, else -1
  int bshift[3];  // log2(b) when b is a power of two, else -1
  float cell;
  int sentinel;
};

__device__ __forceinline__ int cell_key(const KeyGeom& g, const int16_t* __restrict__ pmid,
                                        const float* __restrict__ disp,
                                        const uint8_t* __restrict__ valid, int64_t i) {
  if (valid && !valid[i]) return g.sentinel;
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // containing cell: the CIC base index with offset 0 (cic/pallas.py:227-236)
    float t = disp[i * 3 + a] / g.cell;
    int base = (int)pmid[i * 3 + a] + (int)floorf(t);
    int cell = g.nshift[a] >= 0 ? (base & (g.n[a] - 1)) : pos_mod(base, g.n[a]);
    c[a] = g.bshift[a] >= 0 ? (cell >> g.bshift[a]) : cell / g.b[a];
  }
  return (c[0] * g.nb[1] + c[1]) * g.nb[2] + c[2];
}

template <bool FIRST>
__global__ void __launch_bounds__(kWave * kWavesPerBlock)
    k_digit_hist(const int32_t* __restrict__ keys, const int16_t* __restrict__ pmid,
                 const float* __restrict__ disp, const uint8_t* __restrict__ valid, KeyGeom g,
                 int64_t n, int shift, int dbits, int64_t nseg, int32_t* __restrict__ hist) {
  __shared__ int32_t h[kWavesPerBlock][kMaxRadix];
  const int lane = threadIdx.x % kWave;
  const int w = threadIdx.x / kWave;
  const int radix = 1 << dbits;
  for (int d = lane; d < radix; d += kWave) h[w][d] = 0;
  __syncthreads();
  const int64_t seg = (int64_t)blockIdx.x * kWavesPerBlock + w;
  if (seg < nseg) {
    // equal digits of a 64-item group are counted by their lowest lane (ballot match), so the
    // wave-private row h[w] needs no atomics and nearly sorted keys do not serialise on one bank
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t start = seg * kSeg;
    for (int it = 0; it < kIters; ++it) {
      int64_t i = start + it * kWave + lane;
      bool active = i < n;
      int key = 0;
      if (active) key = FIRST ? cell_key(g, pmid, disp, valid, i) : keys[i];
      int d = (key >> shift) & (radix - 1);
      uint64_t peers = __ballot(active);
      for (int bit = 0; bit < dbits; ++bit) {
        bool set = (d >> bit) & 1;
        uint64_t m = __ballot(set);
        peers &= set ? m : ~m;
      }
      if (active && (peers & lt_mask) == 0) h[w][d] += __popcll(peers);
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  if (seg < nseg)
    for (int d = lane; d < radix; d += kWave) hist[(int64_t)d * nseg + seg] = h[w][d];
}

template <bool FIRST>
__global__ void __launch_bounds__(kWave * kWavesPerBlock)
    k_digit_scatter(const int32_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in,
                    const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                    const uint8_t* __restrict__ valid, KeyGeom g, int64_t n, int shift, int dbits,
                    int64_t nseg, const int32_t* __restrict__ offsets, int32_t* __restrict__ keys_out,
                    int32_t* __restrict__ vals_out) {
  __shared__ int32_t cnt[kWavesPerBlock][kMaxRadix];
  const int lane = threadIdx.x % kWave;
  const int w = threadIdx.x / kWave;
  const int radix = 1 << dbits;
  const int64_t seg = (int64_t)blockIdx.x * kWavesPerBlock + w;
  if (seg >= nseg) return;  // whole wave exits together; no block barriers below
  for (int d = lane; d < radix; d += kWave) cnt[w][d] = offsets[(int64_t)d * nseg + seg];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t start = seg * kSeg;
  for (int it = 0; it < kIters; ++it) {
    int64_t i = start + it * kWave + lane;
    bool active = i < n;
    int key = 0, val = 0;
    if (active) {
      if (FIRST) {
        key = cell_key(g, pmid, disp, valid, i);
        val = (int)i;
      } else {
        key = keys_in[i];
        val = vals_in[i];
      }
    }
    int d = (key >> shift) & (radix - 1);
    uint64_t peers = __ballot(active);
    for (int bit = 0; bit < dbits; ++bit) {
      bool set = (d >> bit) & 1;
      uint64_t m = __ballot(set);
      peers &= set ? m : ~m;
    }
    int rank = __popcll(peers & lt_mask);
    int base = 0;
    if (active) base = cnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (active) {
      int pos = base + rank;
      if (keys_out) keys_out[pos] = key;
      vals_out[pos] = val;
      if (rank == 0) cnt[w][d] = base + __popcll(peers);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
}

// ---- exclusive scan of int32 counts (3 phases) ----
constexpr int kScanBlock = 256;
constexpr int kScanPer = 16;
constexpr int kScanChunk = kScanBlock * kScanPer;

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* sh, int32_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int32_t a = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[kScanBlock - 1];
  int32_t ex = sh[t] - v;
  __syncthreads();
  return ex;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t len,
                                                            int32_t* __restrict__ partial) {
  __shared__ int32_t sh[kScanBlock];
  int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int32_t s = 0;
  for (int j = 0; j < kScanPer; ++j)
    if (base + j < len) s += in[base + j];
  int32_t total;
  block_exclusive_scan(s, sh, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_partials(int32_t* __restrict__ partial, int64_t len) {
  __shared__ int32_t sh[kScanBlock];
  int32_t carry = 0;
  for (int64_t base = 0; base < len; base += kScanBlock) {
    int64_t i = base + threadIdx.x;
    int32_t v = i < len ? partial[i] : 0;
    int32_t total;
    int32_t ex = block_exclusive_scan(v, sh, total);
    if (i < len) partial[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const int32_t* __restrict__ in, int64_t len,
                                                           const int32_t* __restrict__ partial,
                                                           int32_t* __restrict__ out) {
  __shared__ int32_t sh[kScanBlock];
  int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int32_t vals[kScanPer];
  int32_t s = 0;
  for (int j = 0; j < kScanPer; ++j) {
    vals[j] = (base + j < len) ? in[base + j] : 0;
    s += vals[j];
  }
  int32_t total;
  int32_t ex = block_exclusive_scan(s, sh, total) + partial[blockIdx.x];
  for (int j = 0; j < kScanPer; ++j) {
    if (base + j < len) out[base + j] = ex;
    ex += vals[j];
  }
}

// bin_offsets[b] = first sorted position with key >= b, for b in [0, num_keys]
__global__ void k_bin_bounds(const int32_t* __restrict__ keys, int64_t n, int num_keys,
                             int32_t* __restrict__ bounds) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  int lo = (j == 0) ? 0 : keys[j - 1] + 1;
  int hi = (j == n) ? num_keys : keys[j];
  for (int b = lo; b <= hi; ++b) bounds[b] = (int32_t)j;
}

// sorted payload: pmid_s[j] = pmid[perm[j]], disp_s[j] = disp[perm[j]] (either may be null)
__global__ void k_gather_payload(const int32_t* __restrict__ perm, const int16_t* __restrict__ pmid,
                                 const float* __restrict__ disp, int64_t n, int16_t* __restrict__ pmid_s,
                                 float* __restrict__ disp_s) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = perm[j];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (pmid_s) pmid_s[j * 3 + a] = pmid[i * 3 + a];
      if (disp_s) disp_s[j * 3 + a] = disp[i * 3 + a];
    }
  }
}

int scan_exclusive(const int32_t* in, int32_t* out, int64_t len, int32_t* partial, hipStream_t s) {
  if (len == 0) return PM_SUCCESS;
  int64_t blocks = ceil_div(len, kScanChunk);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)blocks), dim3(kScanBlock), 0, s, in, len, partial);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, s, partial, blocks);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)blocks), dim3(kScanBlock), 0, s, in, len, partial, out);
  PM_LAUNCHED("scan", s);
  return PM_SUCCESS;
}

KeyGeom make_key_geom(const int32_t* mesh_shape, float cell, const int32_t* brick) {
  KeyGeom g;
  int64_t nbins = 1;
  auto log2_or_neg = [](int v) {
    if (v <= 0 || (v & (v - 1))) return -1;
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
  };
  for (int a = 0; a < 3; ++a) {
    g.n[a] = mesh_shape[a];
    g.b[a] = brick[a];
    g.nb[a] = (int)ceil_div(mesh_shape[a], brick[a]);
    g.nshift[a] = log2_or_neg(mesh_shape[a]);
    g.bshift[a] = log2_or_neg(brick[a]);
    nbins *= g.nb[a];
  }
  g.cell = cell;
  g.sentinel = (int)nbins;
  return g;
}

// ---- unstable bucketing by key (the gravity step's binning) ----
// The order inside a bucket depends on atomic arrival order, which the tile kernels do not
// see: the scatter sums in exact integer arithmetic and the gather is per particle.
__device__ __forceinline__ int raw_cell_key(const KeyGeom& g, const int p[3], const float d[3], bool ok) {
  if (!ok) return g.sentinel;
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t = d[a] / g.cell;  // the same float32 arithmetic as cell_key and the CIC base cell
    int base = p[a] + (int)floorf(t);
    int cell = g.nshift[a] >= 0 ? (base & (g.n[a] - 1)) : pos_mod(base, g.n[a]);
    c[a] = g.bshift[a] >= 0 ? (cell >> g.bshift[a]) : cell / g.b[a];
  }
  return (c[0] * g.nb[1] + c[1]) * g.nb[2] + c[2];
}

// Bucketing in chunks of 2048 particles per workgroup: keys are aggregated in an LDS hash
// table (2048 slots, so the linear probe always ends), then one global atomic per distinct key
// and chunk reserves that key's range, and the chunk's particles of one key are written
// contiguously by one workgroup.  When the particles are a whole Lagrangian grid in C order
// (the usual PM state), a chunk is a 4 x 8 x 64 block of that grid, which falls into a few
// bricks, so a chunk needs only a handful of global atomics and its writes fill cache lines;
// otherwise chunks are contiguous index ranges.  The mapping only affects speed: every
// partition of the particles gives a valid (unstable) bucketing.
constexpr int kBkThreads = 256;
constexpr int kBkChunk = 2048;
constexpr int kBkPer = kBkChunk / kBkThreads;  // 8 particles per thread
constexpr int kBkSlots = 2048;                // >= particles per chunk
constexpr int kRankBits = 11;

struct ChunkMap {
  int64_t n;
  int64_t nchunks;
  int tiled;     // 1: 4 x 8 x 64 blocks of a (g0, g1, g2) particle grid
  int g1, g2;    // grid extents (tiled)
  int nby, nbz;  // blocks along y and z (tiled)
};

// particle index of slot e of thread t in chunk c
__device__ __forceinline__ int64_t chunk_index(const ChunkMap& m, int64_t c, int e, int t) {
  if (!m.tiled) return c * kBkChunk + (int64_t)e * kBkThreads + t;
  const int bz = (int)(c % m.nbz);
  const int64_t r = c / m.nbz;
  const int by = (int)(r % m.nby);
  const int bx = (int)(r / m.nby);
  const int row = (t >> 6) + 4 * e;  // 32 rows of 64 particles; lane = z offset
  const int ix = bx * 4 + (row >> 3), iy = by * 8 + (row & 7);
  return ((int64_t)ix * m.g1 + iy) * m.g2 + (int64_t)bz * 64 + (t & 63);
}

__device__ __forceinline__ unsigned slot_hash(int key) { return ((unsigned)key * 2654435761u) >> (32 - 11); }

template <bool PLACE>
__global__ void __launch_bounds__(kBkThreads) k_bucket(const int16_t* __restrict__ pmid, const float* __restrict__ disp,
                                                       const uint8_t* __restrict__ valid, KeyGeom g, ChunkMap m,
                                                       int32_t* __restrict__ count, const int32_t* __restrict__ bins,
                                                       int32_t* __restrict__ perm, int16_t* __restrict__ pmid_s,
                                                       float* __restrict__ disp_s, uint8_t* __restrict__ valid_s) {
  __shared__ int s_key[kBkSlots];
  __shared__ int s_cnt[kBkSlots];  // count per slot, then (PLACE) the slot's first destination
  const int t = threadIdx.x;
  for (int h = t; h < kBkSlots; h += kBkThreads) {
    s_key[h] = -1;
    s_cnt[h] = 0;
  }
  __syncthreads();
  for (int64_t c = blockIdx.x; c < m.nchunks; c += gridDim.x) {
    int code[kBkPer];
    // the placing pass keeps each particle's record in registers between its two phases
    // (pmid packed as 2 x 16 bits + 16 bits with the valid flag, disp as 3 floats)
    uint32_t pk0[PLACE ? kBkPer : 1], pk1[PLACE ? kBkPer : 1];
    float pd[PLACE ? kBkPer : 1][3];
#pragma unroll
    for (int e = 0; e < kBkPer; ++e) {
      const int64_t i = chunk_index(m, c, e, t);
      const bool act = i < m.n;
      const int64_t ic = act ? i : m.n - 1;  // clamped: unconditional loads
      int p[3];
      float d[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        p[a] = pmid[ic * 3 + a];
        d[a] = disp[ic * 3 + a];
      }
      const uint8_t v = valid ? valid[ic] : (uint8_t)1;
      const int key = raw_cell_key(g, p, d, v != 0);
      if constexpr (PLACE) {
        pk0[e] = ((uint32_t)p[0] & 0xffffu) | ((uint32_t)p[1] << 16);
        pk1[e] = ((uint32_t)p[2] & 0xffffu) | ((uint32_t)v << 16);
#pragma unroll
        for (int a = 0; a < 3; ++a) pd[e][a] = d[a];
      }
      code[e] = -1;
      if (act) {
        unsigned h = slot_hash(key);
        for (;;) {
          const int prev = atomicCAS(&s_key[h], -1, key);
          if (prev == -1 || prev == key) break;
          h = (h + 1) & (kBkSlots - 1);
        }
        const int r = atomicAdd(&s_cnt[h], 1);
        code[e] = ((int)h << kRankBits) | r;
      }
    }
    __syncthreads();
    for (int h = t; h < kBkSlots; h += kBkThreads) {
      const int k = s_key[h];
      if (k >= 0) {
        const int cnt = s_cnt[h];
        if (PLACE)
          s_cnt[h] = bins[k] + atomicAdd(&count[k], cnt);
        else
          atomicAdd(&count[k], cnt);
      }
    }
    __syncthreads();
    if constexpr (PLACE) {
#pragma unroll
      for (int e = 0; e < kBkPer; ++e) {
        if (code[e] < 0) continue;
        const int64_t i = chunk_index(m, c, e, t);
        const int dest = s_cnt[code[e] >> kRankBits] + (code[e] & ((1 << kRankBits) - 1));
        perm[dest] = (int32_t)i;
        pmid_s[(int64_t)dest * 3 + 0] = (int16_t)(pk0[e] & 0xffffu);
        pmid_s[(int64_t)dest * 3 + 1] = (int16_t)(pk0[e] >> 16);
        pmid_s[(int64_t)dest * 3 + 2] = (int16_t)(pk1[e] & 0xffffu);
#pragma unroll
        for (int a = 0; a < 3; ++a) disp_s[(int64_t)dest * 3 + a] = pd[e][a];
        valid_s[dest] = (uint8_t)(pk1[e] >> 16);
      }
    }
    __syncthreads();
    for (int h = t; h < kBkSlots; h += kBkThreads) {
      s_key[h] = -1;
      s_cnt[h] = 0;
    }
    __syncthreads();
  }
}

ChunkMap make_chunk_map(int64_t n, const int32_t* grid) {
  ChunkMap m{};
  m.n = n;
  if (grid && (int64_t)grid[0] * grid[1] * grid[2] == n && grid[0] % 4 == 0 && grid[1] % 8 == 0 &&
      grid[2] % 64 == 0) {
    m.tiled = 1;
    m.g1 = grid[1];
    m.g2 = grid[2];
    m.nby = grid[1] / 8;
    m.nbz = grid[2] / 64;
    m.nchunks = (int64_t)(grid[0] / 4) * m.nby * m.nbz;
  } else {
    m.nchunks = ceil_div(n, kBkChunk);
  }
  return m;
}

__global__ void k_zero_i32(int32_t* __restrict__ p, int64_t len) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

int key_bits(int64_t max_key) {
  int b = 1;
  while ((int64_t(1) << b) <= max_key) ++b;
  return b;
}

}  // namespace

size_t bucket_workspace_bytes(const int32_t* mesh_shape, const int32_t* brick) {
  int64_t nbins = 1;
  for (int a = 0; a < 3; ++a) nbins *= ceil_div(mesh_shape[a], brick[a]);
  const int64_t len = nbins + 2;
  return align_up(sizeof(int32_t) * len, 256) + align_up(sizeof(int32_t) * (ceil_div(len, kScanChunk) + 1), 256) + 512;
}

int bucket_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n, const int32_t* mesh_shape,
                 float cell_size, const int32_t* brick, const int32_t* ptcl_grid, int32_t* perm, int32_t* bin_offsets,
                 int16_t* pmid_s, float* disp_s, uint8_t* valid_s, Arena& ws, hipStream_t s) {
  KeyGeom g = make_key_geom(mesh_shape, cell_size, brick);
  const int64_t len = (int64_t)g.sentinel + 2;  // bricks, the invalid bucket, and the end
  int32_t* count = ws.take<int32_t>(len);
  int32_t* partial = ws.take<int32_t>(ceil_div(len, kScanChunk) + 1);
  PM_RETURN_IF(!ws.ok(), PM_ERROR_WORKSPACE, "bucket workspace too small");
  const unsigned zb = (unsigned)std::min<int64_t>(ceil_div(len, 256), 1024);
  hipLaunchKernelGGL(k_zero_i32, dim3(zb), dim3(256), 0, s, count, len);
  PM_LAUNCHED("k_zero_i32", s);
  const ChunkMap m = make_chunk_map(n, ptcl_grid);
  const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(m.nchunks, 4096));
  if (n > 0) {
    hipLaunchKernelGGL(k_bucket<false>, dim3(nb), dim3(kBkThreads), 0, s, pmid, disp, valid, g, m, count,
                       (const int32_t*)nullptr, (int32_t*)nullptr, (int16_t*)nullptr, (float*)nullptr,
                       (uint8_t*)nullptr);
    PM_LAUNCHED("k_bucket_count", s);
  }
  PM_TRY(scan_exclusive(count, bin_offsets, len, partial, s));
  hipLaunchKernelGGL(k_zero_i32, dim3(zb), dim3(256), 0, s, count, len);
  PM_LAUNCHED("k_zero_i32", s);
  if (n > 0) {
    hipLaunchKernelGGL(k_bucket<true>, dim3(nb), dim3(kBkThreads), 0, s, pmid, disp, valid, g, m, count,
                       (const int32_t*)bin_offsets, perm, pmid_s, disp_s, valid_s);
    PM_LAUNCHED("k_bucket_place", s);
  }
  return PM_SUCCESS;
}

int check_sort_args(int64_t n, const int32_t* mesh_shape, const int32_t* brick) {
  PM_RETURN_IF(n < 0 || n > INT32_MAX, PM_ERROR_UNSUPPORTED, "particle count must be in [0, 2^31)");
  PM_RETURN_IF(mesh_shape == nullptr || brick == nullptr, PM_ERROR_INVALID_ARGUMENT, "NULL shape");
  int64_t nbins = 1;
  for (int a = 0; a < 3; ++a) {
    PM_RETURN_IF(mesh_shape[a] < 1 || brick[a] < 1, PM_ERROR_INVALID_ARGUMENT,
                 "mesh_shape and brick must be positive");
    nbins *= ceil_div(mesh_shape[a], brick[a]);
  }
  PM_RETURN_IF(nbins >= INT32_MAX, PM_ERROR_UNSUPPORTED,
               "more than 2^31-1 bricks: use a coarser brick for this mesh");
  return PM_SUCCESS;
}

size_t workspace_bytes(int64_t n, const int32_t* mesh_shape, const int32_t* brick) {
  int64_t nseg = ceil_div(n, kSeg);
  int64_t hist = (int64_t)kMaxRadix * nseg;
  size_t b = 0;
  b += align_up(sizeof(int32_t) * n, 256) * 3;  // keys a/b, vals
  b += align_up(sizeof(int32_t) * hist, 256) * 2;
  b += align_up(sizeof(int32_t) * (ceil_div(hist, kScanChunk) + 1), 256);
  (void)mesh_shape;
  (void)brick;
  return b + 256;
}

int digit_bits_for(int bits, int& passes) {
  passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
  return (bits + passes - 1) / passes;
}

int sort_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
               const int32_t* mesh_shape, float cell_size, const int32_t* brick, int32_t* perm,
               int32_t* bin_offsets, int16_t* pmid_s, float* disp_s, Arena& ws, hipStream_t s) {
  KeyGeom g = make_key_geom(mesh_shape, cell_size, brick);
  if (n == 0) {
    if (bin_offsets) {
      hipLaunchKernelGGL(k_bin_bounds, dim3(1), dim3(1), 0, s, (const int32_t*)nullptr, (int64_t)0,
                         g.sentinel + 1, bin_offsets);
      PM_LAUNCHED("k_bin_bounds", s);
    }
    return PM_SUCCESS;
  }
  const int64_t nseg = ceil_div(n, kSeg);
  // without a validity mask no particle takes the sentinel, so the key range ends at nbins-1
  const int bits = key_bits(valid ? g.sentinel : std::max(g.sentinel - 1, 1));
  int passes = 1;
  const int dbits = digit_bits_for(bits, passes);
  const int64_t hist_len = ((int64_t)1 << dbits) * nseg;
  int32_t* keys_a = ws.take<int32_t>(n);
  int32_t* keys_b = ws.take<int32_t>(n);
  int32_t* vals_t = ws.take<int32_t>(n);
  int32_t* hist = ws.take<int32_t>(hist_len);
  int32_t* offs = ws.take<int32_t>(hist_len);
  int32_t* partial = ws.take<int32_t>(ceil_div(hist_len, kScanChunk) + 1);
  PM_RETURN_IF(!ws.ok(), PM_ERROR_WORKSPACE, "sort workspace too small");

  const dim3 grid((unsigned)ceil_div(nseg, kWavesPerBlock)), block(kWave * kWavesPerBlock);
  // value ping-pong arranged so the last pass writes `perm`
  int32_t* vbuf[2] = {(passes % 2 == 1) ? perm : vals_t, (passes % 2 == 1) ? vals_t : perm};
  int32_t* kbuf[2] = {keys_a, keys_b};
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = dbits * pass;
    const bool last = pass == passes - 1;
    int32_t* kin = kbuf[(pass + 1) % 2];
    int32_t* kout = (last && !bin_offsets) ? nullptr : kbuf[pass % 2];
    int32_t* vin = vbuf[(pass + 1) % 2];
    int32_t* vout = vbuf[pass % 2];
    if (pass == 0)
      hipLaunchKernelGGL(k_digit_hist<true>, grid, block, 0, s, nullptr, pmid, disp, valid, g, n, shift,
                         dbits, nseg, hist);
    else
      hipLaunchKernelGGL(k_digit_hist<false>, grid, block, 0, s, kin, pmid, disp, valid, g, n, shift,
                         dbits, nseg, hist);
    PM_LAUNCHED("k_digit_hist", s);
    int st = scan_exclusive(hist, offs, hist_len, partial, s);
    if (st) return st;
    if (pass == 0)
      hipLaunchKernelGGL(k_digit_scatter<true>, grid, block, 0, s, nullptr, nullptr, pmid, disp, valid, g,
                         n, shift, dbits, nseg, offs, kout, vout);
    else
      hipLaunchKernelGGL(k_digit_scatter<false>, grid, block, 0, s, kin, vin, pmid, disp, valid, g, n,
                         shift, dbits, nseg, offs, kout, vout);
    PM_LAUNCHED("k_digit_scatter", s);
    if (last && bin_offsets) {
      hipLaunchKernelGGL(k_bin_bounds, dim3((unsigned)ceil_div(n + 1, 256)), dim3(256), 0, s, kout, n,
                         g.sentinel + 1, bin_offsets);
      PM_LAUNCHED("k_bin_bounds", s);
    }
  }
  if (pmid_s || disp_s) {
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(k_gather_payload, dim3(blocks), dim3(256), 0, s, perm, pmid, disp, n, pmid_s, disp_s);
    PM_LAUNCHED("k_gather_payload", s);
  }
  return PM_SUCCESS;
}

// ------------------------------------------------ allocation-free binning ----
// pm_bucket_cells_noalloc takes its bin counters and scan partials from this file-scope
// device array instead of a caller workspace.  On the host the array's name is only a
// stand-in, so host code never passes it to a kernel or a copy: pm_init resolves the
// array's address on each device once (hipGetSymbolAddress, with that device current) and
// the entry point uses the address of the current device.  Every call on a device shares
// that device's copy, so calls that overlap on one device (two streams) race.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;  // 8 MiB: up to 2,096,446 bricks
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::atomic<void*> g_scratch_ptr[kMaxDevices];

int init_noalloc_scratch() {
  int dev = 0;
  PM_HIP_TRY(hipGetDevice(&dev));
  PM_RETURN_IF(dev < 0 || dev >= kMaxDevices, PM_ERROR_UNSUPPORTED, "device ordinal >= 64");
  if (g_scratch_ptr[dev].load(std::memory_order_acquire)) return PM_SUCCESS;
  void* p = nullptr;
  PM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PM_RETURN_IF(p == nullptr, PM_ERROR_HIP, "hipGetSymbolAddress returned NULL for the binning scratch");
  g_scratch_ptr[dev].store(p, std::memory_order_release);
  return PM_SUCCESS;
}

int check_bucket_args(const int16_t* pmid, const float* disp, int64_t n, const int32_t* mesh_shape, float cell_size,
                      const int32_t* brick, const int32_t* ptcl_grid, const int32_t* perm, const int32_t* bin_offsets,
                      const int16_t* pmid_s, const float* disp_s, const uint8_t* valid_s) {
  int st = check_sort_args(n, mesh_shape, brick);
  if (st) return st;
  PM_RETURN_IF(!(cell_size > 0.f), PM_ERROR_INVALID_ARGUMENT, "cell_size must be positive");
  PM_RETURN_IF(bin_offsets == nullptr, PM_ERROR_INVALID_ARGUMENT, "bin_offsets NULL");
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || perm == nullptr || pmid_s == nullptr ||
                         disp_s == nullptr || valid_s == nullptr),
               PM_ERROR_INVALID_ARGUMENT, "pmid/disp/perm/sorted payload NULL");
  if (ptcl_grid)
    for (int a = 0; a < 3; ++a)
      PM_RETURN_IF(ptcl_grid[a] < 0, PM_ERROR_INVALID_ARGUMENT, "ptcl_grid must be non-negative");
  return PM_SUCCESS;
}

}  // namespace sort
}  // namespace pm

using namespace pm;

extern "C" {

int pm_sort_cells_workspace(int64_t n, const int32_t mesh_shape[3], const int32_t brick[3],
                            size_t* bytes) {
  int st = sort::check_sort_args(n, mesh_shape, brick);
  if (st) return st;
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = sort::workspace_bytes(n, mesh_shape, brick);
  return PM_SUCCESS;
}

int pm_sort_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                  const int32_t mesh_shape[3], float cell_size, const int32_t brick[3],
                  int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted, float* disp_sorted,
                  void* workspace, size_t workspace_bytes, void* stream) {
  int st = sort::check_sort_args(n, mesh_shape, brick);
  if (st) return st;
  PM_RETURN_IF(!(cell_size > 0.f), PM_ERROR_INVALID_ARGUMENT, "cell_size must be positive");
  PM_RETURN_IF(n > 0 && (pmid == nullptr || disp == nullptr || perm == nullptr),
               PM_ERROR_INVALID_ARGUMENT, "pmid/disp/perm NULL");
  size_t need = sort::workspace_bytes(n, mesh_shape, brick);
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < need, PM_ERROR_WORKSPACE,
               "sort workspace too small: need " + std::to_string(need) + " bytes");
  Arena ws(workspace, workspace_bytes);
  return sort::sort_cells(pmid, disp, valid, n, mesh_shape, cell_size, brick, perm, bin_offsets, pmid_sorted,
                          disp_sorted, ws, as_stream(stream));
}

int pm_bucket_cells_workspace(int64_t n, const int32_t mesh_shape[3], const int32_t brick[3], size_t* bytes) {
  int st = sort::check_sort_args(n, mesh_shape, brick);
  if (st) return st;
  PM_RETURN_IF(bytes == nullptr, PM_ERROR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = sort::bucket_workspace_bytes(mesh_shape, brick);
  return PM_SUCCESS;
}

int pm_bucket_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                    const int32_t mesh_shape[3], float cell_size, const int32_t brick[3], const int32_t ptcl_grid[3],
                    int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted, float* disp_sorted,
                    uint8_t* valid_sorted, void* workspace, size_t workspace_bytes, void* stream) {
  int st = sort::check_bucket_args(pmid, disp, n, mesh_shape, cell_size, brick, ptcl_grid, perm, bin_offsets,
                                   pmid_sorted, disp_sorted, valid_sorted);
  if (st) return st;
  const size_t need = sort::bucket_workspace_bytes(mesh_shape, brick);
  PM_RETURN_IF(workspace == nullptr || workspace_bytes < need, PM_ERROR_WORKSPACE,
               "bucket workspace too small: need " + std::to_string(need) + " bytes");
  Arena ws(workspace, workspace_bytes);
  return sort::bucket_cells(pmid, disp, valid, n, mesh_shape, cell_size, brick, ptcl_grid, perm, bin_offsets,
                            pmid_sorted, disp_sorted, valid_sorted, ws, as_stream(stream));
}

int pm_bucket_cells_noalloc(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                            const int32_t mesh_shape[3], float cell_size, const int32_t brick[3],
                            const int32_t ptcl_grid[3], int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted,
                            float* disp_sorted, uint8_t* valid_sorted, void* stream) {
  int st = sort::check_bucket_args(pmid, disp, n, mesh_shape, cell_size, brick, ptcl_grid, perm, bin_offsets,
                                   pmid_sorted, disp_sorted, valid_sorted);
  if (st) return st;
  const size_t need = sort::bucket_workspace_bytes(mesh_shape, brick);
  PM_RETURN_IF(need > sort::kNoAllocScratchBytes, PM_ERROR_UNSUPPORTED,
               "pm_bucket_cells_noalloc: " + std::to_string(need) + " bytes of bin counters exceed its " +
                   std::to_string(sort::kNoAllocScratchBytes) +
                   "-byte static scratch (at most 2,096,446 bricks): use pm_bucket_cells with a workspace");
  int dev = 0;
  PM_HIP_TRY(hipGetDevice(&dev));
  void* p = (dev >= 0 && dev < sort::kMaxDevices) ? sort::g_scratch_ptr[dev].load(std::memory_order_acquire)
                                                  : nullptr;
  PM_RETURN_IF(p == nullptr, PM_ERROR_INVALID_ARGUMENT,
               "pm_init has not run on the current device (it resolves the binning scratch)");
  Arena ws(p, sort::kNoAllocScratchBytes);
  return sort::bucket_cells(pmid, disp, valid, n, mesh_shape, cell_size, brick, ptcl_grid, perm, bin_offsets,
                            pmid_sorted, disp_sorted, valid_sorted, ws, as_stream(stream));
}

}  // extern "C"
