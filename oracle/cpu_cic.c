/*
 * Float32 CIC mass assignment and interpolation on the host cores (OpenMP): the CIC stages of
 * the timed CPU baseline (oracle/cpu_baseline.py).  Test/measurement infrastructure only.
 *
 * The arithmetic is the single-device CIC of the reference with offset 0 and the particle
 * cell size (src/pmpp/cic/enmesh.py:126-147, src/pmpp/cic/pallas.py:227-258): t = disp / cell,
 * base = pmid + floor(t) wrapped mod N, f = t - floor(t), corner c takes the bit (c >> a) & 1
 * per axis and the weight prod_a (bit ? f_a : 1 - f_a) (enmesh.py:98).  The scatter adds with
 * atomics (the order of float additions varies between runs, as in the reference's GPU scatter).
 */
#include <math.h>
#include <stdint.h>

static inline int64_t wrap_index(int64_t v, int64_t n) {
  int64_t r = v % n;
  return r < 0 ? r + n : r;
}

static inline void particle_cell(const int16_t* pmid, const float* disp, int64_t p, const int32_t* shape,
                                 float cell, int64_t base[3], float f[3]) {
  for (int a = 0; a < 3; ++a) {
    const float t = disp[p * 3 + a] / cell;
    const float lf = floorf(t);
    base[a] = wrap_index((int64_t)pmid[p * 3 + a] + (int64_t)lf, shape[a]);
    f[a] = t - lf;
  }
}

static inline void corner(const int64_t base[3], const float f[3], const int32_t* shape, int c, int64_t* flat,
                          float* w) {
  int64_t idx[3];
  float wt = 1.f;
  for (int a = 0; a < 3; ++a) {
    const int bit = (c >> a) & 1;
    idx[a] = base[a] + bit;
    if (idx[a] == shape[a]) idx[a] = 0;
    wt = wt * (bit ? f[a] : 1.f - f[a]);
  }
  *flat = (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2];
  *w = wt;
}

/* mesh += val * W (the mesh is caller-zeroed) */
void cpu_cic_scatter(const int16_t* pmid, const float* disp, int64_t n, const int32_t* shape, float cell,
                     float val, float* mesh) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    int64_t base[3];
    float f[3];
    particle_cell(pmid, disp, p, shape, cell, base, f);
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      float w;
      corner(base, f, shape, c, &flat, &w);
#pragma omp atomic
      mesh[flat] += val * w;
    }
  }
}

/* out[p * stride + comp] = sum_c W_c mesh[corner c] */
void cpu_cic_gather(const int16_t* pmid, const float* disp, int64_t n, const int32_t* shape, float cell,
                    const float* mesh, float* out, int32_t stride, int32_t comp) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    int64_t base[3];
    float f[3];
    particle_cell(pmid, disp, p, shape, cell, base, f);
    float s = 0.f;
    for (int c = 0; c < 8; ++c) {
      int64_t flat;
      float w;
      corner(base, f, shape, c, &flat, &w);
      s = s + mesh[flat] * w;
    }
    out[p * stride + comp] = s;
  }
}
