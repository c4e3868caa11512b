/*
 * pmpp_amd -- C ABI of the MI355X (gfx950) particle-mesh force-step library.
 *
 * Every entry point is extern "C", takes caller-owned device pointers plus plain
 * sizes, enqueues work on the caller's HIP stream (passed as void*, a hipStream_t),
 * never synchronises (pm_init excepted), never allocates device memory and returns a
 * status code; the text of the last error on the calling host thread is available from
 * pm_last_error().
 * Temporaries come only from a caller-provided workspace whose size is reported by the
 * matching *_workspace() query.  This is the ownership/threading contract of the
 * reference's native FFI handlers (cuda/route_kernels.cu:1-7, 738-800; XLA owns every
 * buffer and the handler only enqueues on the supplied stream).
 *
 * Layouts (all row-major, C order):
 *   pmid      (n,3) int16   particle mesh anchors      (nbody/particles.py:33-127)
 *   disp      (n,3) float32 physical displacements     (nbody/particles.py:33-127)
 *   valid     (n)   uint8   1 = active slot, 0 = padding; NULL = all active
 *   mesh      (X,Y,Z[,C]) float32
 *   spectrum  (X,Y,Z/2+1) complex64 interleaved (re,im), the layout of rfftn
 */
#ifndef PMPP_AMD_H
#define PMPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_ABI_VERSION 9

enum pm_status {
  PM_SUCCESS = 0,
  PM_ERROR_INVALID_ARGUMENT = 1, /* ffi::Error::InvalidArgument analogue */
  PM_ERROR_UNSUPPORTED = 2,      /* shape outside what the kernels implement */
  PM_ERROR_WORKSPACE = 3,        /* workspace missing or too small */
  PM_ERROR_HIP = 4               /* a HIP runtime call or launch failed (ffi::Error::Internal) */
};

/* CIC geometry shared by the scatter/gather kernels.
 * Mirrors the keyword arguments of the reference kernel seam
 * pallas_scatter / pallas_gather (src/pmpp/cic/pallas.py:412-478):
 *   local_shape  = mesh.shape[:3]          (corners outside it are dropped, pallas.py:261-265)
 *   global_shape = global_shape            (periodic wrap of the particle mesh)
 *   offset       = offset (physical units, per axis; pallas.py:107-110)
 *   particle_cell= particle_cell_size      (conf.cell_size, a1)
 *   mesh_cell    = cell_size               (explicit a2, <= 0 means None; pallas.py:192-223) */
typedef struct pm_cic_geom {
  int32_t local_shape[3];
  int32_t global_shape[3];
  float offset[3];
  float particle_cell;
  float mesh_cell;
} pm_cic_geom;

/* Description of one PM force evaluation (src/pmpp/nbody/gravity.py:786-814).
 *   mesh_shape      conf.mesh_shape
 *   cell_size       conf.cell_size                       (core/configuration.py:463-466)
 *   ptcl_spacing    conf.ptcl_spacing (particle-Nyquist mask, core/utils.py:15-37)
 *   mask_active     1 when conf.mesh_shape != conf.ptcl_grid_shape (utils.py:31-32)
 *   omega_m         cosmo.Omega_m (cast to float32 as at gravity.py:630)
 *   particle_value  conf.mesh_size / conf.ptcl_num       (cic/scatter.py:258-265)
 *   brick           cell-sort granularity (cells per brick along each axis) */
typedef struct pm_gravity_desc {
  int32_t mesh_shape[3];
  float cell_size;
  float ptcl_spacing;
  int32_t mask_active;
  float omega_m;
  float particle_value;
  int32_t brick[3];
} pm_gravity_desc;

const char* pm_last_error(void);
int pm_abi_version(void);
/* Prepares the per-device constant tables (the FFT twiddle table) on the current device and
 * synchronises it: the only entry point that does either.  It also resolves the current
 * device's address of the static scratch that pm_bucket_cells_noalloc uses.  It must run once per device
 * before any entry point that transforms (pm_rfft3d, pm_irfft3d, the pm_gravity family);
 * those return PM_ERROR_INVALID_ARGUMENT until it has.  Thread-safe; later calls return at
 * once.  (The reference's FFI library has no such step: XLA's FFT needs none.) */
int pm_init(void);

/* Per-kernel timing for measurement harnesses (no counterpart in the reference, whose
 * only timer is timeit around a jitted call, core/utils.py:347-374).  pm_timing_enable(n)
 * creates n HIP events for the calling host thread (0 disables); afterwards every library
 * kernel launch on that thread records the next event on its stream, and pm_timing_mark
 * records one explicitly.  pm_timing_read waits for the last event and returns, for each
 * recorded event after the first, its name and the milliseconds since the previous one. */
int pm_timing_enable(int32_t max_events);
int pm_timing_mark(const char* name, void* stream);
int pm_timing_read(float* ms, const char** names, int32_t capacity, int32_t* count);

/* ---------------------------------------------------------------- CIC ----
 * pm_cic_scatter   replaces pallas_scatter       (src/pmpp/cic/pallas.py:444-478)
 *                  and the JAX fallback _scatter  (src/pmpp/cic/scatter.py:278-338).
 *   mesh += sum_p val_p W_p ; val is (n,C) when non-NULL, else the scalar val_scalar.
 * pm_cic_gather    replaces pallas_gather        (pallas.py:412-441, gather.py:414-472).
 *   out (n,C) = sum_corners W mesh ; invalid particles read 0.
 * pm_cic_gather_bwd replaces pallas_gather_bwd   (pallas.py:481-519, gather.py:475-549).
 *   disp_cot (n,3) is overwritten; mesh_cot is accumulated into (caller zeroes it).
 * pm_cic_scatter_bwd replaces pallas_scatter_bwd (pallas.py:522-571, scatter.py:341-437).
 *   disp_cot (n,3) overwritten; val_cot (n,C) overwritten when val != NULL, otherwise
 *   val_cot[0] is accumulated into (the scalar-value cotangent; caller zeroes it). */
int pm_cic_scatter(const int16_t* pmid, const float* disp, const uint8_t* valid, const float* val,
                   float val_scalar, int64_t n, int32_t channels, const pm_cic_geom* geom, float* mesh,
                   void* stream);
int pm_cic_gather(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                  int32_t channels, const pm_cic_geom* geom, const float* mesh, float* out,
                  void* stream);
int pm_cic_gather_bwd(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                      int32_t channels, const pm_cic_geom* geom, const float* mesh,
                      const float* val_cot, float* disp_cot, float* mesh_cot, void* stream);
int pm_cic_scatter_bwd(const int16_t* pmid, const float* disp, const uint8_t* valid,
                       const float* val, float val_scalar, int64_t n, int32_t channels,
                       const pm_cic_geom* geom, const float* mesh_cot, float* disp_cot,
                       float* val_cot, void* stream);

/* ---------------------------------------------------------- cell sort ----
 * Stable counting sort of particles by the raveled brick key of their containing
 * cell pmid + floor(disp/cell) (wrapped); invalid particles take the sentinel key
 * num_bricks and land last.  Replaces the single-device remnant of the canonical
 * key/order contract: pmid_to_idx (src/pmpp/core/utils.py:292-326), the sentinel
 * (src/pmpp/distributed/routing.py:112-116) and the stable block-scan compaction
 * (cuda/route_kernels.cu:194-225).  perm[j] = input index of the j-th sorted particle;
 * bin_offsets (optional, num_bricks+2 int32) delimits each brick; pmid_sorted (n,3) and
 * disp_sorted (n,3) (each optional) receive the payload in sorted order. */
int pm_sort_cells_workspace(int64_t n, const int32_t mesh_shape[3], const int32_t brick[3],
                            size_t* bytes);
int pm_sort_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                  const int32_t mesh_shape[3], float cell_size, const int32_t brick[3],
                  int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted, float* disp_sorted,
                  void* workspace, size_t workspace_bytes, void* stream);
/* The force step's own binning (pm_gravity runs it internally): the same keys, bins and
 * bin_offsets (num_bricks+2 int32, required) as pm_sort_cells, but the order inside a bin is
 * unspecified (atomic arrival order), which the step's order-independent tile kernels allow.
 * perm, pmid_sorted (n,3), disp_sorted (n,3) and valid_sorted (n) are all written.
 * ptcl_grid (optional) names the Lagrangian grid the particles were generated on in C order;
 * when its product is n and it is a multiple of (4, 8, 64), particles are read in blocks of
 * that grid (speed only; any value gives a valid binning).  No counterpart in the reference,
 * whose single-device path does not order particles (nbody/gravity.py:813). */
int pm_bucket_cells_workspace(int64_t n, const int32_t mesh_shape[3], const int32_t brick[3],
                              size_t* bytes);
int pm_bucket_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                    const int32_t mesh_shape[3], float cell_size, const int32_t brick[3],
                    const int32_t ptcl_grid[3], int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted,
                    float* disp_sorted, uint8_t* valid_sorted, void* workspace, size_t workspace_bytes,
                    void* stream);
/* pm_bucket_cells without a workspace: no counterpart in the reference.  It is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc): the bin counters and scan partials live in a static 8 MiB device
 * array of the library (one per device, held until the process ends), so the call allocates
 * nothing.  Same arguments, outputs and errors as pm_bucket_cells; meshes with more than
 * 2,096,446 bricks return PM_ERROR_UNSUPPORTED, and PM_ERROR_INVALID_ARGUMENT until pm_init
 * has run on the current device.  Every call on a device shares that device's scratch (the
 * counters are cleared by a kernel at the start of each call): calls that overlap on one
 * device, on different streams, race. */
int pm_bucket_cells_noalloc(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                            const int32_t mesh_shape[3], float cell_size, const int32_t brick[3],
                            const int32_t ptcl_grid[3], int32_t* perm, int32_t* bin_offsets, int16_t* pmid_sorted,
                            float* disp_sorted, uint8_t* valid_sorted, void* stream);

/* --------------------------------------------------------------- FFT -----
 * pm_rfft3d  = jnp.fft.rfftn  (unnormalised forward)     used at nbody/gravity.py:436
 * pm_irfft3d = jnp.fft.irfftn (1/N normalised inverse)   used at nbody/gravity.py:528
 * Power-of-two axes from 2 to 2048, last axis even. In-place and out-of-place both work;
 * pm_irfft3d overwrites its input spectrum. */
int pm_fft_workspace(const int32_t shape[3], size_t* bytes);
int pm_rfft3d(const float* in, float* out_complex, const int32_t shape[3], void* workspace,
              size_t workspace_bytes, void* stream);
int pm_irfft3d(float* in_complex, float* out, const int32_t shape[3], void* workspace,
               size_t workspace_bytes, void* stream);

/* Fused spectral force multiplier on a standard-layout spectrum:
 *   out = D_axis(k) * mask(k) * (-1/k^2) * in,  D_axis = -i k_axis (Nyquist derivative 0)
 * i.e. apply_particle_nyquist_filter + laplace_transposed + neg_ik[axis]
 * (nbody/gravity.py:174-196, 263-321; core/configuration.py:365-376).  adjoint != 0
 * multiplies by the conjugate symbol (the spectral adjoint).  axis = -1 gives the
 * potential (no derivative).  Elementwise, so in-place is allowed; the workspace holds the
 * float32 wavenumber/mask tables (pm_spectral_workspace). */
int pm_spectral_workspace(const pm_gravity_desc* desc, size_t* bytes);
int pm_spectral_force_multiply(const float* in_complex, float* out_complex,
                               const pm_gravity_desc* desc, int32_t axis, int32_t adjoint,
                               void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ gravity ----
 * The whole single-device force step (nbody/gravity.py:786-814 on the streamed branch
 * :597-606, 506-543): cell sort, CIC scatter of particle_value, S = 1.5*Omega_m*(rho-1),
 * r2c FFT, particle-Nyquist mask x Poisson x (-i k_i), three c2r FFTs and the CIC gather.
 * acc (n,3) float32 in caller order; invalid particles get 0.
 * pm_gravity_vjp pulls acc_cot back to disp_cot (n,3, overwritten) and accumulates
 * d/dOmega_m into omega_cot[0] (device float, caller zeroes) -- the force VJP walked by
 * force_adj (nbody/integrator.py:620-686).  Each call takes a workspace of at least the size
 * its own query returns (pm_gravity_workspace / pm_gravity_vjp_workspace); the reverse pass
 * needs more (it keeps the forward meshes and three cotangent meshes). */
int pm_gravity_workspace(const pm_gravity_desc* desc, int64_t n, size_t* bytes);
int pm_gravity_vjp_workspace(const pm_gravity_desc* desc, int64_t n, size_t* bytes);
int pm_gravity(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
               const pm_gravity_desc* desc, float* acc, void* workspace, size_t workspace_bytes,
               void* stream);
int pm_gravity_vjp(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                   const pm_gravity_desc* desc, const float* acc_cot, float* disp_cot,
                   float* omega_cot, void* workspace, size_t workspace_bytes, void* stream);

/* The reverse sweep's force pair (nbody/integrator.py:781-837: after drift_adj the force is
 * recomputed at the reconstructed positions, and the next stage's force_adj pulls its
 * cotangent back at those same positions).  pm_gravity_keep is pm_gravity on a workspace of
 * pm_gravity_vjp_workspace bytes that keeps the bucketed particles, the density and the force
 * meshes in it; pm_gravity_vjp_resume is pm_gravity_vjp that starts from them instead of
 * recomputing them.  Its workspace must still hold what the preceding pm_gravity_keep call on
 * the same particles (pmid, disp, valid bit for bit), desc and stream left in it; calls that
 * change the positions or reuse the workspace in between invalidate it.  Results equal
 * pm_gravity / pm_gravity_vjp. */
int pm_gravity_keep(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                    const pm_gravity_desc* desc, float* acc, void* workspace, size_t workspace_bytes, void* stream);
int pm_gravity_vjp_resume(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                          const pm_gravity_desc* desc, const float* acc_cot, float* disp_cot, float* omega_cot,
                          void* workspace, size_t workspace_bytes, void* stream);

/* The force step with a learned correction (nbody/gravity.py:419-430, 635-640 with
 * apply_potential_correction, corrections/core.py:286-334): rho_add (the real mesh, Nx*Ny*Nz
 * float32, C order) is added to the CIC density before S = 1.5*Omega_m*(rho + rho_add - 1),
 * so the correction enters the Poisson solve as a source residual (a potential residual
 * phi_add corresponds to rho_add = -laplacian(phi_add) / (1.5 Omega_m)) instead of costing
 * an extra irfftn + rfftn per force.  rho_add = NULL is pm_gravity.  The VJP also writes
 * rho_cot = dL/d rho_add (the real mesh, overwritten) when rho_cot is not NULL; workspaces
 * are those of pm_gravity_workspace / pm_gravity_vjp_workspace.  pm_gravity_density writes
 * the step's own CIC density (the mesh the correction network reads) and stops there;
 * pm_gravity_corrected_resume then finishes that step with rho_add (the spectral solve and
 * the gather, without bucketing and scattering the particles again): its workspace must
 * still hold what the preceding pm_gravity_density call on the same particles, desc and
 * stream left in it (the network may run in between on other buffers).  Together they are
 * the reference's corrected force (nbody/gravity.py:626-640: density once, then the
 * corrected potential and the force). */
int pm_gravity_density(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                       const pm_gravity_desc* desc, float* rho, void* workspace, size_t workspace_bytes,
                       void* stream);
/* pm_gravity_density_offset  the byte offset inside a pm_gravity_workspace (of the same desc
 *                            and n, 256-byte aligned base) where pm_gravity_density leaves the
 *                            density.  Passing workspace + offset as rho skips the copy out (a
 *                            4 GiB mesh at 1024^3); the mesh there is then overwritten by
 *                            pm_gravity_corrected_resume, so a caller that needs the density
 *                            after the resume passes its own buffer.  No reference counterpart:
 *                            a layout query of this library's workspace. */
int pm_gravity_density_offset(const pm_gravity_desc* desc, int64_t n, size_t* offset);
int pm_gravity_corrected(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                         const pm_gravity_desc* desc, const float* rho_add, float* acc, void* workspace,
                         size_t workspace_bytes, void* stream);
int pm_gravity_corrected_resume(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                                const pm_gravity_desc* desc, const float* rho_add, float* acc, void* workspace,
                                size_t workspace_bytes, void* stream);
int pm_gravity_corrected_vjp(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                             const pm_gravity_desc* desc, const float* rho_add, const float* acc_cot,
                             float* disp_cot, float* omega_cot, float* rho_cot, void* workspace,
                             size_t workspace_bytes, void* stream);

/* pm_gravity_density_cot adds the CIC scatter adjoint's displacement term for a density
 * cotangent dens_cot (the real mesh; cic/pallas.py:358-409 with the particle value N_m/N_p) to
 * disp_cot (n x 3, caller order; invalid particles' rows are left as they are): the particles
 * are binned by brick and the term is taken in the tiled column pass the VJP uses for its own
 * density term.  The workspace is that of pm_gravity_workspace.  The corrected force's reverse
 * uses it for the density cotangent through the network's inputs, known only after the VJP. */
int pm_gravity_density_cot(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
                           const pm_gravity_desc* desc, const float* dens_cot, float* disp_cot, void* workspace,
                           size_t workspace_bytes, void* stream);

/* The particle-Nyquist-filtered potential of a real mesh, on the force step's spectral band:
 *   out = add + irfftn(mask(k) * (-1/k^2) * rfftn(scale * (in + shift)))
 * (nbody/gravity.py:174-196, 263-321, 626-633: the potential the reference's corrected force
 * computes before apply_potential_correction, corrections/core.py:286-334, and feeds to the
 * mesh CNN as its potential channel, corrections/mesh_cnn.py:20-27, 76-84).  shift = -1,
 * scale = 1.5 Omega_m on a density mesh gives the potential of the force's source; the operator
 * is symmetric, so shift = 0 applies its adjoint to a cotangent.  add (NULL: none) may alias
 * out; in may alias out (it is consumed before out is written).  desc gives the mesh, cell size,
 * particle spacing and mask; its omega_m is not used.  Workspace: pm_potential_workspace. */
int pm_potential_workspace(const pm_gravity_desc* desc, size_t* bytes);
int pm_gravity_potential(const float* in, float shift, float scale, const float* add, const pm_gravity_desc* desc,
                         float* out, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ N-body updates ----
 * Elementwise stages of the symplectic integrator on (n,3) float32 arrays (16-byte aligned):
 * pm_kick   vel_out  = vel  + factor * acc   replaces kick      (nbody/integrator.py:512-538)
 * pm_drift  disp_out = disp + factor * vel   replaces drift     (nbody/integrator.py:203-237)
 *           (the output may be the input: an in-place update)
 * Their reverse passes reconstruct the input state from the output and pull cotangents:
 * pm_kick_adj   vel -= factor*acc;  acc_cot += factor*vel_cot;  *dot = sum(vel_cot*acc)
 *               replaces kick_adj                               (nbody/integrator.py:541-582)
 * pm_drift_adj  disp -= factor*vel; vel_cot += factor*disp_cot; *dot = sum(disp_cot*vel)
 *               replaces drift_adj_from_output                  (nbody/integrator.py:429-474)
 * *dot (a device double) is the factor's cotangent, which the caller chains into the
 * cosmology cotangent; it is reduced deterministically in a workspace of
 * pm_nbody_adj_workspace bytes.  acc_cot may be NULL (not accumulated). */
int pm_kick(const float* vel, const float* acc, float* vel_out, int64_t n, float factor, void* stream);
int pm_drift(const float* disp, const float* vel, float* disp_out, int64_t n, float factor, void* stream);
int pm_nbody_adj_workspace(size_t* bytes);
int pm_kick_adj(float* vel, const float* acc, const float* vel_cot, float* acc_cot, int64_t n, float factor,
                double* dot, void* workspace, size_t workspace_bytes, void* stream);
int pm_drift_adj(float* disp, const float* vel, const float* disp_cot, float* vel_cot, int64_t n, float factor,
                 double* dot, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ power spectrum ----
 * Integer-shell sums of an rfft half spectrum (nx, ny, nz/2+1) complex64: the shell
 * reduction of analysis/power_spectrum.py:108-143 (_shell_reduce_local) and :145-176 (its
 * cross variant).  sums[0:num_shells] += sum over modes of the shell of
 * mult * |a|^2 * (sinc sinc sinc)^(-2 mas_power); with spec_b also sums[num_shells:2n] for
 * |b|^2 and sums[2n:3n] for Re(a conj b).  density != 0 treats the inputs as densities
 * (delta_k = rho_k / mean, density_to_pk :348-374).  sums (float64, device) is overwritten;
 * the caller divides by the shell mode counts and multiplies by cell^3 / N_m (:251-256). */
int pm_power_spectrum(const float* spec_a, const float* spec_b, const int32_t shape[3], int32_t mas_power,
                      int32_t num_shells, int32_t density, double* sums, void* stream);
/* Reverse of pm_power_spectrum (auto spectrum): grad_spec = coef[shell] * window * s^2 * spec
 * (s = N / Re spec[0] for a density, else 1), so that with coef[s] = 2 cell^3 dL/dP_s / nmodes_s
 * irfftn(grad_spec) = dL/d(mesh) -- the adjoint the reference takes by autodiff through
 * analysis/power_spectrum.py:196-260.  For a density the constant term of the mean is left
 * out (it carries no particle gradient). */
int pm_power_spectrum_bwd(const float* spec, const int32_t shape[3], int32_t mas_power, int32_t num_shells,
                          int32_t density, const float* shell_coef, float* grad_spec, void* stream);

/* ------------------------------------------------------------ conv3d (learned correction) ----
 * Periodic 3x3x3 convolutions of channels-last bf16 meshes on the matrix cores (BF16 MFMA,
 * float32 accumulation), 16 output channels, for the mesh-CNN potential correction
 * (corrections/mesh_cnn.py:54-105, its Conv3D layers with periodic padding).  Mesh axes are
 * multiples of 8; x / w / res bf16, activations [X][Y][Z][C].
 * pm_conv3d_fwd        pre = conv(x; w[27][16][cin]) + bias (+ res); act = gelu(pre) or pre;
 *                      pre (float32, the first pre_channels = 1 or 16 channels per voxel) and
 *                      act (bf16, 16 channels) are each optional outputs.
 * pm_conv3d_bwd_data   dx = conv^T(dz; w) (+ add), times gelu'(gact) when gact is given
 *                      (16 -> 16 layers; meshes 16-byte aligned; dx may alias add or gact but
 *                      not dz); dx_channels = 16, or 4: only the first four channels of dx are
 *                      written, as an n x 4 mesh (what the input path reads).  bf16_io: 0 =
 *                      dz, add and dx float32; 1 = dz and add bf16, dx float32; 3 = dz, add
 *                      and dx bf16 (the matrix cores read dz as bf16 in every case; gact is
 *                      float32).
 * pm_conv3d_bwd_data_out1  the same for a one-output-channel layer (w[27][16][16], channel 0
 *                      used) from its one-channel gradient dz0 (n floats): dx (n x 16, float32,
 *                      or bf16 when dx_bf16) = conv^T(dz0; w[:, 0, :]) times gelu'(gact) when
 *                      gact is given.
 * pm_conv3d_bwd_weight dw_db = [dW (27*16*cin floats), db (16 floats)] for pre-activation
 *                      gradient dz and layer input x; workspace from the query.
 * pm_conv3d_stem       act = gelu(conv(x0; w[27][16][8]) + bias) for the network input x0 that
 *                      pm_correction_features would write from the density rho (float32
 *                      mesh), the potential phi (NULL: no potential channel) and their stats,
 *                      without writing x0: the constant input channels (a, Omega_m, sigma8) are
 *                      folded into the bias (inference; training keeps x0 for the weight
 *                      gradient).
 * pm_conv3d_out_residual  the output layer fused with pm_correction_residual (inference):
 *                      rho_add = g (rho - 1) + max_residual tanh(conv(x; w[27][16][16]) + bias,
 *                      channel 0) sigma / c, without writing the output pre-activation. */
int pm_conv3d_fwd(const void* x, int32_t cin, const void* w, const float* bias, const void* res, int32_t gelu,
                  float* pre, int32_t pre_channels, void* act, const int32_t shape[3], void* stream);
int pm_conv3d_bwd_data(const void* dz, const void* w, const void* add, const float* gact, void* dx,
                       int32_t dx_channels, int32_t bf16_io, const int32_t shape[3], void* stream);
int pm_conv3d_bwd_data_out1(const float* dz0, const void* w, const float* gact, void* dx, int32_t dx_bf16,
                            const int32_t shape[3], void* stream);
int pm_conv3d_bwd_weight_workspace(const int32_t shape[3], int32_t cin, size_t* bytes);
int pm_conv3d_bwd_weight(const float* dz, const void* x, int32_t cin, const int32_t shape[3], float* dw_db,
                         void* workspace, size_t workspace_bytes, void* stream);
int pm_conv3d_stem(const float* rho, const float* phi, const int32_t shape[3], float omega_m, float a, float sigma8,
                   const float* stats, const void* w, const float* bias, void* act, void* stream);
int pm_conv3d_out_residual(const void* x, const void* w, const float* bias, const float* rho,
                           const int32_t shape[3], float omega_m, float max_residual, const float* stats,
                           float* rho_add, void* stream);
/* pm_conv3d_stem_hidden  the inference stem (as pm_conv3d_stem) fused with a first hidden
 *                        layer that is a plain GELU conv (16 -> 16, weights w_hidden [27][16][16],
 *                        bias b_hidden): act = gelu(conv(gelu(stem(rho))) + b_hidden), bf16
 *                        (n x 16).  The stem's activation stays in LDS.  Meshes whose x extent is
 *                        a multiple of 4 and z extent a multiple of 16 (PM_ERROR_UNSUPPORTED
 *                        otherwise: use the two kernels). */
int pm_conv3d_stem_hidden(const float* rho, const float* phi, const int32_t shape[3], float omega_m, float a,
                          float sigma8, const float* stats, const void* w_stem, const float* b_stem,
                          const void* w_hidden, const float* b_hidden, void* act, void* stream);

/* ------------------------------------------------------------ learned correction ----
 * The elementwise stages around the conv stack of the mesh CNN correction
 * (corrections/mesh_cnn.py:54-105), its training loss and its optimizer
 * (corrections/common.py:136-174).  With c = 1.5 Omega_m and S = c (rho - 1):
 * pm_correction_features  sigma = sqrt(mean S^2 + 1e-6) and the gain
 *                         g = max_residual tanh(gain_params . [a, Omega_m, sigma8, 1]) into
 *                         stats[0..2] (device floats); x0 = bf16 [S/sigma, a, Omega_m, sigma8,
 *                         0, 0, 0, 0] per cell (n x 8 bf16, 16-byte aligned; NULL: stats only).
 *                         With the potential phi (float32 mesh, pm_gravity_potential of rho;
 *                         NULL: no potential channel, mesh_cnn.py:20-27, 81-84) also
 *                         stats[3] = sigma_p = sqrt(mean phi^2 + 1e-6) and x0 = bf16 [S/sigma,
 *                         phi/sigma_p, a, Omega_m, sigma8, 0, 0, 0].  With gain_params NULL,
 *                         stats are read instead (those of the whole mesh when rho and phi are
 *                         slabs of it) and only x0 is written.
 * pm_correction_residual  rho_add = g (rho - 1) + max_residual tanh(out) sigma / c, out the
 *                         output layer's float32 pre-activation of channel 0 (n floats).
 * pm_correction_residual_bwd  the output layer's pre-activation gradient dz_out (n x 16 with
 *                         channel 0 live, or that channel alone, n floats, when dz_channels = 1)
 *                         and the gain parameters' gradient (4 floats) from rho_cot.
 * pm_correction_input_bwd the cotangents through the network's inputs: rho_bar (n floats,
 *                         overwritten) = d<rho_cot, rho_add>/d rho and omega_cot[0] += its
 *                         Omega_m derivative, from dx0 (n x 4 float32, channels 0..3 of the
 *                         gradient of the network input), the forward's out and stats.  With the
 *                         potential channel (phi not NULL) it also writes phi_bar (n floats), the
 *                         cotangent of phi, whose path to the density the caller adds with
 *                         pm_gravity_potential(phi_bar, 0, 1.5 Omega_m, rho_bar, ..., rho_bar);
 *                         rho_bar holds the other paths.  With the scatter adjoint of rho_bar this
 *                         is reverse mode of the corrected force through the density.
 * pm_force_loss           loss[0] = sum |acc - target|^2 / sum |target|^2 (device doubles,
 *                         loss[1] = 1 / sum |target|^2), acc_cot = d loss / d acc if not NULL.
 * pm_adamax_step          clip_by_global_norm(clip_norm) then adamax(learning_rate, b1, b2, eps)
 *                         on n float32 parameters, skipped when the gradient norm is not
 *                         finite (the apply_if_finite wrapper); count (device int32) is the
 *                         step counter; params_bf16 (optional) receives the bf16 mirror.
 * pm_to_bf16              dst = bf16(src).
 * Reductions are deterministic (fixed-order float64 partials); the workspace holds them. */
int pm_correction_workspace(size_t* bytes);
int pm_correction_features(const float* rho, const float* phi, int64_t n, float omega_m, float a, float sigma8,
                           float max_residual, const float* gain_params, void* x0, float* stats, void* workspace,
                           size_t workspace_bytes, void* stream);
int pm_correction_residual(const float* out, const float* rho, int64_t n, float omega_m, float max_residual,
                           const float* stats, float* rho_add, void* stream);
int pm_correction_residual_bwd(const float* rho_cot, const float* out, const float* rho, int64_t n, float omega_m,
                               float a, float sigma8, float max_residual, const float* stats, float* dz_out,
                               int32_t dz_channels, float* gain_grad, void* workspace, size_t workspace_bytes,
                               void* stream);
int pm_correction_input_bwd(const float* rho_cot, const float* out, const float* rho, const float* phi,
                            const float* dx0, int64_t n, float omega_m, float max_residual, const float* stats,
                            const float* gain_params, float* rho_bar, float* phi_bar, float* omega_cot,
                            void* workspace, size_t workspace_bytes, void* stream);
int pm_force_loss(const float* acc, const float* target, int64_t n, double* loss, float* acc_cot, void* workspace,
                  size_t workspace_bytes, void* stream);
int pm_adamax_step(float* params, const float* grads, float* mu, float* nu, int32_t* count, int64_t n,
                   float learning_rate, float b1, float b2, float eps, float clip_norm, void* params_bf16,
                   void* workspace, size_t workspace_bytes, void* stream);
int pm_to_bf16(const float* src, int64_t n, void* dst, void* stream);

/* ------------------------------------------------------------ initial conditions ----
 * Spectral and real-space steps of the IC pipeline on the particle grid (n0, n1, n2), spectra
 * in the rfftn layout (n0, n1, n2/2+1) complex; ktab[a] are the float32 angular wavenumbers
 * of axis a (fftfreq * 2 pi / spacing, rfftfreq for the last axis: n0, n1, n2/2+1 entries).
 * pm_ic_linear_modes   spec *= scale sqrt(coef (k/k_pivot)^n_s T(k)^2), T linearly
 *                      interpolated from (table_k, table_T) and clamped, k = 0 -> 0:
 *                      linear_modes (initial_conditions/modes.py:399-454, boltzmann.py:428-482)
 * pm_ic_spectral_op    out = symbol(k) in: op 0 scale; 1 -1/k^2 (k = 0 -> 0,
 *                      laplace_transposed, gravity.py:263-287); 2 -k_i k_j with the Nyquist
 *                      wavenumber zeroed for i != j (_strain, lpt.py:17-43); 3 -i k_i with the
 *                      Nyquist zeroed (neg_grad, gravity.py:355-391).  A wavenumber is the
 *                      Nyquist when | |k| - nyquist[a] | <= tol[a].
 * pm_ic_product_acc    acc = (init ? 0 : acc) + coef a b elementwise (b = NULL: + coef a); the
 *                      2LPT source (_L, lpt.py:239-302) and the sums of its adjoint
 * pm_ic_lpt_accumulate disp[:, axis] += D grad, vel[:, axis] += V grad (vel may be NULL) for
 *                      particles in grid order (lpt.py:419-436) */
int pm_ic_linear_modes(float* spec, const int32_t shape[3], const float* const ktab[3], const float* table_k,
                       const float* table_T, int32_t table_n, float coef, float k_pivot, float n_s, float scale,
                       void* stream);
int pm_ic_spectral_op(const float* in, float* out, const int32_t shape[3], const float* const ktab[3], int32_t op,
                      int32_t axis_i, int32_t axis_j, float scale, const float nyquist[3], const float tol[3],
                      void* stream);
int pm_ic_product_acc(float* acc, const float* a, const float* b, int64_t n, float coef, int32_t init,
                      void* stream);
int pm_ic_lpt_accumulate(float* disp, float* vel, const float* grad, int64_t n, int32_t axis, float D, float V,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PMPP_AMD_H */
