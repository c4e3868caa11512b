"""Float64 NumPy restatement of PM++'s single-device spectral force (gravity).

Follows (read as text, never imported):

* ``src/pmpp/numerics/fft.py:4-44``           -- ``fftfreq``: angular wavenumbers, rfft last axis.
* ``src/pmpp/core/utils.py:15-37``            -- ``build_particle_nyquist_filter``.
* ``src/pmpp/core/configuration.py:365-376``  -- cached ``neg_ik`` with the Nyquist derivative zeroed.
* ``src/pmpp/nbody/gravity.py:74-119, 263-321`` -- ``k^2`` and ``laplace_transposed`` (-S/k^2, k=0 -> 0).
* ``src/pmpp/nbody/gravity.py:174-196``       -- ``apply_particle_nyquist_filter``.
* ``src/pmpp/nbody/gravity.py:626-633, 597-606, 506-543, 786-814`` -- the single-device
  streamed force: scatter, delta*1.5*Omega_m, rfftn, mask, Poisson, -ik_i, irfftn, gather.
* ``src/pmpp/cic/scatter.py:258-265``        -- default particle value N_m/N_p on valid particles.

The FFTs are pocketfft's ``rfftn``/``irfftn`` (a library routine; ``scipy.fft`` with one worker
thread per allowed core, the same transform ``numpy.fft`` runs on one thread), pinned in
``tests/test_oracle_spectral.py`` against a brute-force O(N^2) DFT on small meshes.  The CIC
stages run over particle chunks of ``CHUNK`` so meshes of 512^3 cells and 16.8M particles fit
in host memory; chunking changes only the float64 summation order.  The
force is pinned by the closed form of a single plane wave (a = -A k0 sin(k0.x)/k0^2 of the
source A cos(k0.x)), by zero force on a uniform lattice, by momentum conservation, and by the
linearity in Omega_m (gravity.py:630).  ``gravity_vjp`` is pinned by central finite differences.
``potential`` is pinned by the closed form of a plane wave (phi = -A cos(k0.x)/k0^2) and by
the force meshes being its spectral gradient, F_i = irfftn(-i k_i phi_hat) term by term.
"""

from __future__ import annotations

import os

import numpy as np
import scipy.fft

from .cic import scatter, gather, gather_many, scatter_adj, gather_adj

F32_EPS = float(np.finfo(np.float32).eps)
CHUNK = 1 << 21  # particles per CIC chunk
_WORKERS = max(1, len(os.sched_getaffinity(0)))


def rfftn(x):
    """Unnormalised forward real 3-D DFT (numpy.fft.rfftn convention)."""
    return scipy.fft.rfftn(np.asarray(x, dtype=np.float64), workers=_WORKERS)


def irfftn(x, s):
    """1/N normalised inverse of ``rfftn`` onto a real mesh of shape ``s``."""
    return scipy.fft.irfftn(x, s=tuple(s), axes=(0, 1, 2), workers=_WORKERS)


def _chunks(n):
    return [(i, min(i + CHUNK, n)) for i in range(0, max(n, 1), CHUNK)] if n else []


def _sel(a, lo, hi):
    return None if a is None else np.asarray(a)[lo:hi]


def scatter_chunked(pmid, disp, val, mesh_shape, cell_size, valid=None):
    """CIC density of per-particle values ``val`` (N,), particle chunk by chunk."""
    dens = np.zeros(tuple(mesh_shape))
    for lo, hi in _chunks(np.asarray(pmid).shape[0]):
        dens = scatter(pmid[lo:hi], disp[lo:hi], val[lo:hi], dens, cell_size, mesh_shape, valid=_sel(valid, lo, hi))
    return dens


def gather_chunked(pmid, disp, meshes, cell_size, mesh_shape, valid=None):
    """CIC gather of each mesh of ``meshes`` at the particles, chunk by chunk: (N, len(meshes))."""
    n = np.asarray(pmid).shape[0]
    out = np.zeros((n, len(meshes)))
    for lo, hi in _chunks(n):
        out[lo:hi] = gather_many(pmid[lo:hi], disp[lo:hi], meshes, cell_size, mesh_shape, valid=_sel(valid, lo, hi))
    return out


def fftfreq_angular(n, spacing, last=False):
    """Angular wavenumbers of one axis (numerics/fft.py:30-41), float64."""
    period = 2 * np.pi / spacing
    f = np.fft.rfftfreq(n) if last else np.fft.fftfreq(n)
    return f * period


def kvec(mesh_shape, cell_size):
    """Per-axis angular wavenumbers (kx, ky, kz_half) in float64."""
    nx, ny, nz = mesh_shape
    return (fftfreq_angular(nx, cell_size), fftfreq_angular(ny, cell_size),
            fftfreq_angular(nz, cell_size, last=True))


def particle_nyquist_masks(mesh_shape, ptcl_grid_shape, ptcl_spacing, cell_size):
    """Per-axis 0/1 masks keeping |k| <= pi/l_p (1 + 8 eps_f32) (core/utils.py:15-37).

    The support is decided in float32 exactly as the reference computes it, so the
    set of kept modes is identical; an empty tuple means "no filter" (mesh == particle grid).
    """
    if tuple(mesh_shape) == tuple(ptcl_grid_shape):
        return ()
    k_nyq = np.float32(np.pi / ptcl_spacing)
    limit = k_nyq + k_nyq * np.float32(8 * F32_EPS)
    out = []
    for k in kvec(mesh_shape, cell_size):
        k32 = k.astype(np.float32)
        out.append((np.abs(k32) <= limit).astype(np.float64))
    return tuple(out)


def neg_k(mesh_shape, cell_size):
    """-k_i with the Nyquist derivative zeroed (configuration.py:369-376), float64.

    The multiplier of the spectral gradient is i * neg_k; the zero test is done in
    float32 as in the reference: | |k| - pi/cell | <= (pi/cell) eps_f32.
    """
    nyq = np.float32(np.pi / cell_size)
    eps = nyq * np.float32(F32_EPS)
    out = []
    for k in kvec(mesh_shape, cell_size):
        k32 = k.astype(np.float32)
        zero = np.abs(np.abs(k32) - nyq) <= eps
        out.append(np.where(zero, 0.0, -k))
    return tuple(out)


def green_multiplier(mesh_shape, cell_size, masks=()):
    """mask(k) * (-1/k^2) with the k=0 mode set to 0 (gravity.py:263-287 + :174-196)."""
    kx, ky, kz = kvec(mesh_shape, cell_size)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    g = np.where(k2 != 0, -1.0 / np.where(k2 == 0, 1.0, k2), 0.0)
    if masks:
        g = g * masks[0][:, None, None] * masks[1][None, :, None] * masks[2][None, None, :]
    return g


def force_meshes(density, omega_m, mesh_shape, cell_size, masks=()):
    """Real-space force components (F_x, F_y, F_z) from a density mesh (gravity.py:626-633, 506-543)."""
    src = (np.asarray(density, dtype=np.float64) - 1.0) * (1.5 * float(omega_m))
    src_hat = rfftn(src)
    del src
    pot = src_hat * green_multiplier(mesh_shape, cell_size, masks)
    del src_hat
    nk = neg_k(mesh_shape, cell_size)
    shapes = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    return tuple(irfftn(1j * nk[i].reshape(shapes[i]) * pot, tuple(mesh_shape)) for i in range(3))


def potential(density, omega_m, mesh_shape, cell_size, masks=()):
    """Real-space potential irfftn(mask * (-1/k^2) * rfftn(1.5 Omega_m (density - 1)))
    (gravity.py:419-430, read back in real space as corrections/core.py:286-334 does for the
    mesh CNN's potential channel, mesh_cnn.py:76-84)."""
    src = (np.asarray(density, dtype=np.float64) - 1.0) * (1.5 * float(omega_m))
    return irfftn(rfftn(src) * green_multiplier(mesh_shape, cell_size, masks), tuple(mesh_shape))


def default_value(n, mesh_shape, ptcl_num, valid=None):
    """Per-particle CIC value valid * N_m / N_p (scatter.py:258-265)."""
    v = np.full((n,), float(np.prod(mesh_shape)) / float(ptcl_num))
    if valid is not None:
        v = np.where(np.asarray(valid, dtype=bool), v, 0.0)
    return v


def gravity(pmid, disp, omega_m, mesh_shape, cell_size, ptcl_num, masks=(), valid=None, rho_add=None):
    """Accelerations (N,3) of the single-device PM force (gravity.py:786-814), float64.

    ``rho_add`` (a real mesh) is added to the CIC density before the source is formed: the
    learned correction applied as a source residual (the potential-residual hook of
    corrections/core.py:286-334 moved in front of the Poisson solve)."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=np.float64)
    n = pmid.shape[0]
    val = default_value(n, mesh_shape, ptcl_num, valid)
    dens = scatter_chunked(pmid, disp, val, mesh_shape, cell_size, valid=valid)
    if rho_add is not None:
        dens = dens + np.asarray(rho_add, dtype=np.float64)
    comps = force_meshes(dens, omega_m, mesh_shape, cell_size, masks)
    del dens
    return gather_chunked(pmid, disp, comps, cell_size, mesh_shape, valid=valid)


def gravity_vjp(pmid, disp, omega_m, acc_cot, mesh_shape, cell_size, ptcl_num, masks=(), valid=None,
                rho_add=None):
    """Reverse of ``gravity`` w.r.t. ``disp`` and ``omega_m`` (call stack (2) of the force VJP).

    Order of the pullbacks: gather^T per component (disp cotangent + mesh cotangent), the
    irfftn transpose, the -ik_i multiply transpose, the self-adjoint Poisson symbol
    (gravity.py:307-318), the mask, the rfftn transpose, the 1.5 Omega_m scale, and the
    scatter^T.  Implemented through the exact linear adjoints of numpy's real FFTs.
    Returns ``(disp_cot (N,3), omega_m_cot)``, and ``rho_cot`` (the density cotangent, which
    is also that of ``rho_add``) as a third item when ``rho_add`` is given.
    """
    mesh_shape = tuple(mesh_shape)
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=np.float64)
    n = pmid.shape[0]
    val = default_value(n, mesh_shape, ptcl_num, valid)
    dens = scatter_chunked(pmid, disp, val, mesh_shape, cell_size, valid=valid)
    if rho_add is not None:
        dens = dens + np.asarray(rho_add, dtype=np.float64)
    comps = force_meshes(dens, omega_m, mesh_shape, cell_size, masks)
    acc_cot = np.asarray(acc_cot, dtype=np.float64)
    disp_cot = np.zeros((n, 3))
    mesh_cots = []
    for i in range(3):
        mc = np.zeros(mesh_shape)
        for lo, hi in _chunks(n):
            dc, m = gather_adj(pmid[lo:hi], disp[lo:hi], comps[i], acc_cot[lo:hi, i], cell_size, mesh_shape,
                               valid=_sel(valid, lo, hi))
            disp_cot[lo:hi] += dc
            mc += m
        mesh_cots.append(mc)
    del comps
    # F_i = irfftn(D_i * G * rfftn(S)), D_i = i neg_k_i.  For a real linear map
    # F = A S its transpose A^T is obtained from the adjoint-of-irfftn / rfftn pair:
    # A^T g = irfftn(conj(D_i) G rfftn(g)) exactly, because the composite is a real
    # convolution whose kernel transposes to its mirror image (conj of a Hermitian symbol).
    g = green_multiplier(mesh_shape, cell_size, masks)
    nk = neg_k(mesh_shape, cell_size)
    shapes = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    src_cot = np.zeros(mesh_shape)
    for i in range(3):
        d = 1j * nk[i].reshape(shapes[i])
        src_cot += irfftn(np.conj(d) * g * rfftn(mesh_cots[i]), mesh_shape)
    src = dens - 1.0
    omega_cot = 1.5 * float(np.sum(src_cot * src))
    dens_cot = src_cot * (1.5 * float(omega_m))
    for lo, hi in _chunks(n):
        dc, _ = scatter_adj(pmid[lo:hi], disp[lo:hi], val[lo:hi], dens_cot, cell_size, mesh_shape,
                            valid=_sel(valid, lo, hi))
        disp_cot[lo:hi] += dc
    if rho_add is not None:
        return disp_cot, omega_cot, dens_cot
    return disp_cot, omega_cot
