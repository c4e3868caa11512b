"""Timed CPU baseline: the reference's single-device force step restated in float32 on host cores.

Test/measurement infrastructure only (``bench.py``'s ``cpu_baseline`` leg).  It is the same
operator sequence as ``src/pmpp/nbody/gravity.py:786-814`` (CIC scatter, 1.5 Omega_m (rho-1),
rfftn, particle-Nyquist mask, -1/k^2, -i k_i, irfftn, CIC gather): the CIC stages in C with
OpenMP (``oracle/cpu_cic.c``, built by ``oracle/build_cpu.py``) and the FFTs as ``scipy.fft``
with ``workers`` threads.  ``nbody_run`` is the whole plain run of the reference's first
workload (white noise -> linear modes -> 2LPT, ``src/pmpp/initial_conditions/lpt.py:339-449``,
then the KDK steps of ``src/pmpp/nbody/integrator.py:689-731``) with this force.  The
reference's own CPU path (JAX/XLA) cannot run here: jax is not installed.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import scipy.fft

from .spectral import green_multiplier, neg_k, particle_nyquist_masks

_LIB = None
_GREEN = {}


def _lib():
    global _LIB
    if _LIB is None:
        path = Path(__file__).resolve().parent / "libpm_cpu_cic.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m oracle.build_cpu`")
        lib = ctypes.CDLL(str(path))
        vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
        lib.cpu_cic_scatter.argtypes = [vp, vp, i64, vp, f32, f32, vp]
        lib.cpu_cic_gather.argtypes = [vp, vp, i64, vp, f32, vp, vp, ctypes.c_int32, ctypes.c_int32]
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _symbols(mesh_shape, cell, ptcl_grid_shape, ptcl_spacing):
    """float32 Green x mask and the three -k_i vectors (constant per geometry, cached)."""
    key = (tuple(mesh_shape), float(cell), tuple(ptcl_grid_shape), float(ptcl_spacing))
    if key not in _GREEN:
        masks = particle_nyquist_masks(mesh_shape, ptcl_grid_shape, ptcl_spacing, cell)
        g = green_multiplier(mesh_shape, cell, masks).astype(np.float32)
        _GREEN[key] = (g, [k.astype(np.float32) for k in neg_k(mesh_shape, cell)])
    return _GREEN[key]


def force_step(pmid, disp, omega_m, mesh_shape, cell, ptcl_grid_shape, ptcl_spacing, workers=1):
    """One float32 PM force evaluation on the CPU; returns acc (N,3) float32."""
    lib = _lib()
    mesh_shape = tuple(int(s) for s in mesh_shape)
    pmid = np.ascontiguousarray(pmid, dtype=np.int16)
    disp = np.ascontiguousarray(disp, dtype=np.float32)
    n = pmid.shape[0]
    shape = np.asarray(mesh_shape, dtype=np.int32)
    val = np.float32(np.prod(mesh_shape) / np.prod(ptcl_grid_shape))
    rho = np.zeros(mesh_shape, dtype=np.float32)
    lib.cpu_cic_scatter(_p(pmid), _p(disp), n, _p(shape), float(cell), float(val), _p(rho))
    rho -= np.float32(1)
    rho *= np.float32(1.5 * omega_m)
    pot = scipy.fft.rfftn(rho, workers=workers)
    del rho
    green, nks = _symbols(mesh_shape, cell, ptcl_grid_shape, ptcl_spacing)
    pot *= green
    acc = np.empty((n, 3), dtype=np.float32)
    shapes = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    for i in range(3):
        spec = pot * (1j * nks[i].reshape(shapes[i])).astype(np.complex64)
        F = np.ascontiguousarray(scipy.fft.irfftn(spec, s=mesh_shape, workers=workers), dtype=np.float32)
        del spec
        lib.cpu_cic_gather(_p(pmid), _p(disp), n, _p(shape), float(cell), _p(F), _p(acc), 3, i)
    return acc


def nbody_run(ng, ratio, box, n_steps, seed, cosmo, workers=1):
    """The plain run of the reference's first workload on the CPU: white noise (seed) -> linear
    modes -> 2LPT at a = 1/64 (float64, oracle.ic) -> ``n_steps`` KDK steps linear in a to a = 1
    with the float32 force above.  ``cosmo`` supplies Omega_m and the host transfer/growth
    tables.  Returns (disp, vel, acc) float32 of the final state."""
    import math

    from pmpp_amd.core.configuration import Configuration
    from pmpp_amd.cosmology import boltzmann
    from pmpp_amd.cosmology.background import E2
    from pmpp_amd.nbody.integrator import growth_table

    from . import cosmology as oc
    from . import ic as oic
    from . import nbody as onb

    a0 = 1 / 64
    conf = Configuration(box / ng, (ng,) * 3, mesh_shape=ratio, a_start=a0, a_nbody_maxstep=(1 - a0) / n_steps)
    grid, spacing = conf.ptcl_grid_shape, conf.ptcl_spacing
    tk, tT = boltzmann.transfer_table(cosmo)
    coef = boltzmann.power_coef(cosmo, cosmo.A_s_1e9)
    modes = oic.linear_modes(oic.ortho_rfftn(oic.white_noise(seed, grid)), grid, spacing, tk, tT, coef,
                             boltzmann.k_pivot(cosmo), cosmo.n_s)
    tab = growth_table(cosmo, conf)
    growth = [(float(tab.growth(a0, m)), a0 * a0 * math.sqrt(float(E2(a0, cosmo))) * float(tab.growth(a0, m, 1)))
              for m in (1, 2)]
    d, v = oic.lpt(modes, grid, spacing, growth, order=2)
    # particles on the Lagrangian grid: anchors on the mesh, lattice offsets in the displacement
    pm1, off1 = [], []
    for sp, sm in zip(grid, conf.mesh_shape):
        p = np.rint(np.linspace(0, sm, num=sp, endpoint=False)).astype(np.int64)
        pm1.append(p)
        off1.append((np.arange(sp) * sm - p * sp).astype(np.float64) * (conf.cell_size / sp))
    pmid = np.stack(np.meshgrid(*pm1, indexing="ij"), -1).reshape(-1, 3).astype(np.int16)
    lattice = np.stack(np.meshgrid(*off1, indexing="ij"), -1).reshape(-1, 3)
    disp = (lattice + d.reshape(-1, 3)).astype(np.float32)
    vel = v.reshape(-1, 3).astype(np.float32)
    om = float(cosmo.Omega_m)
    table = oc.growth_table(conf.growth_a, om, conf.a_lpt_step)

    def force(x):
        return force_step(pmid, x, om, conf.mesh_shape, conf.cell_size, grid, spacing, workers)

    acc = force(disp)
    a = conf.a_nbody
    for a_prev, a_next in zip(a[:-1], a[1:]):
        for st in onb.stages(a_prev, a_next, conf.symp_splits):
            if st[0] == "drift":
                _, av, ap, an, _ = st
                disp = disp + vel * np.float32(oc.drift_factor(av, ap, an, table, conf.growth_a, om))
                acc = force(disp)
            else:
                _, af, ap, an = st
                vel = vel + acc * np.float32(oc.kick_factor(af, ap, an, table, conf.growth_a, om))
    return disp, vel, acc
