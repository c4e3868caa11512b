"""Kernel seam: HIP CIC scatter/gather and their adjoints.

Same contract as the reference's Pallas kernels ``pallas_scatter``, ``pallas_gather``,
``pallas_gather_bwd`` and ``pallas_scatter_bwd`` (``src/pmpp/cic/pallas.py:412-571``):
``offset`` is a scalar or (3,) physical offset (:107-110), ``particle_cell_size`` is
``conf.cell_size``, ``cell_size`` an explicit mesh cell size (None: same as particles),
``global_shape`` the periodic wrap shape, corners outside the local mesh are dropped
(:261-265), ``valid_mask`` is (N,) bool and invalid particles neither read nor write
(:133-143).  float32 and three spatial dimensions only, as in the reference (:87-100, :163).
"""

from __future__ import annotations

import math

import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, require_device, stream_handle


def _check(pmid, disp, mesh):
    require_device(pmid, "pmid")
    require_device(disp, "disp")
    if pmid.ndim != 2 or pmid.shape[1] != 3 or disp.shape != pmid.shape:
        raise NotImplementedError("HIP CIC currently supports (N, 3) particles")
    if mesh.dtype != torch.float32:
        raise RuntimeError(
            f"HIP CIC supports float32 meshes only (got {mesh.dtype}); PM++'s GPU kernels are float32 too"
        )
    if mesh.ndim < 3:
        raise NotImplementedError("HIP CIC currently supports three spatial dimensions")


def _geom(mesh_spatial, global_shape, offset, particle_cell_size, cell_size):
    off = torch.as_tensor(offset, dtype=torch.float64).cpu().reshape(-1)
    if off.numel() == 1:
        off = off.repeat(3)
    if off.numel() != 3:
        raise ValueError("offset must be a scalar or have 3 entries")
    g = _lib.CicGeom()
    for a in range(3):
        g.local_shape[a] = int(mesh_spatial[a])
        g.global_shape[a] = int(global_shape[a])
        g.offset[a] = float(off[a])
    g.particle_cell = float(particle_cell_size)
    g.mesh_cell = 0.0 if cell_size is None else float(cell_size)
    if cell_size is not None and not g.mesh_cell > 0:
        raise ValueError("cell_size must be positive")
    return g


def _valid(valid_mask, n, device):
    if valid_mask is None:
        return None
    v = torch.as_tensor(valid_mask, device=device)
    if v.ndim != 1 or v.shape[0] != n:
        raise ValueError(
            f"HIP CIC valid_mask must have shape (particle_count,), got {tuple(v.shape)} for particle_count={n}."
        )
    return v.to(torch.uint8).contiguous()


def _is_scalar(val):
    return not isinstance(val, torch.Tensor) or val.ndim == 0


def hip_scatter(pmid, disp, val, mesh, *, offset, particle_cell_size, cell_size=None, global_shape,
                valid_mask=None):
    """mesh + sum_p val_p W_p (returns a new mesh)."""
    _check(pmid, disp, mesh)
    ensure_init(mesh.device)
    n = pmid.shape[0]
    chan_shape = tuple(mesh.shape[3:])
    C = math.prod(chan_shape) if chan_shape else 1
    out = as_contig(mesh, torch.float32, "mesh").clone()
    pm = as_contig(pmid, torch.int16, "pmid")
    dp = as_contig(disp, torch.float32, "disp")
    if _is_scalar(val):
        vt, vs = None, float(val)
    else:
        vt = as_contig(val, torch.float32, "val")
        if tuple(vt.shape[1:]) != chan_shape:
            raise ValueError(f"channel shape mismatch: {chan_shape} != {tuple(vt.shape[1:])}")
        vs = 0.0
    valid = _valid(valid_mask, n, mesh.device)
    g = _geom(mesh.shape[:3], global_shape, offset, particle_cell_size, cell_size)
    _lib.call("pm_cic_scatter", ptr(pm), ptr(dp), ptr(valid), ptr(vt), vs, n, C, g, ptr(out),
              stream_handle(mesh.device))
    return out


def hip_gather(pmid, disp, mesh, *, offset, particle_cell_size, cell_size=None, global_shape,
               valid_mask=None):
    """(N,)+channels values sum_corners W mesh; invalid particles read 0."""
    _check(pmid, disp, mesh)
    ensure_init(mesh.device)
    n = pmid.shape[0]
    chan_shape = tuple(mesh.shape[3:])
    C = math.prod(chan_shape) if chan_shape else 1
    m = as_contig(mesh, torch.float32, "mesh")
    pm = as_contig(pmid, torch.int16, "pmid")
    dp = as_contig(disp, torch.float32, "disp")
    out = torch.empty((n,) + chan_shape, dtype=torch.float32, device=mesh.device)
    valid = _valid(valid_mask, n, mesh.device)
    g = _geom(mesh.shape[:3], global_shape, offset, particle_cell_size, cell_size)
    _lib.call("pm_cic_gather", ptr(pm), ptr(dp), ptr(valid), n, C, g, ptr(m), ptr(out),
              stream_handle(mesh.device))
    return out


def hip_gather_bwd(pmid, disp, mesh, val_cot, *, offset, particle_cell_size, cell_size=None, global_shape,
                   valid_mask=None):
    """Gather adjoint: (disp_cot (N,3), mesh_cot like mesh)."""
    _check(pmid, disp, mesh)
    ensure_init(mesh.device)
    n = pmid.shape[0]
    chan_shape = tuple(mesh.shape[3:])
    C = math.prod(chan_shape) if chan_shape else 1
    m = as_contig(mesh, torch.float32, "mesh")
    vc = as_contig(val_cot, torch.float32, "val_cot")
    pm = as_contig(pmid, torch.int16, "pmid")
    dp = as_contig(disp, torch.float32, "disp")
    disp_cot = torch.empty((n, 3), dtype=torch.float32, device=mesh.device)
    mesh_cot = torch.zeros_like(m)
    valid = _valid(valid_mask, n, mesh.device)
    g = _geom(mesh.shape[:3], global_shape, offset, particle_cell_size, cell_size)
    _lib.call("pm_cic_gather_bwd", ptr(pm), ptr(dp), ptr(valid), n, C, g, ptr(m), ptr(vc), ptr(disp_cot),
              ptr(mesh_cot), stream_handle(mesh.device))
    return disp_cot, mesh_cot


def hip_scatter_bwd(pmid, disp, val, mesh_cot, *, offset, particle_cell_size, cell_size=None, global_shape,
                    valid_mask=None):
    """Scatter adjoint: (disp_cot (N,3), val_cot) with a 0-d val_cot for a scalar val."""
    _check(pmid, disp, mesh_cot)
    ensure_init(mesh_cot.device)
    n = pmid.shape[0]
    chan_shape = tuple(mesh_cot.shape[3:])
    C = math.prod(chan_shape) if chan_shape else 1
    mc = as_contig(mesh_cot, torch.float32, "mesh_cot")
    pm = as_contig(pmid, torch.int16, "pmid")
    dp = as_contig(disp, torch.float32, "disp")
    scalar = _is_scalar(val)
    if scalar:
        vt, vs = None, float(val)
        val_cot = torch.zeros((), dtype=torch.float32, device=mc.device)
    else:
        vt = as_contig(val, torch.float32, "val")
        vs = 0.0
        val_cot = torch.empty((n,) + chan_shape, dtype=torch.float32, device=mc.device)
    disp_cot = torch.empty((n, 3), dtype=torch.float32, device=mc.device)
    valid = _valid(valid_mask, n, mc.device)
    g = _geom(mesh_cot.shape[:3], global_shape, offset, particle_cell_size, cell_size)
    _lib.call("pm_cic_scatter_bwd", ptr(pm), ptr(dp), ptr(valid), ptr(vt), vs, n, C, g, ptr(mc),
              ptr(disp_cot), ptr(val_cot), stream_handle(mc.device))
    return disp_cot, val_cot
