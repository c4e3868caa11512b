"""CIC scatter with a hand-written VJP (drop-in for ``pmpp.cic.scatter``).

Follows ``src/pmpp/cic/scatter.py``: ``scatter`` (:235-275), the custom-VJP primitive
``_scatter`` (:278-312) and its rules ``_scatter_fwd``/``_scatter_bwd`` (:375-437).  The
primitive's forward and backward run the HIP kernels of ``pmpp_amd.cic.kernels``.
"""

from __future__ import annotations

import torch

from .kernels import hip_scatter, hip_scatter_bwd


class _ScatterFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, mesh, val_t, pmid, val_scalar, offset, particle_cell, cell_size, global_shape, valid):
        val = val_scalar if val_t is None else val_t
        out = hip_scatter(pmid, disp, val, mesh, offset=offset, particle_cell_size=particle_cell,
                          cell_size=cell_size, global_shape=global_shape, valid_mask=valid)
        ctx.save_for_backward(disp, pmid, val_t, valid)
        ctx.meta = (val_scalar, offset, particle_cell, cell_size, global_shape)
        return out

    @staticmethod
    def backward(ctx, mesh_cot):
        disp, pmid, val_t, valid = ctx.saved_tensors
        val_scalar, offset, particle_cell, cell_size, global_shape = ctx.meta
        val = val_scalar if val_t is None else val_t
        disp_cot, val_cot = hip_scatter_bwd(pmid, disp, val, mesh_cot.contiguous(), offset=offset,
                                            particle_cell_size=particle_cell, cell_size=cell_size,
                                            global_shape=global_shape, valid_mask=valid)
        val_grad = None
        if val_t is not None and ctx.needs_input_grad[2]:
            val_grad = val_cot.reshape(val_t.shape)
        return disp_cot, mesh_cot, val_grad, None, None, None, None, None, None, None


def _scatter(pmid, disp, conf, mesh, val, offset, cell_size, valid_mask=None):
    """Local multilinear scatter primitive with a custom VJP (scatter.py:278-312)."""
    if val is None:
        val = conf.mesh_size / conf.ptcl_num
    val_t = val if isinstance(val, torch.Tensor) and val.ndim > 0 else None
    val_scalar = 0.0 if val_t is not None else float(val)
    chan = () if val_t is None else tuple(val_t.shape[1:])
    if mesh is None:
        mesh = torch.zeros(tuple(conf.mesh_shape) + chan, dtype=torch.float32, device=disp.device)
    if tuple(mesh.shape[3:]) != chan:
        raise ValueError(f"channel shape mismatch: {tuple(mesh.shape[3:])} != {chan}")
    valid = None if valid_mask is None else valid_mask.detach()
    if valid is not None and val_t is not None:
        val_t = torch.where(valid.reshape(valid.shape + (1,) * (val_t.ndim - 1)), val_t, torch.zeros_like(val_t))
    return _ScatterFn.apply(disp, mesh, val_t, pmid.detach(), val_scalar, offset, conf.cell_size, cell_size,
                            tuple(conf.mesh_shape), valid)


def scatter(ptcl, conf, mesh=None, val=None, offset=0, cell_size=None):
    """Scatter particle values to the mesh multilinearly (scatter.py:235-275).

    ``val`` defaults to ``conf.mesh_size / conf.ptcl_num`` on valid particles, so a uniform
    lattice scatters to a mesh of mean 1; ``mesh`` defaults to zeros.
    """
    valid_mask = None if ptcl.unused_index is None else ~ptcl.unused_index.detach()
    if val is None:
        val = conf.mesh_size / conf.ptcl_num
        if valid_mask is None:
            valid = torch.ones(ptcl.disp.shape[0], dtype=torch.float32, device=ptcl.disp.device)
        else:
            valid = valid_mask.to(torch.float32)
        val = valid * val
    return _scatter(ptcl.pmid, ptcl.disp, conf, mesh, val, offset, cell_size, valid_mask)
