"""CIC gather with a hand-written VJP (drop-in for ``pmpp.cic.gather``).

Follows ``src/pmpp/cic/gather.py``: ``gather`` (:323-360), the custom-VJP primitive
``_gather`` (:446-449) and its rules (:511-549), on the HIP kernels.
"""

from __future__ import annotations

import torch

from .kernels import hip_gather, hip_gather_bwd


class _GatherFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, mesh, pmid, offset, particle_cell, cell_size, global_shape, valid):
        out = hip_gather(pmid, disp, mesh, offset=offset, particle_cell_size=particle_cell, cell_size=cell_size,
                         global_shape=global_shape, valid_mask=valid)
        ctx.save_for_backward(disp, mesh, pmid, valid)
        ctx.meta = (offset, particle_cell, cell_size, global_shape)
        return out

    @staticmethod
    def backward(ctx, val_cot):
        disp, mesh, pmid, valid = ctx.saved_tensors
        offset, particle_cell, cell_size, global_shape = ctx.meta
        disp_cot, mesh_cot = hip_gather_bwd(pmid, disp, mesh, val_cot.contiguous(), offset=offset,
                                            particle_cell_size=particle_cell, cell_size=cell_size,
                                            global_shape=global_shape, valid_mask=valid)
        return disp_cot, mesh_cot, None, None, None, None, None, None


def _gather(pmid, disp, conf, mesh, val, offset, cell_size, valid_mask=None):
    """Local multilinear gather primitive with a custom VJP (gather.py:446-449).

    ``val`` is accepted for signature parity and ignored, as in the reference (:466).
    """
    del val
    mesh = torch.as_tensor(mesh)
    valid = None if valid_mask is None else valid_mask.detach()
    return _GatherFn.apply(disp, mesh, pmid.detach(), offset, conf.cell_size, cell_size, tuple(conf.mesh_shape),
                           valid)


def gather(ptcl, conf, mesh):
    """Gather particle values from the mesh multilinearly (gather.py:323-360)."""
    valid_mask = None if ptcl.unused_index is None else ~ptcl.unused_index.detach()
    return _gather(ptcl.pmid.detach(), ptcl.disp, conf, mesh, 0, 0, None, valid_mask)
