"""Particle state: integer mesh anchors plus float displacements (single device).

Mirrors ``src/pmpp/nbody/particles.py``: the fields of ``Particles`` (:33-127), ``gen_grid``
(:615-727, single-device branch), ``from_pos`` (:440-478) and ``pos`` (:773-798).  Arrays are
torch tensors; ``pmid`` is (N,3) int16, ``disp``/``vel``/``acc`` are (N,3) float32 and
``unused_index`` is (N,) bool (True marks padding; all False on one device, :686-692).
"""

from __future__ import annotations

from dataclasses import dataclass, replace as dc_replace
from typing import Any, Optional

import numpy as np
import torch


@dataclass(frozen=True, eq=False)
class Particles:
    conf: Any
    pmid: torch.Tensor
    disp: torch.Tensor
    vel: Optional[torch.Tensor] = None
    acc: Optional[torch.Tensor] = None
    unused_index: Optional[torch.Tensor] = None

    def __post_init__(self):
        object.__setattr__(self, "pmid", self.pmid.to(torch.int16))
        for name in ("disp", "vel", "acc"):
            v = getattr(self, name)
            if v is not None and v.dtype != torch.float32:
                object.__setattr__(self, name, v.to(torch.float32))
        if self.unused_index is not None and self.unused_index.dtype != torch.bool:
            object.__setattr__(self, "unused_index", self.unused_index.to(torch.bool))

    def __len__(self):
        return self.pmid.shape[0]

    def replace(self, **changes):
        return dc_replace(self, **changes)

    @property
    def valid_mask(self):
        return None if self.unused_index is None else ~self.unused_index

    @classmethod
    def gen_grid(cls, conf, vel=False, acc=False, device="cuda"):
        """Particles on the uniform Lagrangian grid with exact integer anchors (:615-727)."""
        pmid, disp = [], []
        for sp, sm in zip(conf.ptcl_grid_shape, conf.mesh_shape):
            p = np.rint(np.linspace(0, sm, num=sp, endpoint=False)).astype(np.int64)
            d = (np.arange(sp) * sm - p * sp).astype(np.float64) * (conf.cell_size / sp)
            pmid.append(p.astype(np.int16))
            disp.append(d.astype(np.float32))
        pm = np.stack(np.meshgrid(*pmid, indexing="ij"), axis=-1).reshape(-1, conf.dim)
        dp = np.stack(np.meshgrid(*disp, indexing="ij"), axis=-1).reshape(-1, conf.dim)
        pm_t = torch.from_numpy(pm).to(device)
        dp_t = torch.from_numpy(dp).to(device)
        unused = torch.zeros(pm.shape[0], dtype=torch.bool, device=device)
        v = torch.zeros_like(dp_t) if vel else None
        a = torch.zeros_like(dp_t) if acc else None
        return cls(conf, pm_t, dp_t, vel=v, acc=a, unused_index=unused)

    @classmethod
    def from_pos(cls, conf, pos, vel=None, acc=None, wrap=True):
        """pmid = rint(pos / cell), disp = pos - pmid * cell (:470-478)."""
        pos = torch.as_tensor(pos)
        cell = conf.cell_size
        pmid = torch.round(pos / cell)
        disp = pos - pmid * cell
        pmid = pmid.to(torch.int16)
        if wrap:
            shape = torch.tensor(conf.mesh_shape, dtype=torch.int16, device=pmid.device)
            pmid = torch.remainder(pmid, shape)
        return cls(conf, pmid, disp.to(torch.float32), vel=vel, acc=acc, unused_index=None)

    def pos(self, dtype=torch.float64, wrap=True):
        """pmid * cell + disp, wrapped into the box (:773-798)."""
        p = self.pmid.to(dtype) * self.conf.cell_size + self.disp.to(dtype)
        if wrap:
            box = torch.tensor(self.conf.box_size, dtype=dtype, device=p.device)
            p = torch.remainder(p, box)
        return p
