"""Frozen configuration: the geometry/precision/schedule subset of PM++'s Configuration.

Mirrors ``src/pmpp/core/configuration.py`` (field names, defaults and validation messages of
:172-304, the derived properties :463-574) for the single-device force path.  Multi-GPU fields,
transfer-function and LPT knobs that the single-device force path does not use are not carried.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional, Tuple, Union

import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)


def _fftfreq_angular(n: int, spacing: float, last: bool) -> np.ndarray:
    """numerics/fft.py:30-41: numpy fftfreq/rfftfreq times 2 pi / spacing, in float64."""
    period = 2 * np.pi / spacing
    return (np.fft.rfftfreq(n) if last else np.fft.fftfreq(n)) * period


@dataclass(frozen=True, eq=False)
class Configuration:
    ptcl_spacing: float
    ptcl_grid_shape: Tuple[int, ...]
    mesh_shape: Union[float, Tuple[int, ...]] = 1
    a_start: float = 1 / 64
    a_stop: float = 1.0
    a_nbody_maxstep: float = 1 / 64
    a_lpt_maxstep: float = 1 / 128
    a_custom: Optional[Tuple[float, ...]] = None
    symp_splits: Tuple[Tuple[float, float], ...] = ((0, 0.5), (1, 0.5))
    lpt_order: int = 2
    # cell-sort / CIC-tile granularity (cells per brick along each axis); None: (16, 16, 32)
    brick: Optional[Tuple[int, int, int]] = None
    # the PM step is float32 with int16 anchors, as PM++'s defaults (configuration.py:183-185)
    float_dtype: str = "float32"
    pmid_dtype: str = "int16"
    init_mesh_shape: Union[float, Tuple[int, ...], None] = field(default=None, repr=False)

    def __post_init__(self):
        object.__setattr__(self, "ptcl_grid_shape", tuple(int(s) for s in self.ptcl_grid_shape))
        if isinstance(self.mesh_shape, (int, float)):
            object.__setattr__(self, "init_mesh_shape", self.mesh_shape)
            ms = tuple(round(s * self.mesh_shape) for s in self.ptcl_grid_shape)
            object.__setattr__(self, "mesh_shape", ms)
        else:
            object.__setattr__(self, "mesh_shape", tuple(int(s) for s in self.mesh_shape))
        if len(self.ptcl_grid_shape) != len(self.mesh_shape):
            raise ValueError("particle and mesh grid dimensions differ")
        if any(sm < sp for sp, sm in zip(self.ptcl_grid_shape, self.mesh_shape)):
            raise ValueError("mesh grid cannot be smaller than particle grid")
        if any(self.ptcl_grid_shape[0] * sm != self.mesh_shape[0] * sp
               for sp, sm in zip(self.ptcl_grid_shape[1:], self.mesh_shape[1:])):
            raise ValueError("particle and mesh grid aspect ratios differ")
        if self.float_dtype != "float32":
            raise NotImplementedError("the HIP PM step computes in float32 (as PM++'s GPU kernels)")
        if any(len(s) != 2 for s in self.symp_splits):
            raise ValueError(f"symp_splits={self.symp_splits} not supported")
        sums = tuple(sum(s) for s in zip(*self.symp_splits))
        if sums != (1, 1):
            raise ValueError(f"sum of symplectic splits = {sums} != (1, 1)")
        if self.a_custom is not None:
            object.__setattr__(self, "a_custom", tuple(float(a) for a in self.a_custom))
        if self.brick is None:
            object.__setattr__(self, "brick", default_brick(self.mesh_shape))
        else:
            object.__setattr__(self, "brick", tuple(int(b) for b in self.brick))

    def replace(self, **changes):
        if "mesh_shape" not in changes and self.init_mesh_shape is not None:
            changes["mesh_shape"] = self.mesh_shape
        return replace(self, **changes)

    # -- geometry (configuration.py:463-507)
    @property
    def dim(self) -> int:
        return len(self.ptcl_grid_shape)

    @property
    def ptcl_num(self) -> int:
        return math.prod(self.ptcl_grid_shape)

    @property
    def box_size(self):
        return tuple(self.ptcl_spacing * s for s in self.ptcl_grid_shape)

    @property
    def box_vol(self) -> float:
        return float(np.prod(self.box_size))

    @property
    def cell_size(self) -> float:
        return self.ptcl_spacing * self.ptcl_grid_shape[0] / self.mesh_shape[0]

    @property
    def mesh_size(self) -> int:
        return math.prod(self.mesh_shape)

    @property
    def nMesh(self) -> int:
        return self.mesh_shape[0]

    # -- spectral caches (configuration.py:301-307, 365-376; utils.py:15-37), float32 values
    @property
    def kvec(self):
        return tuple(_fftfreq_angular(n, self.cell_size, a == self.dim - 1).astype(np.float32)
                     for a, n in enumerate(self.mesh_shape))

    @property
    def mask_active(self) -> bool:
        return tuple(self.mesh_shape) != tuple(self.ptcl_grid_shape)

    @property
    def particle_nyquist_masks(self):
        if not self.mask_active:
            return ()
        k_nyq = np.float32(np.pi / self.ptcl_spacing)
        limit = k_nyq + k_nyq * np.float32(8 * F32_EPS)
        return tuple((np.abs(k) <= limit).astype(np.float32) for k in self.kvec)

    @property
    def neg_ik(self):
        nyq = np.float32(np.pi / self.cell_size)
        eps = nyq * np.float32(F32_EPS)
        return tuple((1j * np.where(np.abs(np.abs(k) - nyq) <= eps, np.float32(0), -k)).astype(np.complex64)
                     for k in self.kvec)

    # -- time schedule (configuration.py:530-555)
    @property
    def a_nbody_num(self) -> int:
        return math.ceil((self.a_stop - self.a_start) / self.a_nbody_maxstep)

    @property
    def a_nbody(self) -> np.ndarray:
        if self.a_custom is not None:
            return np.asarray(self.a_custom, dtype=np.float64)
        return np.linspace(self.a_start, self.a_stop, num=1 + self.a_nbody_num, dtype=np.float64)

    @property
    def a_lpt_num(self) -> int:
        return math.ceil(self.a_start / self.a_lpt_maxstep)

    @property
    def a_lpt_step(self) -> float:
        return self.a_start / self.a_lpt_num

    @property
    def a_lpt(self) -> np.ndarray:
        return np.linspace(0.0, self.a_start, num=self.a_lpt_num + 1, dtype=np.float64)

    @property
    def growth_a(self) -> np.ndarray:
        """Nodes of the growth table: LPT steps then N-body steps (configuration.py:558-563)."""
        return np.concatenate((self.a_lpt, self.a_nbody[1:]))


def default_brick(mesh_shape) -> Tuple[int, int, int]:
    """(16, 16, 32)-cell bricks (clipped to the mesh): a 17x17x34 int64 LDS tile for the
    fixed-point scatter (78.6 KB, two workgroups per CU), 128-byte rows along z, and 2^17
    bricks on a 1024^3 mesh (two 9-bit radix passes)."""
    return tuple(min(int(n), b) for n, b in zip(mesh_shape, (16, 16, 32)))
