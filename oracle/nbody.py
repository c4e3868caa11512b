"""Float64 restatement of PM++'s KDK N-body integration and its reverse sweep.

Test infrastructure only.  Follows (read as text, never imported):

* ``src/pmpp/nbody/integrator.py:114-170``  -- G_D, G_K, drift and kick factors
  (``oracle.cosmology``);
* ``:203-237`` drift, ``:512-538`` kick, ``:585-617`` force, ``:689-731`` integrate with
  ``symp_splits`` and the stage schedule of ``:751-770``;
* ``:429-474`` drift_adj_from_output, ``:541-582`` kick_adj, ``:620-686`` force_adj,
  ``:781-837`` integrate_adj (state reconstruction + force recomputation);
* ``src/pmpp/nbody/solver.py:21-63, 253-273, 385-444`` -- nbody_init, the scan over
  ``a_nbody`` and nbody_adj (the initial force is pulled through last).

The force is ``oracle.spectral.gravity`` (float64) and its VJP ``oracle.spectral.gravity_vjp``.
The Omega_m cotangent of a step factor is its central difference in Omega_m.

Pinned in ``tests/test_oracle_nbody.py``: Einstein-de Sitter closed forms of the factors,
exact reversibility of forward + reconstruction, the adjoint against central finite
differences of the forward map, and linear growth of a single Zel'dovich mode.
"""

from __future__ import annotations

import numpy as np

from . import cosmology as oc
from . import spectral


class Setup:
    """Geometry, cosmology and growth table of one run (all float64)."""

    def __init__(self, mesh_shape, cell_size, ptcl_num, masks, Om, growth_a, a_lpt_step, symp_splits=((0, 0.5), (1, 0.5))):
        self.mesh_shape = tuple(mesh_shape)
        self.cell = float(cell_size)
        self.ptcl_num = int(ptcl_num)
        self.masks = masks
        self.Om = float(Om)
        self.growth_a = np.asarray(growth_a, dtype=np.float64)
        self.a_lpt_step = float(a_lpt_step)
        self.symp_splits = symp_splits
        self._tables = {}

    def table(self, Om=None):
        Om = self.Om if Om is None else Om
        if Om not in self._tables:
            self._tables[Om] = oc.growth_table(self.growth_a, Om, self.a_lpt_step)
        return self._tables[Om]

    def drift_factor(self, av, a0, a1, Om=None):
        Om = self.Om if Om is None else Om
        return float(oc.drift_factor(av, a0, a1, self.table(Om), self.growth_a, Om))

    def kick_factor(self, af, a0, a1, Om=None):
        Om = self.Om if Om is None else Om
        return float(oc.kick_factor(af, a0, a1, self.table(Om), self.growth_a, Om))

    def dfactor(self, kind, a, a0, a1, rel=1e-6):
        h = rel * self.Om
        f = getattr(self, kind)
        return (f(a, a0, a1, self.Om + h) - f(a, a0, a1, self.Om - h)) / (2 * h)

    def force(self, pmid, disp):
        return spectral.gravity(pmid, disp, self.Om, self.mesh_shape, self.cell, self.ptcl_num, self.masks)

    def force_vjp(self, pmid, disp, acc_cot):
        return spectral.gravity_vjp(pmid, disp, self.Om, acc_cot, self.mesh_shape, self.cell, self.ptcl_num,
                                    self.masks)


def stages(a_prev, a_next, symp_splits):
    D = K = 0.0
    a_disp = a_vel = a_acc = a_prev
    out = []
    for d, k in symp_splits:
        if d != 0:
            D += d
            a_disp_next = a_prev * (1 - D) + a_next * D
            out.append(("drift", a_vel, a_disp, a_disp_next, a_acc))
            a_disp = a_disp_next
            a_acc = a_disp
        if k != 0:
            K += k
            a_vel_next = a_prev * (1 - K) + a_next * K
            out.append(("kick", a_acc, a_vel, a_vel_next))
            a_vel = a_vel_next
    return out


def nbody(S: Setup, pmid, disp, vel, a_nbody):
    """Forward integration; returns (disp, vel, acc) at a_nbody[-1]."""
    disp = np.array(disp, dtype=np.float64)
    vel = np.array(vel, dtype=np.float64)
    acc = S.force(pmid, disp)
    for a0, a1 in zip(a_nbody[:-1], a_nbody[1:]):
        for st in stages(a0, a1, S.symp_splits):
            if st[0] == "drift":
                _, av, ap, an, _ = st
                disp = disp + vel * S.drift_factor(av, ap, an)
                acc = S.force(pmid, disp)
            else:
                _, af, ap, an = st
                vel = vel + acc * S.kick_factor(af, ap, an)
    return disp, vel, acc


def nbody_adj(S: Setup, pmid, disp, vel, acc, disp_cot, vel_cot, acc_cot, a_nbody):
    """Reverse sweep from the final state; returns (disp_in, vel_in, disp_cot, vel_cot, Om_cot)."""
    disp, vel, acc = (np.array(x, dtype=np.float64) for x in (disp, vel, acc))
    dc, vc, ac = (np.array(x, dtype=np.float64) for x in (disp_cot, vel_cot, acc_cot))
    om_cot = 0.0
    for a0, a1 in zip(a_nbody[:-1][::-1], a_nbody[1:][::-1]):
        for st in reversed(stages(a0, a1, S.symp_splits)):
            if st[0] == "kick":
                _, af, ap, an = st
                K = S.kick_factor(af, ap, an)
                vel = vel - acc * K
                ac = ac + vc * K
                om_cot += S.dfactor("kick_factor", af, ap, an) * float(np.sum(vc * acc))
                continue
            _, av, ap, an, _ = st
            d_disp, d_om = S.force_vjp(pmid, disp, ac)
            dc = dc + d_disp
            om_cot += d_om
            ac = np.zeros_like(ac)
            D = S.drift_factor(av, ap, an)
            disp = disp - vel * D
            vc = vc + dc * D
            om_cot += S.dfactor("drift_factor", av, ap, an) * float(np.sum(dc * vel))
            acc = S.force(pmid, disp)
    d_disp, d_om = S.force_vjp(pmid, disp, ac)
    dc = dc + d_disp
    om_cot += d_om
    return disp, vel, dc, vc, om_cot
