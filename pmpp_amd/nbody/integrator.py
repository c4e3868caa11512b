"""Symplectic kick-drift-kick integration on the HIP library, and its hand-written reverse.

Mirrors ``src/pmpp/nbody/integrator.py``:

* ``drift_factor`` / ``kick_factor`` (:125-170): growth-matched step factors
  (D(a1) - D(a0)) / G_D(a_vel) and (G_D(a1) - G_D(a0)) / G_K(a_acc), evaluated in float64 on
  the host from the interpolated growth table and cast to float32 as the reference does.
* ``drift`` (:203-237), ``kick`` (:512-538), ``force`` (:585-617), ``integrate`` (:689-731)
  with the configuration's ``symp_splits``.
* ``kick_adj`` (:541-582), ``drift_adj_from_output`` (:429-474), ``force_adj`` (:620-686) and
  ``integrate_adj`` (:781-837): the reverse sweep reconstructs each earlier state from the
  later one (no stored trajectory) and recomputes the force the earlier kick consumed.

The elementwise updates run in ``pm_kick`` / ``pm_drift`` (and their adjoints); forces in
``pm_gravity`` / ``pm_gravity_vjp``.  The cosmology cotangent is carried for ``Omega_m``
(through the force's 1.5 Omega_m source and through the step factors, whose derivatives
are central differences of the float64 factor functions).
"""

from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from .. import _lib
from .._torch import as_contig, ensure_init, ptr, stream_handle, workspace
from ..cosmology import GrowthTable

_TABLES: Dict[tuple, GrowthTable] = {}
_FD_REL = 1e-6  # relative step of the factor derivatives in Omega_m


def _cosmo_key(cosmo):
    return (float(cosmo.Omega_m), float(getattr(cosmo, "Omega_k", 0.0)), float(getattr(cosmo, "w_0", -1.0)),
            float(getattr(cosmo, "w_a", 0.0)))


class _Background:
    """Omega_m, Omega_k, w_0, w_a view used by the growth table (the rest does not enter)."""

    def __init__(self, om, ok, w0, wa):
        self.Omega_m, self.Omega_k, self.w_0, self.w_a = om, ok, w0, wa

    @property
    def Omega_de(self):
        return 1.0 - self.Omega_m - self.Omega_k


def growth_table(cosmo, conf, omega_m=None) -> GrowthTable:
    """The growth table at ``conf.growth_a`` (cached per background and node set)."""
    om, ok, w0, wa = _cosmo_key(cosmo)
    if omega_m is not None:
        om = float(omega_m)
    key = (om, ok, w0, wa, tuple(np.asarray(conf.growth_a).tolist()), float(conf.a_lpt_step))
    tab = _TABLES.get(key)
    if tab is None:
        tab = GrowthTable(_Background(om, ok, w0, wa), conf.growth_a, conf.a_lpt_step)
        _TABLES[key] = tab
    return tab


def drift_factor(a_vel, a_prev, a_next, cosmo, conf) -> float:
    return growth_table(cosmo, conf).drift_factor(a_vel, a_prev, a_next)


def kick_factor(a_acc, a_prev, a_next, cosmo, conf) -> float:
    return growth_table(cosmo, conf).kick_factor(a_acc, a_prev, a_next)


def _factor_dOm(kind, a0, a1, a2, cosmo, conf) -> float:
    """d factor / d Omega_m by a central difference of the float64 factor function."""
    om = float(cosmo.Omega_m)
    h = _FD_REL * max(om, 1e-3)
    fp = getattr(growth_table(cosmo, conf, om + h), kind)(a0, a1, a2)
    fm = getattr(growth_table(cosmo, conf, om - h), kind)(a0, a1, a2)
    return (fp - fm) / (2 * h)


def _f32(x) -> float:
    return float(np.float32(x))


def _n(ptcl) -> int:
    return int(ptcl.pmid.shape[0])


def drift(a_vel, a_prev, a_next, ptcl, cosmo, conf):
    """disp += D vel (integrator.py:203-237), into a new array (one pass)."""
    src = as_contig(ptcl.disp, torch.float32, "disp")
    disp = torch.empty_like(src)
    _lib.call("pm_drift", ptr(src), ptr(as_contig(ptcl.vel, torch.float32, "vel")), ptr(disp), _n(ptcl),
              _f32(drift_factor(a_vel, a_prev, a_next, cosmo, conf)), stream_handle(disp.device))
    return ptcl.replace(disp=disp)


def kick(a_acc, a_prev, a_next, ptcl, cosmo, conf):
    """vel += K acc (integrator.py:512-538), into a new array (one pass)."""
    src = as_contig(ptcl.vel, torch.float32, "vel")
    vel = torch.empty_like(src)
    _lib.call("pm_kick", ptr(src), ptr(as_contig(ptcl.acc, torch.float32, "acc")), ptr(vel), _n(ptcl),
              _f32(kick_factor(a_acc, a_prev, a_next, cosmo, conf)), stream_handle(vel.device))
    return ptcl.replace(vel=vel)


def force(a, ptcl, cosmo, conf, correction=None):
    """Overwrite ``acc`` with the PM acceleration at ``a`` (integrator.py:600-617)."""
    from .gravity import gravity

    with torch.no_grad():
        acc = gravity(a, ptcl, cosmo, conf, correction=correction)
    return ptcl.replace(acc=acc)


def stage_schedule(a_prev, a_next, conf):
    """The substeps of one macro step, as recorded by integrator.py:751-770."""
    D = K = 0.0
    a_disp = a_vel = a_acc = a_prev
    stages = []
    for d, k in conf.symp_splits:
        if d != 0:
            D += d
            a_disp_next = a_prev * (1 - D) + a_next * D
            stages.append(("drift", a_vel, a_disp, a_disp_next, a_acc))
            a_disp = a_disp_next
            a_acc = a_disp
        if k != 0:
            K += k
            a_vel_next = a_prev * (1 - K) + a_next * K
            stages.append(("kick", a_acc, a_vel, a_vel_next))
            a_vel = a_vel_next
    return stages


def integrate(a_prev, a_next, ptcl, cosmo, conf, correction=None):
    """One macro step with the configured splitting (integrator.py:689-731)."""
    for st in stage_schedule(a_prev, a_next, conf):
        if st[0] == "drift":
            _, a_vel, a_disp, a_disp_next, _ = st
            ptcl = drift(a_vel, a_disp, a_disp_next, ptcl, cosmo, conf)
            ptcl = force(a_disp_next, ptcl, cosmo, conf, correction=correction)
        else:
            _, a_acc, a_vel, a_vel_next = st
            ptcl = kick(a_acc, a_vel, a_vel_next, ptcl, cosmo, conf)
    return ptcl


class AdjointState:
    """Mutable reverse-sweep state: particles, their cotangents and the Omega_m cotangent."""

    def __init__(self, ptcl, disp_cot, vel_cot, acc_cot):
        dev = ptcl.disp.device
        ensure_init(dev)
        self.ptcl = ptcl
        self.disp = ptcl.disp.clone().contiguous()
        self.vel = ptcl.vel.clone().contiguous()
        self.acc = ptcl.acc.clone().contiguous()
        self.disp_cot = disp_cot.clone().contiguous()
        self.vel_cot = vel_cot.clone().contiguous()
        self.acc_cot = acc_cot.clone().contiguous()
        # the Omega_m cotangent accumulates on the device (float64): reading it back after every
        # stage would stall the stream between the per-step kernels
        self.omega_cot_dev = torch.zeros((), dtype=torch.float64, device=dev)
        self.dot = torch.zeros(1, dtype=torch.float64, device=dev)
        nbytes = _lib.size_query("pm_nbody_adj_workspace")
        self.ws = workspace(nbytes, dev)
        self.ws_bytes = nbytes
        self.stream = stream_handle(dev)
        from .gravity import KeptForce

        self.kept = KeptForce()  # forward meshes of the last force recompute, for the next VJP
        from ..corrections.training import KeptCorrection

        self.kept_corr = KeptCorrection()  # the same for a corrected force: density, residual, tape
        # with a learned correction: accumulate the gradient of its parameters here (flat
        # float32, the parameter layout) when set
        self.param_grads = None

    def particles(self):
        return self.ptcl.replace(disp=self.disp, vel=self.vel, acc=self.acc)

    @property
    def omega_cot(self) -> float:
        """The accumulated Omega_m cotangent (one host read)."""
        return float(self.omega_cot_dev.item())

    def add_dot(self, coef: float) -> None:
        """omega_cot += coef * (the factor cotangent the last adjoint update wrote), on the device."""
        self.omega_cot_dev.add_(self.dot[0], alpha=float(coef))


def kick_adj(a_acc, a_prev, a_next, st: AdjointState, cosmo, conf):
    """Reverse a kick: vel -= K acc, acc_cot += K vel_cot (integrator.py:541-582)."""
    K = kick_factor(a_acc, a_prev, a_next, cosmo, conf)
    _lib.call("pm_kick_adj", ptr(st.vel), ptr(st.acc), ptr(st.vel_cot), ptr(st.acc_cot), _n(st.ptcl), _f32(K),
              ptr(st.dot), ptr(st.ws), st.ws_bytes, st.stream)
    st.add_dot(_factor_dOm("kick_factor", a_acc, a_prev, a_next, cosmo, conf))


def drift_adj_from_output(a_vel, a_prev, a_next, st: AdjointState, cosmo, conf):
    """Reverse a drift from its output: disp -= D vel, vel_cot += D disp_cot (:429-474)."""
    D = drift_factor(a_vel, a_prev, a_next, cosmo, conf)
    st.kept.key = None  # the positions change under any kept force meshes
    st.kept_corr.clear()
    _lib.call("pm_drift_adj", ptr(st.disp), ptr(st.vel), ptr(st.disp_cot), ptr(st.vel_cot), _n(st.ptcl), _f32(D),
              ptr(st.dot), ptr(st.ws), st.ws_bytes, st.stream)
    st.add_dot(_factor_dOm("drift_factor", a_vel, a_prev, a_next, cosmo, conf))


def force_adj(a, st: AdjointState, cosmo, conf, correction=None):
    """Pull acc_cot through the force at the current state (integrator.py:620-686); with a
    learned correction also through the network and its density input."""
    if correction is None:
        from .gravity import gravity_vjp

        d_disp, d_om = gravity_vjp(st.particles(), cosmo, conf, st.acc_cot, kept=st.kept, device_omega=True)
    else:
        from ..corrections.training import corrected_gravity_vjp

        d_disp, d_om = corrected_gravity_vjp(a, st.particles(), cosmo, conf, correction, st.acc_cot,
                                             param_grads=st.param_grads,
                                             kept=None if st.param_grads is not None else st.kept_corr)
    st.disp_cot += d_disp
    st.omega_cot_dev += d_om
    st.acc_cot.zero_()


def integrate_adj(a_prev, a_next, st: AdjointState, cosmo, conf, correction=None):
    """Reverse one macro step from its output state (integrator.py:781-837)."""
    for stage in reversed(stage_schedule(a_prev, a_next, conf)):
        if stage[0] == "kick":
            _, a_acc, a_vel, a_vel_next = stage
            kick_adj(a_acc, a_vel, a_vel_next, st, cosmo, conf)
            continue
        _, a_vel, a_disp, a_disp_next, a_acc_in = stage
        force_adj(a_disp_next, st, cosmo, conf, correction=correction)
        drift_adj_from_output(a_vel, a_disp, a_disp_next, st, cosmo, conf)
        # restore the acceleration the earlier kick consumed; uncorrected, the force keeps its
        # meshes for the VJP at these same positions in the next stage
        if correction is None:
            from .gravity import gravity_keep

            st.acc = gravity_keep(st.particles(), cosmo, conf, st.kept).contiguous()
        else:
            from ..corrections.training import corrected_gravity

            # the density, residual and network tape stay for the next stage's VJP at these positions
            keep = st.kept_corr if st.param_grads is None else None
            with torch.no_grad():
                st.acc = corrected_gravity(a_acc_in, st.particles(), cosmo, conf, correction, kept=keep).contiguous()
