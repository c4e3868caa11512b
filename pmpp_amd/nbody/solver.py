"""N-body time integration with a hand-written adjoint (drop-in for ``pmpp.nbody.solver``).

Mirrors ``src/pmpp/nbody/solver.py``: ``nbody_init`` (:21-43) computes the starting force,
``nbody_step`` (:46-63) one macro step, ``nbody`` (:253-273) the scan over ``conf.a_nbody``,
``nbody_adj`` (:385-444) the reverse sweep from the final state, and the public entry point is
differentiable through a custom backward (:447-527) that reconstructs the trajectory
step by step instead of storing it.  Here the custom backward is a
``torch.autograd.Function``: gradients flow to the initial ``disp``, ``vel`` and (when it is
a tensor with ``requires_grad``) ``Omega_m``.
"""

from __future__ import annotations

import torch

from .integrator import AdjointState, force, force_adj, integrate, integrate_adj


def nbody_init(a, ptcl, cosmo, conf, correction=None):
    return force(a, ptcl, cosmo, conf, correction=correction)


def nbody_step(a_prev, a_next, ptcl, cosmo, conf, correction=None):
    return integrate(a_prev, a_next, ptcl, cosmo, conf, correction=correction)


def _schedule(conf, reverse):
    a = conf.a_nbody
    return a[::-1] if reverse else a


def _run(ptcl, cosmo, conf, reverse=False, correction=None, observer=None):
    a = _schedule(conf, reverse)
    ptcl = nbody_init(float(a[0]), ptcl, cosmo, conf, correction=correction)
    if observer is not None:
        observer(float(a[0]), ptcl)
    for a0, a1 in zip(a[:-1], a[1:]):
        ptcl = nbody_step(float(a0), float(a1), ptcl, cosmo, conf, correction=correction)
        if observer is not None:
            observer(float(a1), ptcl)
    return ptcl


def nbody_adj(ptcl, ptcl_cot, cosmo, conf, reverse=False, correction=None, param_grads=None):
    """Reverse sweep from the final state (solver.py:385-444).

    ``ptcl`` is the final state (disp, vel, acc), ``ptcl_cot`` a Particles-like object whose
    ``disp``/``vel``/``acc`` hold the cotangents of that state.  Returns
    ``(ptcl_in, ptcl_cot_in, omega_m_cot)``: the reconstructed initial state, the cotangent
    of the initial state and the Omega_m cotangent.  With a ``correction`` and
    ``param_grads`` (a flat float32 tensor in its parameter layout), the gradient with respect
    to the correction's parameters is added to ``param_grads`` (training through the run).
    """
    a = _schedule(conf, reverse)
    st = AdjointState(ptcl, ptcl_cot.disp, ptcl_cot.vel, ptcl_cot.acc)
    st.param_grads = param_grads
    for a0, a1 in zip(a[:-1][::-1], a[1:][::-1]):
        integrate_adj(float(a0), float(a1), st, cosmo, conf, correction=correction)
    # the initial force (computed by nbody_init before the first step) is pulled through last
    force_adj(float(a[0]), st, cosmo, conf, correction=correction)
    ptcl_in = st.particles()
    cot = ptcl.replace(disp=st.disp_cot, vel=st.vel_cot, acc=st.acc_cot)
    return ptcl_in, cot, st.omega_cot


class _Cosmo:
    def __init__(self, base, omega_m):
        self._base, self.Omega_m = base, omega_m

    def __getattr__(self, name):
        return getattr(self._base, name)


class _NBodyFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, vel, omega_m, ptcl, cosmo, conf, reverse, correction):
        c = _Cosmo(cosmo, float(omega_m))
        p = ptcl.replace(disp=disp.detach().contiguous(), vel=vel.detach().contiguous())
        out = _run(p, c, conf, reverse, correction=correction)
        ctx.state = (out, c, conf, reverse, omega_m, correction)
        return out.disp, out.vel, out.acc

    @staticmethod
    def backward(ctx, disp_cot, vel_cot, acc_cot):
        out, c, conf, reverse, omega_m, correction = ctx.state
        zeros = torch.zeros_like(out.disp)
        cot = out.replace(disp=disp_cot if disp_cot is not None else zeros,
                          vel=vel_cot if vel_cot is not None else zeros,
                          acc=acc_cot if acc_cot is not None else zeros)
        _, cin, om_cot = nbody_adj(out, cot, c, conf, reverse, correction=correction)
        om = torch.as_tensor(om_cot, dtype=omega_m.dtype, device=omega_m.device).reshape(omega_m.shape)
        return cin.disp, cin.vel, om, None, None, None, None, None


def nbody(ptcl, cosmo, conf, reverse=False, correction=None, observer=None):
    """Integrate ``ptcl`` over ``conf.a_nbody`` with KDK steps; returns the final particles.

    ``ptcl.vel`` must be set (zeros for a cold start).  Differentiable with respect to
    ``ptcl.disp``, ``ptcl.vel`` and a tensor ``cosmo.Omega_m`` through the reverse sweep.
    A ``correction`` (a learned force correction, ``pmpp_amd.corrections``) is applied at every
    force; the reverse sweep then also runs through the network and its density input.
    ``observer(a, ptcl)`` (optional, forward only) sees the state after the initial force and
    after every step (the reference's observers, solver.py:21-63).
    """
    if ptcl.vel is None:
        raise ValueError("nbody needs particle velocities (ptcl.vel)")
    if observer is not None:
        om = cosmo.Omega_m
        if ptcl.disp.requires_grad or ptcl.vel.requires_grad or (isinstance(om, torch.Tensor) and om.requires_grad):
            raise NotImplementedError("reverse mode with an observer is not implemented")
        with torch.no_grad():
            return _run(ptcl, cosmo, conf, reverse, correction=correction, observer=observer)
    om = cosmo.Omega_m
    om_t = om if isinstance(om, torch.Tensor) else torch.tensor(float(om), dtype=torch.float64,
                                                                device=ptcl.disp.device)
    disp, vel, acc = _NBodyFn.apply(ptcl.disp, ptcl.vel, om_t, ptcl, cosmo, conf, reverse, correction)
    return ptcl.replace(disp=disp, vel=vel, acc=acc)
