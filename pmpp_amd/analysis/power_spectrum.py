"""Integer-shell power spectra and cross-correlations (drop-in for ``pmpp.analysis.power_spectrum``).

Mirrors ``src/pmpp/analysis/power_spectrum.py``: shells floor(|n| + 1e-6) of the integer mode
vector up to floor(sqrt(3) N/2) (:38-43), Hermitian multiplicity (:46-71), the CIC (or NGP,
TSC, PCS) window deconvolution sinc^(-2p) (:108-121), normalisation P = sum / nmodes *
cell^3 / N_m over shells 1.. (:251-256), ``delta_to_pk`` (:287-310), ``density_to_pk``
(:348-374) and the cross-correlation r = P_ab / sqrt(max(P_aa P_bb, eps)) (:259-268).  The
FFT is ``pm_rfft3d``; the shell reduction ``pm_power_spectrum`` (float64 sums); the host
keeps the shell centres and mode counts, as the reference does (:46-71).
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
import torch

from .. import _lib
from .._torch import ensure_init, ptr, require_device, stream_handle
from ..numerics import rfftn

_MAS_POWER = {None: 0, "NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}


def _mas(mas):
    key = None if mas is None else str(mas).upper()
    if key not in _MAS_POWER:
        raise ValueError(f"unknown mass-assignment scheme {mas!r}")
    return _MAS_POWER[key]


def num_shells(mesh_shape) -> int:
    """Largest integer shell of the mesh plus one (power_spectrum.py:38-43)."""
    return math.floor(math.sqrt(3) * (int(mesh_shape[0]) // 2)) + 1


@lru_cache(maxsize=None)
def shell_statistics(mesh_shape, cell_size):
    """(k_shell[1:], nmodes[1:]) on the host (power_spectrum.py:46-71)."""
    shape = tuple(int(s) for s in mesh_shape)
    kf = 2 * math.pi / (shape[0] * float(cell_size))
    mx = np.fft.fftfreq(shape[0]) * shape[0]
    my = np.fft.fftfreq(shape[1]) * shape[1]
    mz = np.fft.rfftfreq(shape[2]) * shape[2]
    X, Y, Z = np.meshgrid(mx, my, mz, indexing="ij")
    mag = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    shell = np.floor(mag + 1e-6).astype(np.int64)
    mult = np.full(shell.shape, 2.0)
    mult[..., 0] = 1
    if shape[2] % 2 == 0:
        mult[..., -1] = 1
    ns = int(shell.max()) + 1
    nmodes = np.bincount(shell.ravel(), weights=mult.ravel(), minlength=ns)
    ksum = np.bincount(shell.ravel(), weights=(mag * mult).ravel(), minlength=ns)
    k = ksum / np.maximum(nmodes, 1) * kf
    return k[1:], nmodes[1:].astype(np.int64)


def _shell_sums(a, b, mas, density):
    require_device(a, "mesh")
    if a.ndim != 3 or a.dtype != torch.float32:
        raise NotImplementedError("power spectra take (X, Y, Z) float32 meshes")
    if b is not None and b.shape != a.shape:
        raise ValueError("both meshes must have the same shape")
    dev = a.device
    ensure_init(dev)
    shape = tuple(a.shape)
    ns = num_shells(shape)
    sa = rfftn(a.contiguous())
    sb = rfftn(b.contiguous()) if b is not None else None
    sums = torch.empty(3 * ns if b is not None else ns, dtype=torch.float64, device=dev)
    _lib.call("pm_power_spectrum", ptr(torch.view_as_real(sa)), ptr(torch.view_as_real(sb)) if sb is not None else None,
              _lib.int3(shape), _mas(mas), ns, int(density), ptr(sums), stream_handle(dev))
    return shape, ns, sums.cpu().numpy()


def _finalize(psum, shape, cell_size):
    k, nmodes = shell_statistics(shape, float(cell_size))
    n = min(len(k), len(psum) - 1)
    pk = psum[1:1 + n] / nmodes[:n] * (float(cell_size) ** 3 / float(np.prod(shape)))
    return k[:n], pk, nmodes[:n]


def _to_torch(x, dev):
    return torch.as_tensor(np.asarray(x), device=dev)


def delta_to_pk(delta, conf, mas="CIC"):
    """(k, P(k), nmodes) of an overdensity mesh (power_spectrum.py:287-310)."""
    shape, ns, s = _shell_sums(delta, None, mas, density=False)
    k, pk, nm = _finalize(s, shape, conf.cell_size)
    dev = delta.device
    return _to_torch(k, dev).float(), _to_torch(pk, dev).float(), _to_torch(nm, dev).int()


def density_to_pk(density, conf, mas="CIC"):
    """(k, P(k), nmodes) of a density mesh; delta = rho / mean - 1 (power_spectrum.py:348-374)."""
    shape, ns, s = _shell_sums(density, None, mas, density=True)
    k, pk, nm = _finalize(s, shape, conf.cell_size)
    dev = density.device
    return _to_torch(k, dev).float(), _to_torch(pk, dev).float(), _to_torch(nm, dev).int()


def _cross(a, b, conf, mas, eps, density):
    shape, ns, s = _shell_sums(a, b, mas, density=density)
    k, pa, nm = _finalize(s[:ns], shape, conf.cell_size)
    _, pb, _ = _finalize(s[ns:2 * ns], shape, conf.cell_size)
    _, pab, _ = _finalize(s[2 * ns:], shape, conf.cell_size)
    r = pab / np.sqrt(np.maximum(pa * pb, eps))
    dev = a.device
    f = lambda x: _to_torch(x, dev).float()  # noqa: E731
    return f(k), f(r), f(pab), f(pa), f(pb), _to_torch(nm, dev).int()


def delta_to_cross_correlation(delta_a, delta_b, conf, mas="CIC", eps=1e-30):
    """(k, r, P_ab, P_aa, P_bb, nmodes) of two overdensity meshes (power_spectrum.py:313-345)."""
    return _cross(delta_a, delta_b, conf, mas, eps, density=False)


def density_to_cross_correlation(density_a, density_b, conf, mas="CIC", eps=1e-30):
    """(k, r, P_ab, P_aa, P_bb, nmodes) of two density meshes (power_spectrum.py:377-405)."""
    return _cross(density_a, density_b, conf, mas, eps, density=True)


def particles_to_pk(ptcl, conf, mas="CIC"):
    """CIC density of the particles on the configuration's mesh, then ``density_to_pk``."""
    from ..cic import scatter

    mesh = scatter(ptcl, conf)
    return density_to_pk(mesh, conf, mas=mas)


def _pk_vjp(mesh, conf, pk_cot, mas, density):
    require_device(mesh, "mesh")
    if mesh.ndim != 3 or mesh.dtype != torch.float32:
        raise NotImplementedError("power spectra take (X, Y, Z) float32 meshes")
    from ..numerics import irfftn

    dev = mesh.device
    ensure_init(dev)
    shape = tuple(mesh.shape)
    ns = num_shells(shape)
    _, nmodes = shell_statistics(shape, float(conf.cell_size))
    cot = np.asarray(pk_cot.detach().cpu() if isinstance(pk_cot, torch.Tensor) else pk_cot, dtype=np.float64)
    n = min(len(nmodes), ns - 1)
    if cot.shape != (n,):
        raise ValueError(f"pk_cot must have {n} entries (shells 1..{n})")
    coef = np.zeros(ns, dtype=np.float64)
    coef[1:1 + n] = 2.0 * float(conf.cell_size) ** 3 * cot / nmodes[:n]
    coef_d = torch.from_numpy(coef.astype(np.float32)).to(dev)
    spec = rfftn(mesh.contiguous())
    grad = torch.empty_like(spec)
    _lib.call("pm_power_spectrum_bwd", ptr(torch.view_as_real(spec)), _lib.int3(shape), _mas(mas), ns, int(density),
              ptr(coef_d), ptr(torch.view_as_real(grad)), stream_handle(dev))
    return irfftn(grad, s=shape)


def _cross_vjp(mesh, other, conf, pab_cot, mas, density):
    """d/d(mesh) of <pab_cot, P_ab(k)> with ``other`` (b) fixed: per mode the shell coefficient
    times the window times b's spectrum (half the auto-spectrum's 2 a)."""
    from ..numerics import irfftn

    dev = mesh.device
    shape = tuple(mesh.shape)
    ns = num_shells(shape)
    _, nmodes = shell_statistics(shape, float(conf.cell_size))
    cot = np.asarray(pab_cot, dtype=np.float64)
    n = min(len(nmodes), ns - 1)
    coef = np.zeros(ns, dtype=np.float64)
    coef[1:1 + n] = float(conf.cell_size) ** 3 * cot / nmodes[:n]
    coef_d = torch.from_numpy(coef.astype(np.float32)).to(dev)
    spec = rfftn(other.contiguous())
    grad = torch.empty_like(spec)
    _lib.call("pm_power_spectrum_bwd", ptr(torch.view_as_real(spec)), _lib.int3(shape), _mas(mas), ns, int(density),
              ptr(coef_d), ptr(torch.view_as_real(grad)), stream_handle(dev))
    return irfftn(grad, s=shape)


def density_to_cross_correlation_vjp(density_a, density_b, conf, r_cot, mas="CIC", eps=1e-30):
    """d/d(density_a) of <r_cot, r(k)> for ``density_to_cross_correlation`` with ``density_b``
    fixed: r = P_ab / sqrt(P_aa P_bb), so dr = dP_ab / sqrt(P_aa P_bb) - r dP_aa / (2 P_aa)
    (up to a constant field, as ``density_to_pk_vjp``)."""
    _, r, pab, pa, pb, _ = density_to_cross_correlation(density_a, density_b, conf, mas, eps)
    r, pa, pb = (x.double().cpu().numpy() for x in (r, pa, pb))
    rc = np.asarray(r_cot.detach().cpu() if isinstance(r_cot, torch.Tensor) else r_cot, dtype=np.float64)
    c_ab = rc / np.sqrt(np.maximum(pa * pb, eps))
    c_aa = -rc * r / (2.0 * np.maximum(pa, eps))
    return (_pk_vjp(density_a, conf, c_aa, mas, density=True) +
            _cross_vjp(density_a, density_b, conf, c_ab, mas, density=True))


def delta_to_pk_vjp(delta, conf, pk_cot, mas="CIC"):
    """d/d(delta) of <pk_cot, P(k)> for ``delta_to_pk`` (the reverse of the shell sums)."""
    return _pk_vjp(delta, conf, pk_cot, mas, density=False)


def density_to_pk_vjp(density, conf, pk_cot, mas="CIC"):
    """d/d(density) of <pk_cot, P(k)> for ``density_to_pk``, up to a constant field (the
    derivative through the mean density, which moves no particle)."""
    return _pk_vjp(density, conf, pk_cot, mas, density=True)
