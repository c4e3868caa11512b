"""Writes the golden fixtures of the PM force path (``golden_*.npz`` beside this script).

Every expected output here is computed by a brute-force routine written for this script
alone -- per-particle/per-corner Python loops for CIC, explicit O(N^2) DFT sums for the
transforms, Python's ``sorted`` for the stable sort -- so a fixture is an independent pin of
the float64 oracle (``tests/test_golden.py``) and of the HIP library (``tests/test_gpu_golden.py``).
The reference stores no golden vectors of its own for this path; the inputs restate the ones
its tests build:

* CIC: ``tests/test_pallas_cic.py:22-28`` (pmid rows, disp ~ U(-0.35, 0.35), val = 1..4 on a
  4^3 grid) and ``:69-83`` (8^3 mesh, offset (1,0,0), explicit mesh cell 0.5), with numpy's
  PCG64 in place of JAX's PRNG.
* sort: keys of the containing cell's brick, duplicates and empty bins (the contract of
  ``docs/source/internals/architecture.md:224-241``).
* gravity: ``nbody/gravity.py:626-640`` on 4^3 particles / 8^3 mesh with the particle-Nyquist
  filter active (``mesh_shape=2``, ``nbody/gravity.py:174-196``).

    python tests/golden/make_golden.py        # rewrites the .npz files
"""

from __future__ import annotations

import itertools
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EPS32 = float(np.finfo(np.float32).eps)


# ----------------------------------------------------------------- CIC, brute force ----
def _corners_periodic(y, n):
    """Textbook CIC of one particle at grid coordinate y (3,) on a periodic grid n (3,)."""
    base = [math.floor(y[a]) for a in range(3)]
    f = [y[a] - base[a] for a in range(3)]
    out = []
    for bits in itertools.product((0, 1), repeat=3):  # any order: sums only
        w = 1.0
        idx = []
        for a in range(3):
            w *= f[a] if bits[a] else 1.0 - f[a]
            idx.append((base[a] + bits[a]) % n[a])
        out.append((tuple(idx), w))
    return out


def _grid_coord(pmid, disp, a1, period, offset, a2):
    """Grid coordinate of a particle on the mesh of cell a2 whose origin sits at `offset`."""
    y = []
    for a in range(3):
        x = (pmid[a] * a1 + disp[a] - offset[a]) % period[a]
        y.append(x / a2)
    return y


def bf_scatter(pmid, disp, val, a1, period, offset, a2, shape):
    mesh = np.zeros(shape, dtype=np.float64)
    for p in range(len(pmid)):
        y = _grid_coord(pmid[p], disp[p], a1, period, offset, a2)
        for idx, w in _corners_periodic(y, shape):
            mesh[idx] += val[p] * w
    return mesh


def bf_gather(pmid, disp, mesh, a1, period, offset, a2):
    shape = mesh.shape[:3]
    out = np.zeros((len(pmid),) + mesh.shape[3:], dtype=np.float64)
    for p in range(len(pmid)):
        y = _grid_coord(pmid[p], disp[p], a1, period, offset, a2)
        for idx, w in _corners_periodic(y, shape):
            out[p] += w * mesh[idx]
    return out


# ----------------------------------------------------------------- DFT, brute force ----
def bf_dft3(x, inverse=False):
    """Full complex 3-D DFT by explicit sums over all index pairs (O(N^2))."""
    shape = x.shape
    idx = np.array(list(itertools.product(*[range(s) for s in shape])), dtype=np.float64)
    phase = np.zeros((len(idx), len(idx)))
    for a, s in enumerate(shape):
        phase += np.outer(idx[:, a], idx[:, a]) / s
    sgn = 1.0 if inverse else -1.0
    mat = np.exp(sgn * 2j * np.pi * phase)
    out = mat @ x.reshape(-1).astype(np.complex128)
    if inverse:
        out /= len(idx)
    return out.reshape(shape)


# ---------------------------------------------------------------------- fixtures ----
def golden_cic():
    rng = np.random.default_rng(17)
    pmid = np.array([[0, 0, 0], [1, 2, 3], [3, 0, 2], [2, 3, 1]], dtype=np.int16)
    disp = rng.uniform(-0.35, 0.35, (4, 3)).astype(np.float32)
    mesh = rng.normal(size=(4, 4, 4, 2)).astype(np.float32)
    val = np.arange(1, 5, dtype=np.float32)
    pm64, dp64 = pmid.astype(np.int64), disp.astype(np.float64)
    period = (4.0, 4.0, 4.0)
    zero = (0.0, 0.0, 0.0)
    gathered = bf_gather(pm64, dp64, mesh.astype(np.float64), 1.0, period, zero, 1.0)
    scattered = bf_scatter(pm64, dp64, val.astype(np.float64), 1.0, period, zero, 1.0, (4, 4, 4))
    # explicit mesh cell: 8^3 mesh of cell 0.5 over the same 4^3 periodic box, origin shifted by (1,0,0)
    mesh8 = np.random.default_rng(9).normal(size=(8, 8, 8)).astype(np.float32)
    off = (1.0, 0.0, 0.0)
    gathered8 = bf_gather(pm64, dp64, mesh8.astype(np.float64), 1.0, period, off, 0.5)
    scattered8 = bf_scatter(pm64, dp64, val.astype(np.float64), 1.0, period, off, 0.5, (8, 8, 8))
    np.savez(os.path.join(HERE, "golden_cic.npz"), pmid=pmid, disp=disp, mesh=mesh, val=val, gathered=gathered,
             scattered=scattered, mesh8=mesh8, offset8=np.array(off), gathered8=gathered8, scattered8=scattered8)


def golden_sort():
    rng = np.random.default_rng(5)
    mesh_shape, cell, brick = (8, 4, 16), 0.5, (2, 2, 4)
    n = 300
    pmid = rng.integers(-3, 20, (n, 3)).astype(np.int16)
    pmid[:40] = pmid[0]  # a crowded cell
    disp = rng.uniform(-0.75, 0.75, (n, 3)).astype(np.float32)
    disp[50:60] = 0.0  # particles exactly on cell corners
    disp[60:70, 0] = -0.5  # exactly on a lower cell face: floor(-1) = -1
    valid = rng.random(n) > 0.15
    nb = [m // b for m, b in zip(mesh_shape, brick)]
    keys = np.empty(n, dtype=np.int64)
    for i in range(n):
        c = []
        for a in range(3):
            t = float(np.float32(disp[i, a]) / np.float32(cell))  # the containing cell, in float32
            base = int(pmid[i, a]) + math.floor(t)
            c.append((base % mesh_shape[a]) // brick[a])
        keys[i] = (c[0] * nb[1] + c[1]) * nb[2] + c[2]
    sentinel = nb[0] * nb[1] * nb[2]
    keys_v = np.where(valid, keys, sentinel)
    perm = np.array(sorted(range(n), key=lambda i: (keys_v[i], i)), dtype=np.int32)
    perm_all = np.array(sorted(range(n), key=lambda i: (keys[i], i)), dtype=np.int32)
    bins = np.array([sum(1 for k in keys_v if k < b) for b in range(sentinel + 2)], dtype=np.int32)
    np.savez(os.path.join(HERE, "golden_sort.npz"), pmid=pmid, disp=disp, valid=valid,
             mesh_shape=np.array(mesh_shape), cell=np.float32(cell), brick=np.array(brick), keys=keys,
             perm=perm, perm_all=perm_all, bins=bins)


def golden_fft():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(4, 8, 16)).astype(np.float32)
    X = bf_dft3(x.astype(np.float64))[:, :, : 16 // 2 + 1]
    np.savez(os.path.join(HERE, "golden_fft.npz"), x=x, X=X)


def golden_gravity():
    ng, ratio, spacing, omega = 4, 2, 2.0, 0.3
    nm = ng * ratio
    cell = spacing * ng / nm
    L = spacing * ng
    rng = np.random.default_rng(11)
    q = np.array(list(itertools.product(range(ng), repeat=3)), dtype=np.int64) * ratio  # pmid on the mesh grid
    disp = (rng.normal(size=q.shape) * 0.4 * spacing).astype(np.float32)
    pmid = q.astype(np.int16)
    n = len(pmid)
    period = (L, L, L)
    rho = bf_scatter(q, disp.astype(np.float64), np.full(n, nm ** 3 / n), cell, period, (0.0,) * 3, cell, (nm,) * 3)
    src = 1.5 * omega * (rho - 1.0)
    S = bf_dft3(src)
    k1 = [2 * np.pi * (((i + nm // 2) % nm) - nm // 2) / (nm * cell) for i in range(nm)]
    kmax = np.pi / spacing * (1 + 8 * EPS32)
    knyq = np.pi / cell
    acc = np.zeros((n, 3))
    for comp in range(3):
        A = np.zeros_like(S)
        for i, j, l in itertools.product(range(nm), repeat=3):
            kv = (k1[i], k1[j], k1[l])
            k2 = sum(v * v for v in kv)
            if k2 == 0 or any(abs(v) > kmax for v in kv):
                continue
            kc = kv[comp]
            if abs(abs(kc) - knyq) <= EPS32 * knyq:
                continue  # Nyquist derivative is zero
            A[i, j, l] = (-1j * kc) * (-S[i, j, l] / k2)
        F = bf_dft3(A, inverse=True).real
        acc[:, comp] = bf_gather(q, disp.astype(np.float64), F, cell, period, (0.0,) * 3, cell)
    np.savez(os.path.join(HERE, "golden_gravity.npz"), pmid=pmid, disp=disp, omega_m=omega,
             ptcl_grid=ng, mesh=nm, ptcl_spacing=spacing, acc=acc)


if __name__ == "__main__":
    golden_cic()
    golden_sort()
    golden_fft()
    golden_gravity()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
