"""Float64 NumPy restatement of PM++'s CIC geometry, scatter, gather and adjoints.

Follows (read as text, never imported):

* ``src/pmpp/cic/enmesh.py:34-165``        -- ``enmesh`` (both the a2=None and explicit
  cell-size branches; corner order ``bits = (i>>0&1, i>>1&1, i>>2&1)`` at :98).
* ``src/pmpp/cic/pallas.py:157-258``        -- per-lane restatement of the same math; its
  bounds mask (:261-265) and validity semantics (:133-143) define the kernel contract.
* ``src/pmpp/cic/scatter.py:278-338``      -- ``_scatter`` / ``_scatter_chunk``.
* ``src/pmpp/cic/scatter.py:341-437``      -- ``_scatter_chunk_adj`` / ``_scatter_bwd``.
* ``src/pmpp/cic/gather.py:414-472``       -- ``_gather_impl`` / ``_gather_chunk``.
* ``src/pmpp/cic/gather.py:475-549``       -- ``_gather_chunk_adj`` / ``_gather_bwd``.

Pinned in ``tests/test_oracle_cic.py`` by: partition of unity and zero gradient sum
(``tests/test_deep_scatter_fallback.py:75-95``), exact unit density of a uniform lattice
(``:57-62``), brute-force single-particle weights, the adjoint dot-product identities
<scatter(v), M> = <v, gather(M)>, and 4th-order central finite differences of both
adjoints (``tests/test_finite_diff.py:206-255``).
"""

from __future__ import annotations

import numpy as np

# Corner bit patterns in the reference order (enmesh.py:98): corner c has bit
# (c >> axis) & 1 on each axis.
CORNER_BITS = np.array([[(c >> a) & 1 for a in range(3)] for c in range(8)], dtype=np.int64)


def enmesh(pmid, disp, cell_size, mesh_shape, offset=0.0, mesh_cell_size=None, spatial_shape=None,
           grad=False):
    """Multilinear mesh indices, fractions and fraction gradients (enmesh.py:34-165).

    Parameters mirror the reference: ``pmid`` (N,3) integer anchors on grid 1,
    ``disp`` (N,3) physical displacements, ``cell_size`` = a1, ``mesh_shape`` = s1
    (periodic shape of grid 1, or None), ``offset`` = b12, ``mesh_cell_size`` = a2
    (None: grid 2 shares a1), ``spatial_shape`` = s2.

    Returns ``idx`` (N,8,3) int64, ``frac`` (N,8) and optionally ``frac_grad`` (N,8,3),
    all float64.  Indices that fall outside ``spatial_shape`` are returned as-is; the
    callers drop them, as ``.at[].add``/``get(mode='drop')`` and the Pallas bounds mask do.
    """
    i1 = np.asarray(pmid, dtype=np.int64)
    d1 = np.asarray(disp, dtype=np.float64)
    n, dim = i1.shape
    a1 = float(cell_size)
    b12 = np.broadcast_to(np.asarray(offset, dtype=np.float64), (dim,))
    s1 = None if mesh_shape is None else np.asarray(mesh_shape, dtype=np.int64)
    nb = CORNER_BITS[:, :dim]

    if mesh_cell_size is not None:
        # enmesh.py:107-125 -- resample onto a grid of another cell size
        a2 = float(mesh_cell_size)
        P = (i1 * a1 + d1 - b12)[:, None, :]
        i2 = P + nb * a2
        if s1 is not None:
            L = s1 * a1
            i2 = np.mod(i2, L)
        i2 = np.floor(i2 / a2)
        d2 = P - i2 * a2
        if s1 is not None:
            d2 = d2 - np.rint(d2 / L) * L
        i2 = i2.astype(np.int64)
        d2 = d2 / a2
    else:
        # enmesh.py:126-147 -- both grids share the cell size
        i12 = np.floor(b12 / a1)
        d12 = b12 - i12 * a1
        i1 = i1 - i12.astype(np.int64)
        d1 = (d1 - d12) / a1
        d1 = d1[:, None, :]
        i2 = np.floor(d1).astype(np.int64) + nb
        d2 = d1 - i2
        i2 = i2 + i1[:, None, :]
        if s1 is not None:
            i2 = np.mod(i2, s1)

    f2 = 1.0 - np.abs(d2)
    if s1 is None and spatial_shape is not None:
        s2 = np.asarray(spatial_shape, dtype=np.int64)
        i2 = np.where(i2 < 0, s2, i2)

    frac = f2.prod(axis=-1)
    if not grad:
        return i2, frac
    sign = np.sign(-d2)
    fg = np.empty_like(d2)
    for i in range(dim):
        others = [(i + 1 + j) % dim for j in range(dim - 1)]
        fg[..., i] = sign[..., i] * np.prod(f2[..., others], axis=-1)
    return i2, frac, fg


def _in_bounds(idx, spatial_shape):
    s = np.asarray(spatial_shape, dtype=np.int64)
    return np.all((idx >= 0) & (idx < s), axis=-1)


def _flat(idx, spatial_shape):
    s = np.asarray(spatial_shape, dtype=np.int64)
    safe = np.clip(idx, 0, s - 1)
    return (safe[..., 0] * s[1] + safe[..., 1]) * s[2] + safe[..., 2]


def _accumulate(out, flat, contrib):
    """out[flat[i]] += contrib[i] for every i (repeated indices accumulate), per channel.

    ``numpy.bincount`` sums its weights in input order, as ``numpy.add.at`` does, at a
    fraction of the cost."""
    n = out.shape[0]
    if out.ndim == 1:
        out += np.bincount(flat, weights=contrib, minlength=n)
        return
    rest = out.reshape(n, -1)
    cc = contrib.reshape(contrib.shape[0], -1)
    for c in range(rest.shape[1]):
        rest[:, c] += np.bincount(flat, weights=cc[:, c], minlength=n)


def _broadcast_val(val, n, valid):
    v = np.asarray(val, dtype=np.float64)
    scalar = v.ndim == 0
    if scalar:
        v = np.full((n,), float(v))
    if valid is not None:
        m = np.asarray(valid, dtype=bool).reshape((n,) + (1,) * (v.ndim - 1))
        v = np.where(m, v, 0.0)
    return v, scalar


def scatter(pmid, disp, val, mesh, cell_size, mesh_shape, offset=0.0, mesh_cell_size=None,
            valid=None):
    """M_g = M_in,g + sum_p v_p W_gp (scatter.py:278-338, pallas.py:268-303).

    ``mesh`` is (X,Y,Z) or (X,Y,Z,C); ``val`` is scalar, (N,) or (N,C).  Invalid
    particles contribute nothing; corners outside ``mesh.shape[:3]`` are dropped.
    """
    mesh = np.array(mesh, dtype=np.float64, copy=True)
    spatial = mesh.shape[:3]
    chan = mesh.shape[3:]
    n = np.asarray(pmid).shape[0]
    v, _ = _broadcast_val(val, n, valid)
    if v.shape[1:] != chan:
        raise ValueError(f"channel shape mismatch: {chan} != {v.shape[1:]}")
    idx, frac = enmesh(pmid, disp, cell_size, mesh_shape, offset, mesh_cell_size, spatial)
    ok = _in_bounds(idx, spatial)
    flat = _flat(idx, spatial)
    w = np.where(ok, frac, 0.0)
    out = mesh.reshape((-1,) + chan)
    contrib = v[:, None, ...] * w.reshape(w.shape + (1,) * len(chan))
    _accumulate(out, flat.reshape(-1), contrib.reshape((-1,) + chan))
    return out.reshape(mesh.shape)


def gather(pmid, disp, mesh, cell_size, mesh_shape, offset=0.0, mesh_cell_size=None, valid=None):
    """f_p = sum_g W_gp F_g (gather.py:452-472); invalid particles read 0 (pallas.py:305-318)."""
    mesh = np.asarray(mesh, dtype=np.float64)
    spatial = mesh.shape[:3]
    chan = mesh.shape[3:]
    idx, frac = enmesh(pmid, disp, cell_size, mesh_shape, offset, mesh_cell_size, spatial)
    ok = _in_bounds(idx, spatial)
    flat = _flat(idx, spatial)
    vals = mesh.reshape((-1,) + chan)[flat]  # (N,8,C...)
    w = np.where(ok, frac, 0.0)
    out = (vals * w.reshape(w.shape + (1,) * len(chan))).sum(axis=1)
    if valid is not None:
        m = np.asarray(valid, dtype=bool).reshape((-1,) + (1,) * len(chan))
        out = np.where(m, out, 0.0)
    return out


def gather_many(pmid, disp, meshes, cell_size, mesh_shape, valid=None):
    """``gather`` of several scalar meshes at the same particles, sharing one ``enmesh``;
    returns (N, len(meshes)).  Equal to stacking ``gather`` of each mesh (test-pinned)."""
    spatial = np.asarray(meshes[0]).shape[:3]
    idx, frac = enmesh(pmid, disp, cell_size, mesh_shape, 0.0, None, spatial)
    ok = _in_bounds(idx, spatial)
    flat = _flat(idx, spatial)
    w = np.where(ok, frac, 0.0)
    out = np.stack([(np.asarray(m, dtype=np.float64).reshape(-1)[flat] * w).sum(axis=1) for m in meshes], axis=-1)
    if valid is not None:
        out = np.where(np.asarray(valid, dtype=bool)[:, None], out, 0.0)
    return out


def scatter_adj(pmid, disp, val, mesh_cot, cell_size, mesh_shape, offset=0.0, mesh_cell_size=None,
                valid=None):
    """Adjoint of scatter (scatter.py:341-375, 386-437; pallas.py:358-409).

    Returns ``(disp_cot (N,3), val_cot)``; ``val_cot`` is a scalar when ``val`` is
    scalar, else shaped like ``val``.  The mesh cotangent passes through unchanged.
    """
    mesh_cot = np.asarray(mesh_cot, dtype=np.float64)
    spatial = mesh_cot.shape[:3]
    chan = mesh_cot.shape[3:]
    n = np.asarray(pmid).shape[0]
    v, scalar = _broadcast_val(val, n, valid)
    idx, frac, fg = enmesh(pmid, disp, cell_size, mesh_shape, offset, mesh_cell_size, spatial, True)
    ok = _in_bounds(idx, spatial)
    flat = _flat(idx, spatial)
    corner = mesh_cot.reshape((-1,) + chan)[flat]  # (N,8,C...)
    corner = np.where(ok.reshape(ok.shape + (1,) * len(chan)), corner, 0.0)
    if valid is not None:
        vm = np.asarray(valid, dtype=bool).reshape((n, 1) + (1,) * len(chan))
        corner = np.where(vm, corner, 0.0)
    inner = corner * v[:, None, ...]
    if len(chan):
        inner = inner.sum(axis=tuple(range(2, inner.ndim)))
    scale = float(mesh_cell_size) if mesh_cell_size is not None else float(cell_size)
    disp_cot = (inner[..., None] * fg).sum(axis=1) / scale
    val_cot = (corner * frac.reshape(frac.shape + (1,) * len(chan))).sum(axis=1)
    if scalar:
        val_cot = val_cot.sum()
    # scatter.py:424-435: no displacement gradient where the value is zero / invalid
    nz = np.asarray(v != 0)
    if nz.ndim > 1:
        nz = nz.any(axis=tuple(range(1, nz.ndim)))
    disp_cot = np.where(nz[:, None], disp_cot, 0.0)
    return disp_cot, val_cot


def gather_adj(pmid, disp, mesh, val_cot, cell_size, mesh_shape, offset=0.0, mesh_cell_size=None,
               valid=None):
    """Adjoint of gather (gather.py:475-508, 517-549; pallas.py:321-355).

    Returns ``(disp_cot (N,3), mesh_cot)`` with ``mesh_cot`` shaped like ``mesh``.
    Invalid particles neither read nor write (the kernel-seam contract).
    """
    mesh = np.asarray(mesh, dtype=np.float64)
    spatial = mesh.shape[:3]
    chan = mesh.shape[3:]
    n = np.asarray(pmid).shape[0]
    vc = np.asarray(val_cot, dtype=np.float64).reshape((n,) + chan)
    if valid is not None:
        vc = np.where(np.asarray(valid, dtype=bool).reshape((n,) + (1,) * len(chan)), vc, 0.0)
    idx, frac, fg = enmesh(pmid, disp, cell_size, mesh_shape, offset, mesh_cell_size, spatial, True)
    ok = _in_bounds(idx, spatial)
    flat = _flat(idx, spatial)
    corner = mesh.reshape((-1,) + chan)[flat]
    corner = np.where(ok.reshape(ok.shape + (1,) * len(chan)), corner, 0.0)
    inner = vc[:, None, ...] * corner
    if len(chan):
        inner = inner.sum(axis=tuple(range(2, inner.ndim)))
    scale = float(mesh_cell_size) if mesh_cell_size is not None else float(cell_size)
    disp_cot = (inner[..., None] * fg).sum(axis=1) / scale
    w = np.where(ok, frac, 0.0)
    mesh_cot = np.zeros(mesh.shape, dtype=np.float64).reshape((-1,) + chan)
    contrib = vc[:, None, ...] * w.reshape(w.shape + (1,) * len(chan))
    _accumulate(mesh_cot, flat.reshape(-1), contrib.reshape((-1,) + chan))
    return disp_cot, mesh_cot.reshape(mesh.shape)
