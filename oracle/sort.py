"""Cell-sort contract: containing-cell keys and a stable counting sort.

The reference's key/order contract (read as text, never imported):

* ``src/pmpp/core/utils.py:292-326`` -- ``pmid_to_idx``: raveled key ((ix*Ny+iy)*Nz+iz) with
  each index wrapped mod N; masked (padding) slots get a sentinel.
* ``src/pmpp/distributed/routing.py:1-15, 112-116`` -- valid particles are packed in
  nondecreasing raveled key, invalid ones after them through the sentinel ``mesh_size``.
* ``cuda/route_kernels.cu:194-225`` -- order within a key follows input order (stable
  block-scan compaction).

Because a single device never re-anchors ``pmid`` after a drift
(``src/pmpp/nbody/integrator.py:225-231``), the key is taken from the *containing* cell
``pmid + floor(disp / cell)`` computed with the CIC base-index arithmetic of
``src/pmpp/cic/pallas.py:227-236`` (offset 0), wrapped mod N.  Cells are grouped into
bricks of ``brick`` cells per axis (``brick = (1,1,1)`` is the raveled cell key itself);
the key is the raveled brick index and invalid particles take the sentinel ``num_bricks``.

Pinned in ``tests/test_oracle_sort.py`` by equality with ``numpy.argsort(kind='stable')``
on the keys, a brute-force insertion sort on small inputs, and the key's agreement with the
CIC base index of ``oracle.cic.enmesh`` (corner 0) on every particle.
"""

from __future__ import annotations

import numpy as np


def containing_cell(pmid, disp, cell_size, mesh_shape, work_dtype=np.float32):
    """Base CIC cell of each particle, wrapped mod N (pallas.py:227-236 with offset 0).

    ``work_dtype=float32`` reproduces the float32 kernel arithmetic bit for bit
    (IEEE division then floor); float64 gives the exact-arithmetic cell.
    """
    pm = np.asarray(pmid, dtype=np.int64)
    d = np.asarray(disp).astype(work_dtype)
    a = work_dtype(cell_size)
    t = d / a
    base = pm + np.floor(t).astype(np.int64)
    return np.mod(base, np.asarray(mesh_shape, dtype=np.int64))


def cell_keys(pmid, disp, cell_size, mesh_shape, brick=(1, 1, 1), valid=None,
              work_dtype=np.float32):
    """Raveled brick key of the containing cell; sentinel = number of bricks if invalid."""
    cell = containing_cell(pmid, disp, cell_size, mesh_shape, work_dtype)
    b = np.asarray(brick, dtype=np.int64)
    nb = -(-np.asarray(mesh_shape, dtype=np.int64) // b)
    c = cell // b
    key = (c[:, 0] * nb[1] + c[:, 1]) * nb[2] + c[:, 2]
    sentinel = int(np.prod(nb))
    if valid is not None:
        key = np.where(np.asarray(valid, dtype=bool), key, sentinel)
    return key.astype(np.int64), sentinel


def sort_cells(pmid, disp, cell_size, mesh_shape, brick=(1, 1, 1), valid=None):
    """Stable counting sort by cell key.

    Returns ``(perm, offsets)``: ``perm[j]`` is the input index of the j-th particle in key
    order (stable), ``offsets`` has ``num_bricks + 2`` entries so that bin b occupies
    ``perm[offsets[b]:offsets[b+1]]`` (the last bin is the invalid sentinel).
    """
    key, sentinel = cell_keys(pmid, disp, cell_size, mesh_shape, brick, valid)
    counts = np.bincount(key, minlength=sentinel + 1)
    offsets = np.zeros(sentinel + 2, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    # A stable counting sort places particle i at offsets[key_i] + #{j < i : key_j == key_i},
    # which is exactly numpy's stable argsort of the keys.
    perm = np.argsort(key, kind="stable").astype(np.int64)
    return perm, offsets


def sort_cells_bruteforce(keys):
    """Insertion sort of (key, index) pairs -- the definition of a stable sort, for tiny inputs."""
    items = []
    for i, k in enumerate(keys):
        j = len(items)
        while j > 0 and items[j - 1][0] > k:
            j -= 1
        items.insert(j, (int(k), i))
    return np.array([i for _, i in items], dtype=np.int64)
