"""The allocation-free binning entry point (pm_bucket_cells_noalloc) against pm_bucket_cells.

Both are called through the C ABI on the same inputs, twice in a row (the second call on a
smaller mesh and fewer particles, so counters left in the shared static scratch by the first
call would show).  The order inside a bin is unspecified for both, so the comparison is on
what the binning fixes bit for bit: the bin offsets, the set of particles in every bin, and
the payload gathered through perm.
"""

import ctypes

import numpy as np
import pytest
import torch

from pmpp_amd import _lib
from pmpp_amd._torch import ensure_init, ptr, stream_handle, workspace

pytestmark = pytest.mark.gpu


def _inputs(grid, seed, valid_frac):
    rng = np.random.default_rng(seed)
    n = int(np.prod(grid))
    pm = np.stack(np.meshgrid(*[2 * np.arange(g) for g in grid], indexing="ij"), -1).reshape(-1, 3)
    disp = (rng.normal(size=(n, 3)) * 2.0).astype(np.float32)
    valid = None if valid_frac is None else rng.random(n) < valid_frac
    return pm.astype(np.int16), disp, valid


def _bucket(entry, pmid, disp, valid, mesh, cell, brick, grid, dev):
    n = pmid.shape[0]
    P = torch.as_tensor(pmid, device=dev)
    D = torch.as_tensor(disp, device=dev)
    V = None if valid is None else torch.as_tensor(valid.astype(np.uint8), device=dev)
    nb = int(np.prod([-(-m // b) for m, b in zip(mesh, brick)]))
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    offs = torch.full((nb + 2,), -7, dtype=torch.int32, device=dev)
    pm_s = torch.empty((n, 3), dtype=torch.int16, device=dev)
    dp_s = torch.empty((n, 3), dtype=torch.float32, device=dev)
    v_s = torch.empty(n, dtype=torch.uint8, device=dev)
    args = [ptr(P), ptr(D), ptr(V), n, _lib.int3(mesh), ctypes.c_float(cell), _lib.int3(brick),
            None if grid is None else _lib.int3(grid), ptr(perm), ptr(offs), ptr(pm_s), ptr(dp_s), ptr(v_s)]
    if entry == "pm_bucket_cells":
        nbytes = _lib.size_query("pm_bucket_cells_workspace", n, _lib.int3(mesh), _lib.int3(brick))
        ws = workspace(nbytes, dev)
        _lib.call(entry, *args, ptr(ws), nbytes, stream_handle(dev))
    else:
        _lib.call(entry, *args, stream_handle(dev))
    torch.cuda.synchronize()
    return perm.cpu().numpy(), offs.cpu().numpy(), pm_s.cpu().numpy(), dp_s.cpu().numpy(), v_s.cpu().numpy()


def _canonical(perm, offs):
    """perm with every bin's entries in increasing order (the binning fixes only the sets)."""
    out = perm.copy()
    for b in range(len(offs) - 1):
        lo, hi = offs[b], offs[b + 1]
        out[lo:hi] = np.sort(perm[lo:hi])
    return out


def test_noalloc_binning_matches_workspace_binning(device):
    ensure_init(device)
    cases = [
        ((16, 16, 64), (4, 4, 8), True, 0.9, 0),    # particle-grid chunks, invalid particles, 8194 bins
        ((5, 7, 9), (2, 2, 2), False, None, 1),     # smaller mesh and input, index chunks, 80 bins
    ]
    for grid, brick, hint, vf, seed in cases:
        print(f"[noalloc] grid {grid} brick {brick}", flush=True)
        mesh = tuple(2 * g for g in grid)
        pmid, disp, valid = _inputs(grid, seed, vf)
        n = pmid.shape[0]
        ref = _bucket("pm_bucket_cells", pmid, disp, valid, mesh, 0.75, brick, grid if hint else None, device)
        got = _bucket("pm_bucket_cells_noalloc", pmid, disp, valid, mesh, 0.75, brick, grid if hint else None,
                      device)
        perm, offs, pm_s, dp_s, v_s = got
        np.testing.assert_array_equal(offs, ref[1])
        assert offs[-1] == n
        assert np.array_equal(np.sort(perm), np.arange(n))
        np.testing.assert_array_equal(_canonical(perm, offs), _canonical(ref[0], ref[1]))
        np.testing.assert_array_equal(pm_s, pmid[perm])
        np.testing.assert_array_equal(dp_s.view(np.uint32), disp[perm].view(np.uint32))
        np.testing.assert_array_equal(v_s, np.ones(n, np.uint8) if valid is None else valid[perm].astype(np.uint8))


def test_noalloc_binning_rejects_meshes_beyond_its_scratch(device):
    ensure_init(device)
    # 2048 x 2048 x 2048 cells in 1 x 1 x 64 bricks: 134M bins, far beyond the 8 MiB scratch;
    # rejected on the host before anything is enqueued
    offs = torch.empty(8, dtype=torch.int32, device=device)
    with pytest.raises(NotImplementedError, match="static scratch"):
        _lib.call("pm_bucket_cells_noalloc", None, None, None, 0, _lib.int3((2048, 2048, 2048)), ctypes.c_float(1.0),
                  _lib.int3((1, 1, 64)), None, None, ptr(offs), None, None, None, stream_handle(device))
