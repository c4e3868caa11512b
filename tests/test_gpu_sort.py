"""HIP cell sort vs the oracle: bit-exact permutation and bin offsets (integer work)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import sort as osort
from pmpp_amd import _lib
from pmpp_amd._torch import ensure_init, ptr, stream_handle, workspace

pytestmark = pytest.mark.gpu


def hip_sort(pmid, disp, valid, mesh_shape, cell, brick, dev):
    n = pmid.shape[0]
    P = torch.as_tensor(pmid, device=dev)
    D = torch.as_tensor(disp, device=dev)
    V = None if valid is None else torch.as_tensor(valid.astype(np.uint8), device=dev)
    ensure_init(dev)
    nbytes = _lib.size_query("pm_sort_cells_workspace", n, _lib.int3(mesh_shape), _lib.int3(brick))
    ws = workspace(nbytes, dev)
    nb = int(np.prod([-(-m // b) for m, b in zip(mesh_shape, brick)]))
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    offs = torch.empty(nb + 2, dtype=torch.int32, device=dev)
    pm_s = torch.empty((n, 3), dtype=torch.int16, device=dev)
    dp_s = torch.empty((n, 3), dtype=torch.float32, device=dev)
    _lib.call("pm_sort_cells", ptr(P), ptr(D), ptr(V), n, _lib.int3(mesh_shape), ctypes.c_float(cell),
              _lib.int3(brick), ptr(perm), ptr(offs), ptr(pm_s), ptr(dp_s), ptr(ws), nbytes, stream_handle(dev))
    torch.cuda.synchronize()
    perm_np = perm.cpu().numpy()
    # the sorted payload is the input in sorted order, bit for bit
    np.testing.assert_array_equal(pm_s.cpu().numpy(), np.asarray(pmid)[perm_np])
    np.testing.assert_array_equal(dp_s.cpu().numpy().view(np.uint32), np.asarray(disp)[perm_np].view(np.uint32))
    return perm_np, offs.cpu().numpy()


@pytest.mark.parametrize("n,mesh,brick,spread,seed", [
    (1, (4, 4, 4), (1, 1, 1), 1.0, 0),
    (63, (4, 4, 4), (1, 1, 1), 1.0, 1),
    (1000, (8, 8, 8), (2, 2, 2), 3.0, 2),
    (20000, (64, 64, 64), (1, 1, 1), 10.0, 3),     # 262145 keys: three 8-bit passes
    (70000, (32, 16, 8), (4, 2, 8), 40.0, 4),      # several wave segments, one pass
    (131077, (128, 128, 128), (4, 4, 4), 2.0, 5),  # two passes, ragged last group
])
def test_sort_is_bit_exact(device, n, mesh, brick, spread, seed):
    rng = np.random.default_rng(seed)
    cell = 0.5
    pmid = rng.integers(0, min(mesh), (n, 3)).astype(np.int16)
    disp = (rng.normal(size=(n, 3)) * spread * cell).astype(np.float32)
    valid = rng.random(n) > 0.1
    perm, offs = hip_sort(pmid, disp, valid, mesh, cell, brick, device)
    ref_perm, ref_offs = osort.sort_cells(pmid, disp, cell, mesh, brick, valid)
    np.testing.assert_array_equal(perm, ref_perm)
    np.testing.assert_array_equal(offs, ref_offs)


def test_sort_adversarial_inputs(device):
    mesh, brick, cell = (16, 16, 16), (1, 1, 1), 1.0
    # every particle in one cell: order must stay the input order
    n = 10000
    pmid = np.full((n, 3), 5, np.int16)
    disp = np.zeros((n, 3), np.float32)
    perm, offs = hip_sort(pmid, disp, None, mesh, cell, brick, device)
    np.testing.assert_array_equal(perm, np.arange(n))
    # all invalid: identity, every bin empty
    valid = np.zeros(n, bool)
    perm, offs = hip_sort(pmid, disp, valid, mesh, cell, brick, device)
    np.testing.assert_array_equal(perm, np.arange(n))
    assert np.all(offs[:-1] == 0) and offs[-1] == n
    # keys at the extremes of the range, displacements on the cell knots (floor exactness)
    pmid = np.array([[15, 15, 15], [0, 0, 0], [15, 0, 15], [-1, -1, -1], [16, 16, 16]], np.int16)
    disp = np.array([[0, 0, 0], [-0.0, 1e-7, -1e-7], [1, -1, 0.999999], [0, 0, 0], [-1, -1, -1]], np.float32)
    perm, offs = hip_sort(pmid, disp, None, mesh, cell, brick, device)
    ref_perm, ref_offs = osort.sort_cells(pmid, disp, cell, mesh, brick)
    np.testing.assert_array_equal(perm, ref_perm)
    np.testing.assert_array_equal(offs, ref_offs)


def test_sort_empty(device):
    perm, offs = hip_sort(np.zeros((0, 3), np.int16), np.zeros((0, 3), np.float32), None, (4, 4, 4), 1.0,
                          (2, 2, 2), device)
    assert perm.shape == (0,)
    np.testing.assert_array_equal(offs, np.zeros(10, np.int32))


def hip_bucket(pmid, disp, valid, mesh_shape, cell, brick, grid, dev):
    n = pmid.shape[0]
    P = torch.as_tensor(pmid, device=dev)
    D = torch.as_tensor(disp, device=dev)
    V = None if valid is None else torch.as_tensor(valid.astype(np.uint8), device=dev)
    nbytes = _lib.size_query("pm_bucket_cells_workspace", n, _lib.int3(mesh_shape), _lib.int3(brick))
    ws = workspace(nbytes, dev)
    nb = int(np.prod([-(-m // b) for m, b in zip(mesh_shape, brick)]))
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    offs = torch.empty(nb + 2, dtype=torch.int32, device=dev)
    pm_s = torch.empty((n, 3), dtype=torch.int16, device=dev)
    dp_s = torch.empty((n, 3), dtype=torch.float32, device=dev)
    v_s = torch.empty(n, dtype=torch.uint8, device=dev)
    _lib.call("pm_bucket_cells", ptr(P), ptr(D), ptr(V), n, _lib.int3(mesh_shape), ctypes.c_float(cell),
              _lib.int3(brick), None if grid is None else _lib.int3(grid), ptr(perm), ptr(offs), ptr(pm_s),
              ptr(dp_s), ptr(v_s), ptr(ws), nbytes, stream_handle(dev))
    torch.cuda.synchronize()
    return perm.cpu().numpy(), offs.cpu().numpy(), pm_s.cpu().numpy(), dp_s.cpu().numpy(), v_s.cpu().numpy()


@pytest.mark.parametrize("grid,hint,valid_frac", [
    ((8, 8, 64), True, None),     # particle-grid chunks (4 x 8 x 64 blocks)
    ((8, 8, 64), False, 0.9),     # contiguous index chunks, invalid particles
    ((4, 16, 128), True, 0.8),    # particle-grid chunks with invalid particles
    ((5, 7, 9), True, None),      # the hint does not qualify: index chunks
])
def test_force_step_bucketing_bins_every_particle(device, grid, hint, valid_frac):
    # the unstable binning the force step runs (k_bucket), in both chunk modes: perm is a
    # permutation, every bin is contiguous and holds exactly the particles of its key (the
    # stable sort's set, in any order), offsets equal the oracle's, the payload follows perm
    ms = tuple(2 * g for g in grid)
    brick = (4, 4, 8)
    cell = 0.75
    rng = np.random.default_rng(sum(grid))
    n = int(np.prod(grid))
    pm = np.stack(np.meshgrid(*[2 * np.arange(g) for g in grid], indexing="ij"), -1).reshape(-1, 3)
    pmid = pm.astype(np.int16)
    disp = (rng.normal(size=(n, 3)) * 2.5 * cell).astype(np.float32)
    valid = None if valid_frac is None else rng.random(n) < valid_frac
    perm, offs, pm_s, dp_s, v_s = hip_bucket(pmid, disp, valid, ms, cell, brick, grid if hint else None, device)
    ref_perm, ref_offs = osort.sort_cells(pmid, disp, cell, ms, brick, valid)
    np.testing.assert_array_equal(offs, ref_offs)
    assert np.array_equal(np.sort(perm), np.arange(n))
    for b in range(len(offs) - 1):
        lo, hi = offs[b], offs[b + 1]
        if hi > lo:
            assert np.array_equal(np.sort(perm[lo:hi]), np.sort(ref_perm[lo:hi])), b
    np.testing.assert_array_equal(pm_s, pmid[perm])
    np.testing.assert_array_equal(dp_s.view(np.uint32), disp[perm].view(np.uint32))
    np.testing.assert_array_equal(v_s, np.ones(n, np.uint8) if valid is None else valid[perm].astype(np.uint8))
