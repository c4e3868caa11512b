"""Pins the cell-sort oracle (oracle/sort.py)."""

import numpy as np

from oracle import cic, sort


def test_stable_sort_equals_insertion_sort():
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 7, 60)
    perm_ref = sort.sort_cells_bruteforce(keys)
    np.testing.assert_array_equal(np.argsort(keys, kind="stable"), perm_ref)


def test_keys_follow_cic_base_cell_and_sentinel():
    rng = np.random.default_rng(1)
    ms = (8, 8, 8)
    pmid = rng.integers(-2, 10, (300, 3))
    disp = rng.uniform(-3, 3, (300, 3)).astype(np.float32)
    idx, _ = cic.enmesh(pmid, disp.astype(np.float64), 0.5, ms)
    cell = sort.containing_cell(pmid, disp, 0.5, ms)
    # corner 0 of enmesh is the containing cell (float64 vs float32 floor agree away from knots)
    agree = np.all(idx[:, 0, :] == cell, axis=1)
    assert agree.mean() > 0.99
    valid = rng.random(300) > 0.2
    key, sentinel = sort.cell_keys(pmid, disp, 0.5, ms, (2, 4, 1), valid)
    assert sentinel == 4 * 2 * 8
    assert np.all(key[~valid] == sentinel)
    c = cell[valid]
    np.testing.assert_array_equal(key[valid], ((c[:, 0] // 2) * 2 + c[:, 1] // 4) * 8 + c[:, 2])
    perm, offsets = sort.sort_cells(pmid, disp, 0.5, ms, (2, 4, 1), valid)
    np.testing.assert_array_equal(perm, sort.sort_cells_bruteforce(key))
    assert offsets[-1] == 300 and offsets[-2] == valid.sum()
    sk = key[perm]
    for b in range(sentinel + 1):
        assert np.all(sk[offsets[b]:offsets[b + 1]] == b)
