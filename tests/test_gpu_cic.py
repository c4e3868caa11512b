"""HIP CIC kernels vs the float64 oracle (the reference's Pallas-vs-JAX gate,
tests/test_pallas_cic.py:31-154, restated: rtol/atol 2e-6, 3e-5 for explicit cell sizes)."""

import numpy as np
import pytest
import torch

from oracle import cic
from pmpp_amd.cic import hip_gather, hip_gather_bwd, hip_scatter, hip_scatter_bwd

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype=torch.float32):
    return torch.as_tensor(np.asarray(x), device=dev).to(dtype)


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


CASES = [
    # (mesh spatial, global, channels, offset, mesh_cell, particle_cell, n, valid)
    ((4, 4, 4), (4, 4, 4), (), 0.0, None, 1.0, 4, False),
    ((4, 4, 4), (4, 4, 4), (2,), 0.0, None, 1.0, 5, False),
    ((4, 4, 4), (4, 4, 4), (3,), 0.0, None, 1.0, 5, True),
    ((16, 8, 32), (16, 8, 32), (), [0.3, -0.7, 0.1], None, 0.5, 3000, True),
    ((8, 8, 8), (4, 4, 4), (), [1.0, 0.0, 0.0], 0.5, 1.0, 50, False),
    ((6, 4, 4), (8, 4, 4), (), 0.0, None, 1.0, 200, False),  # local mesh smaller than global: drops
]


def _inputs(case, seed):
    ls, gs, chan, off, mcell, pcell, n, use_valid = case
    rng = np.random.default_rng(seed)
    pmid = rng.integers(0, min(gs), (n, 3)).astype(np.int16)
    disp = rng.uniform(-0.45, 0.45, (n, 3)).astype(np.float32) * np.float32(pcell)
    mesh = rng.normal(size=ls + chan).astype(np.float32)
    val = rng.normal(size=(n,) + chan).astype(np.float32)
    valid = (rng.random(n) > 0.3) if use_valid else None
    return pmid, disp, mesh, val, valid


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_forward_and_adjoints_match_oracle(device, ci):
    case = CASES[ci]
    ls, gs, chan, off, mcell, pcell, n, _ = case
    pmid, disp, mesh, val, valid = _inputs(case, 17 + ci)
    tol = dict(rtol=3e-5, atol=3e-5) if mcell is not None else dict(rtol=2e-6, atol=2e-6)
    kw = dict(offset=off, particle_cell_size=pcell, cell_size=mcell, global_shape=gs)
    okw = dict(cell_size=pcell, mesh_shape=gs, offset=off, mesh_cell_size=mcell, valid=valid)
    vm = None if valid is None else _t(valid, device, torch.bool)
    P, D, M, V = _t(pmid, device, torch.int16), _t(disp, device), _t(mesh, device), _t(val, device)

    g = hip_gather(P, D, M, valid_mask=vm, **kw)
    np.testing.assert_allclose(_np(g), cic.gather(pmid, disp, mesh, **okw), **tol)

    base = np.ones(ls + chan, np.float32)
    s = hip_scatter(P, D, V, _t(base, device), valid_mask=vm, **kw)
    np.testing.assert_allclose(_np(s), cic.scatter(pmid, disp, val, base, **okw), **tol)

    vc = np.random.default_rng(5).normal(size=(n,) + chan).astype(np.float32)
    dcot, mcot = hip_gather_bwd(P, D, M, _t(vc, device), valid_mask=vm, **kw)
    dref, mref = cic.gather_adj(pmid, disp, mesh, vc, **okw)
    np.testing.assert_allclose(_np(dcot), dref, **tol)
    np.testing.assert_allclose(_np(mcot), mref, **tol)

    dcot, vcot = hip_scatter_bwd(P, D, V, M, valid_mask=vm, **kw)
    dref, vref = cic.scatter_adj(pmid, disp, val, mesh, **okw)
    np.testing.assert_allclose(_np(dcot), dref, **tol)
    np.testing.assert_allclose(_np(vcot), vref, **tol)

    # scalar value: scatter and the summed value cotangent
    s = hip_scatter(P, D, 1.7, _t(np.zeros(ls), device), valid_mask=vm, offset=off, particle_cell_size=pcell,
                    cell_size=mcell, global_shape=gs) if not chan else None
    if s is not None:
        np.testing.assert_allclose(_np(s), cic.scatter(pmid, disp, 1.7, np.zeros(ls), **okw), **tol)
        m0 = mesh if not chan else mesh[..., 0]
        _, vcs = hip_scatter_bwd(P, D, 1.7, _t(m0, device), valid_mask=vm, **kw)
        _, vref = cic.scatter_adj(pmid, disp, 1.7, m0, **okw)
        assert float(vcs) == pytest.approx(float(vref), rel=2e-5, abs=2e-5)


def test_empty_and_single_particle_edge_cases(device):
    kw = dict(offset=0.0, particle_cell_size=1.0, global_shape=(4, 4, 4))
    P = torch.zeros((0, 3), dtype=torch.int16, device=device)
    D = torch.zeros((0, 3), dtype=torch.float32, device=device)
    M = torch.ones((4, 4, 4), device=device)
    assert hip_gather(P, D, M, **kw).shape == (0,)
    assert torch.equal(hip_scatter(P, D, 1.0, M, **kw), M)
    # a particle exactly on a node puts all its mass there; gradients use sign(0) = 0
    P1 = torch.tensor([[1, 2, 3]], dtype=torch.int16, device=device)
    D1 = torch.zeros((1, 3), device=device)
    s = hip_scatter(P1, D1, 2.0, torch.zeros((4, 4, 4), device=device), **kw)
    assert float(s[1, 2, 3]) == 2.0 and float(s.sum()) == 2.0
    # periodic wrap: a particle at the last node spreads into node 0
    D2 = torch.tensor([[0.5, 0.0, 0.0]], device=device)
    P2 = torch.tensor([[3, 0, 0]], dtype=torch.int16, device=device)
    s = hip_scatter(P2, D2, 1.0, torch.zeros((4, 4, 4), device=device), **kw)
    assert float(s[3, 0, 0]) == pytest.approx(0.5) and float(s[0, 0, 0]) == pytest.approx(0.5)


def test_bad_arguments_raise(device):
    P = torch.zeros((3, 3), dtype=torch.int16, device=device)
    D = torch.zeros((3, 3), dtype=torch.float32, device=device)
    with pytest.raises(RuntimeError, match="float32"):
        hip_gather(P, D, torch.zeros((4, 4, 4), dtype=torch.float64, device=device), offset=0,
                   particle_cell_size=1.0, global_shape=(4, 4, 4))
    with pytest.raises(ValueError, match="valid_mask"):
        hip_gather(P, D, torch.zeros((4, 4, 4), device=device), offset=0, particle_cell_size=1.0,
                   global_shape=(4, 4, 4), valid_mask=torch.ones(2, dtype=torch.bool, device=device))
    with pytest.raises(ValueError, match="channel shape mismatch"):
        hip_scatter(P, D, torch.ones((3, 2), device=device), torch.zeros((4, 4, 4, 3), device=device), offset=0,
                    particle_cell_size=1.0, global_shape=(4, 4, 4))
