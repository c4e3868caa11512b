"""The HIP library against the brute-force golden fixtures of ``tests/golden/``.

Tolerances: float32 kernels vs float64 brute force -- CIC 2e-6 (3e-5 with an explicit mesh
cell, as the reference's Pallas gate, tests/test_pallas_cic.py:45,79), FFT 2e-6 x log2(N)
of the spectrum's rms, force field 1e-5 relative; the sort is integer work and bit-exact.
"""

import ctypes
import os

import numpy as np
import pytest
import torch

from pmpp_amd import Configuration, Particles, _lib, gravity
from pmpp_amd._torch import ensure_init, ptr, stream_handle, workspace
from pmpp_amd.cic import hip_gather, hip_scatter
from pmpp_amd.numerics import rfftn

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def test_cic_matches_golden(device):
    g = load("cic")
    P = torch.as_tensor(g["pmid"], device=device)
    D = torch.as_tensor(g["disp"], device=device)
    V = torch.as_tensor(g["val"], device=device)
    kw = dict(offset=0.0, particle_cell_size=1.0, cell_size=None, global_shape=(4, 4, 4))
    for c in range(2):  # the two channels of the fixture mesh, one at a time and together
        m = torch.as_tensor(np.ascontiguousarray(g["mesh"][..., c]), device=device)
        np.testing.assert_allclose(_np(hip_gather(P, D, m, **kw)), g["gathered"][:, c], rtol=2e-6, atol=2e-6)
    both = hip_gather(P, D, torch.as_tensor(g["mesh"], device=device), **kw)
    np.testing.assert_allclose(_np(both), g["gathered"], rtol=2e-6, atol=2e-6)
    s = hip_scatter(P, D, V, torch.zeros((4, 4, 4), device=device), **kw)
    np.testing.assert_allclose(_np(s), g["scattered"], rtol=2e-6, atol=2e-6)

    kw8 = dict(offset=g["offset8"], particle_cell_size=1.0, cell_size=0.5, global_shape=(4, 4, 4))
    m8 = torch.as_tensor(g["mesh8"], device=device)
    np.testing.assert_allclose(_np(hip_gather(P, D, m8, **kw8)), g["gathered8"], rtol=3e-5, atol=3e-5)
    s8 = hip_scatter(P, D, V, torch.zeros((8, 8, 8), device=device), **kw8)
    np.testing.assert_allclose(_np(s8), g["scattered8"], rtol=3e-5, atol=3e-5)


@pytest.mark.parametrize("with_valid", [True, False])
def test_sort_matches_golden(device, with_valid):
    g = load("sort")
    ms, brick, n = [int(v) for v in g["mesh_shape"]], [int(v) for v in g["brick"]], len(g["pmid"])
    P = torch.as_tensor(g["pmid"], device=device)
    D = torch.as_tensor(g["disp"], device=device)
    V = torch.as_tensor(g["valid"].astype(np.uint8), device=device) if with_valid else None
    ensure_init(device)
    nbytes = _lib.size_query("pm_sort_cells_workspace", n, _lib.int3(ms), _lib.int3(brick))
    ws = workspace(nbytes, device)
    perm = torch.empty(n, dtype=torch.int32, device=device)
    offs = torch.empty(len(g["bins"]), dtype=torch.int32, device=device)
    _lib.call("pm_sort_cells", ptr(P), ptr(D), ptr(V), n, _lib.int3(ms), ctypes.c_float(float(g["cell"])),
              _lib.int3(brick), ptr(perm), ptr(offs), None, None, ptr(ws), nbytes, stream_handle(device))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(perm.cpu().numpy(), g["perm"] if with_valid else g["perm_all"])
    if with_valid:
        np.testing.assert_array_equal(offs.cpu().numpy(), g["bins"])


def test_fft_matches_golden(device):
    g = load("fft")
    X = rfftn(torch.as_tensor(g["x"], device=device)).cpu().numpy()
    scale = np.sqrt(g["x"].size)
    assert np.abs(X - g["X"]).max() / scale < 2e-6 * np.log2(g["x"].size)


def test_gravity_matches_golden(device):
    g = load("gravity")
    ng, nm, sp = int(g["ptcl_grid"]), int(g["mesh"]), float(g["ptcl_spacing"])
    conf = Configuration(sp, (ng,) * 3, mesh_shape=nm // ng)
    ptcl = Particles(conf, torch.as_tensor(g["pmid"], device=device), torch.as_tensor(g["disp"], device=device))
    acc = gravity(1.0, ptcl, float(g["omega_m"]), conf).cpu().numpy()
    ref = g["acc"]
    assert np.linalg.norm(acc - ref) / np.linalg.norm(ref) < 1e-5
    np.testing.assert_allclose(acc, ref, atol=3e-5 * np.abs(ref).max())
