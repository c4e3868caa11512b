"""HIP real 3D FFTs vs numpy's pocketfft in float64 on the same float32 input."""

import numpy as np
import pytest
import torch

from pmpp_amd.numerics import irfftn, rfftn, spectral_force_multiply
from pmpp_amd.nbody.gravity import gravity_desc
from pmpp_amd import Configuration
from oracle import spectral

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2, 2), (4, 4, 4), (8, 16, 4), (2, 64, 8), (32, 32, 32), (64, 128, 32), (16, 8, 256),
          (4, 2048, 2), (2048, 2, 4), (4, 4, 2048), (128, 64, 512), (1024, 4, 16), (2, 2, 1024),
          (8, 4, 1024), (16, 2, 1024)]


@pytest.mark.parametrize("shape", SHAPES)
def test_rfftn_irfftn_match_numpy(device, shape):
    rng = np.random.default_rng(abs(hash(shape)) % 2**32)
    x = rng.normal(size=shape).astype(np.float32)
    X = rfftn(torch.from_numpy(x).to(device)).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64))
    scale = np.sqrt(np.prod(shape))  # rms of |X| for unit white noise
    err = np.abs(X - ref).max() / scale
    assert err < 2e-6 * np.log2(np.prod(shape)), err
    y = irfftn(torch.from_numpy(ref.astype(np.complex64)).to(device), s=shape).cpu().numpy()
    np.testing.assert_allclose(y, np.fft.irfftn(ref.astype(np.complex64).astype(np.complex128), s=shape,
                                                axes=(0, 1, 2)), atol=3e-6 * np.log2(np.prod(shape)))


def test_irfftn_ignores_imaginary_dc_and_nyquist(device):
    shape = (4, 4, 8)
    rng = np.random.default_rng(1)
    X = (rng.normal(size=(4, 4, 5)) + 1j * rng.normal(size=(4, 4, 5))).astype(np.complex64)
    y = irfftn(torch.from_numpy(X).to(device), s=shape).cpu().numpy()
    np.testing.assert_allclose(y, np.fft.irfftn(X.astype(np.complex128), s=shape, axes=(0, 1, 2)), atol=2e-6)


@pytest.mark.parametrize("ratio", [1, 2])
@pytest.mark.parametrize("axis", [-1, 0, 1, 2])
def test_spectral_force_multiply_matches_oracle(device, ratio, axis):
    conf = Configuration(1.0, (8, 8, 8), mesh_shape=ratio)
    ms = conf.mesh_shape
    rng = np.random.default_rng(3)
    X = (rng.normal(size=(ms[0], ms[1], ms[2] // 2 + 1)) + 1j * rng.normal(size=(ms[0], ms[1], ms[2] // 2 + 1)))
    X = X.astype(np.complex64)
    masks = spectral.particle_nyquist_masks(ms, conf.ptcl_grid_shape, conf.ptcl_spacing, conf.cell_size)
    G = spectral.green_multiplier(ms, conf.cell_size, masks)
    if axis >= 0:
        nk = spectral.neg_k(ms, conf.cell_size)[axis]
        G = G * (1j * nk.reshape([-1 if a == axis else 1 for a in range(3)]))
    for adjoint in (False, True):
        got = spectral_force_multiply(torch.from_numpy(X).to(device), gravity_desc(conf, 0.3), axis, adjoint)
        ref = (np.conj(G) if adjoint else G) * X.astype(np.complex128)
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=2e-6, atol=1e-7)
