"""CPU checks of the C ABI: the library loads, exports every declared symbol, rejects bad input."""

import ctypes
import re
from pathlib import Path

import pytest

from pmpp_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "pmpp_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(pm_\w+)\s*\(", text, flags=re.M)


def test_header_declares_what_the_binding_types():
    names = declared_functions()
    assert len(names) == len(set(names))
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_symbol():
    lib = _lib.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pm_abi_version() == _lib.ABI_VERSION


def test_queries_validate_without_a_gpu():
    lib = _lib.load_library()
    out = ctypes.c_size_t(0)
    assert lib.pm_fft_workspace(_lib.int3((8, 8, 8)), ctypes.byref(out)) == _lib.PM_SUCCESS
    assert lib.pm_fft_workspace(_lib.int3((6, 8, 8)), ctypes.byref(out)) == _lib.PM_ERROR_UNSUPPORTED
    assert b"powers of two" in lib.pm_last_error()
    assert lib.pm_sort_cells_workspace(1000, _lib.int3((8, 8, 8)), _lib.int3((2, 2, 2)),
                                       ctypes.byref(out)) == _lib.PM_SUCCESS
    assert out.value > 4 * 1000
    assert lib.pm_sort_cells_workspace(10, _lib.int3((8, 0, 8)), _lib.int3((2, 2, 2)),
                                       ctypes.byref(out)) == _lib.PM_ERROR_INVALID_ARGUMENT
    from pmpp_amd import Configuration
    from pmpp_amd.nbody.gravity import gravity_desc

    d = gravity_desc(Configuration(1.0, (16, 16, 16), mesh_shape=2), 0.3)
    assert lib.pm_gravity_workspace(d, 4096, ctypes.byref(out)) == _lib.PM_SUCCESS
    assert out.value > 3 * 4 * 32 ** 3
    fwd = out.value
    assert lib.pm_gravity_vjp_workspace(d, 4096, ctypes.byref(out)) == _lib.PM_SUCCESS
    assert out.value > fwd + 4 * 4 * 32 ** 3  # plus the forward meshes kept and 3 cotangent meshes
    # the density's place in the forward workspace: 256-byte aligned, the whole mesh inside it
    for conf_, n_ in [(Configuration(1.0, (16, 16, 16), mesh_shape=2), 4096),
                      (Configuration(1.0, (16, 16, 16), mesh_shape=2, brick=(32, 32, 32)), 4096),
                      (Configuration(1000 / 512, (512,) * 3, mesh_shape=2), 512 ** 3)]:
        dd = gravity_desc(conf_, 0.3)
        ws_bytes = _lib.size_query("pm_gravity_workspace", dd, n_)
        off = _lib.size_query("pm_gravity_density_offset", dd, n_)
        assert off % 256 == 0 and off + 4 * conf_.mesh_size <= ws_bytes
    with pytest.raises(NotImplementedError, match="powers of two"):
        _lib.call("pm_gravity_workspace", gravity_desc(Configuration(1.0, (12, 12, 12), mesh_shape=1), 0.3),
                  10, ctypes.byref(out))


def test_null_arguments_are_rejected_before_any_launch():
    lib = _lib.load_library()
    assert lib.pm_gravity(None, None, None, 10, None, None, None, 0, None) == _lib.PM_ERROR_INVALID_ARGUMENT
    assert lib.pm_cic_gather(None, None, None, 5, 1, None, None, None, None) == _lib.PM_ERROR_INVALID_ARGUMENT


def test_transforms_require_pm_init():
    # pm_init is the one entry point that uploads tables and synchronises; before it has run on
    # a device, the transforming entry points refuse with an error instead of doing it implicitly
    import torch

    if torch.cuda.is_available():
        pytest.skip("checked where no device has been initialised")
    lib = _lib.load_library()
    dummy = ctypes.c_void_p(256)
    st = lib.pm_rfft3d(dummy, ctypes.c_void_p(512), _lib.int3((8, 8, 8)), None, 0, None)
    assert st == _lib.PM_ERROR_INVALID_ARGUMENT
    assert b"pm_init" in lib.pm_last_error()
