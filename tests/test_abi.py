"""CPU checks of the C ABI: the library loads, exports every symbol the headers
declare, and rejects bad requests before touching a device."""

import ctypes
import os
import re

import pytest

from cryojax_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cryomi_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert "cryomi_simulate" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) >= names


def test_abi_version_and_struct_size():
    assert _lib.load().cryomi_abi_version() == 1
    assert _lib.N_PARAM * 8 == 160


@pytest.mark.parametrize("box", [6, 14, 22, 42, 514, 1024, 7])
def test_unsupported_box(box):
    lib = _lib.load()
    st = lib.cryomi_simulate(None, box, 1, None, 3, None, 0, None, None)
    assert st == _lib.CRYOMI_ERR_UNSUPPORTED
    assert "box" in lib.cryomi_last_error().decode()


def test_invalid_arguments():
    lib = _lib.load()
    assert lib.cryomi_simulate(None, 64, -1, None, 3, None, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate(None, 64, 1, None, 0x80, None, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate(None, 64, 4, None, 3, None, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    # zero images is a valid no-op
    assert lib.cryomi_simulate(None, 64, 0, None, 3, None, 0, None, None) == _lib.CRYOMI_OK
    with pytest.raises(ValueError):
        _lib.check(_lib.CRYOMI_ERR_INVALID_ARGUMENT, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.CRYOMI_ERR_UNSUPPORTED, "x")


def test_fourier_output_needs_no_workspace():
    assert _lib.load().cryomi_simulate_workspace_bytes(64, 10, _lib.OUTPUT_FOURIER) == 0


def test_groups_argument_checks():
    import ctypes

    lib = _lib.load()
    G = _lib.VolumeGroupStruct
    arr = (G * 2)(G(None, None, None, 96, 1, 3, 0), G(None, None, None, 64, 1, 3, 0))
    p = ctypes.cast(arr, ctypes.c_void_p)
    assert lib.cryomi_simulate_groups(p, 2, None, 0, None) == _lib.CRYOMI_ERR_UNSUPPORTED
    assert "box" in lib.cryomi_last_error().decode()
    assert lib.cryomi_simulate_groups_workspace_bytes(p, 2) == 0
    assert lib.cryomi_simulate_groups(p, 0, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate_groups(p, 9, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate_groups(None, 1, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    arr[0].box = 128
    arr[0].n_images = 0  # an empty group needs no pointers
    arr[0].flags = 3 | _lib.GRID_BRICKED  # layouts disagree
    assert lib.cryomi_simulate_groups(p, 2, None, 0, None) == _lib.CRYOMI_ERR_UNSUPPORTED
    arr[0].flags = 3 | _lib.OUTPUT_FOURIER
    assert lib.cryomi_simulate_groups(p, 2, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    arr[0].flags = 3
    arr[0].n_images = -1
    assert lib.cryomi_simulate_groups(p, 2, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    arr[0].n_images = 1  # NULL device pointers
    assert lib.cryomi_simulate_groups(p, 2, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    assert lib.cryomi_simulate_groups_graph_create(p, 2, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate_groups_graph_create(p, 2, None, 0, ctypes.byref(h)) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert h.value is None
    assert lib.cryomi_graph_launch(None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    lib.cryomi_graph_destroy(None)


def test_downsample_argument_checks():
    lib = _lib.load()
    assert lib.cryomi_fourier_crop_operator_floats(64, 64, 65, 32, 0) == 0
    assert lib.cryomi_fourier_crop_operator_floats(64, 64, 32, 32, 0x4) == 0
    assert lib.cryomi_fourier_crop_operator_floats(8192, 64, 32, 32, 0) == 0
    assert lib.cryomi_fourier_crop_operator_floats(64, 64, 32, 32, 0) > 0
    assert lib.cryomi_fourier_crop_downsample_workspace_bytes(-1, 64, 64, 32, 32, 0) == 0
    assert lib.cryomi_fourier_crop_operators(64, 64, 0, 32, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_fourier_crop_operators(64, 64, 32, 32, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    args = (64, 64, 32, 32, 0, None, None, 0, None, None)
    assert lib.cryomi_fourier_crop_downsample(None, 0, *args) == _lib.CRYOMI_OK  # no images: no-op
    assert lib.cryomi_fourier_crop_downsample(None, 2, *args) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    b = _lib.BgemmStruct(None, 0, 0, 4, None, 0, 0, 4, None, 0, 4, 4, 4, 8, 1, 1)  # lda < k
    assert lib.cryomi_bgemm_nt(ctypes.addressof(b), None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    b.lda = b.ldb = 8
    assert lib.cryomi_bgemm_nt(ctypes.addressof(b), None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT  # NULL pointers
    b.batch = 0
    assert lib.cryomi_bgemm_nt(ctypes.addressof(b), None) == _lib.CRYOMI_OK


def test_fft_argument_checks():
    lib = _lib.load()
    assert lib.cryomi_fourier_voxel_grid_workspace_bytes(8, 8, 8, 7, 8, 8) == 0  # padded smaller than input
    assert lib.cryomi_fourier_voxel_grid_workspace_bytes(8, 8, 8, 8, 8, 2048) == 0
    assert lib.cryomi_fourier_voxel_grid_workspace_bytes(8, 8, 8, 12, 12, 12) > 0
    assert lib.cryomi_fourier_voxel_grid(None, 0, 8, 8, 8, 8, 8, 8, None, None, None, 0, None) == \
        _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_fourier_voxel_grid(None, 5, 8, 8, 8, 8, 8, 8, None, None, None, 0, None) == \
        _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_rfft2_f64_workspace_bytes(2, 8, 8192) == 0
    assert lib.cryomi_rfft2_f64_workspace_bytes(4096, 256, 256) <= 256  # FFT sizes work in place
    assert lib.cryomi_rfft2_f64(None, 0, 0, 8, 8, None, None, 0, None) == _lib.CRYOMI_OK  # no images
    assert lib.cryomi_rfft2_f64(None, 0, 2, 8, 8, None, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_rfft2_f64(None, 2, 2, 8, 8, None, None, 0, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT


def test_reuse_prepared_refuses_a_workspace_that_would_grow():
    """A REUSE_PREPARED candidate search on the shared workspace must not reallocate it: the prepared
    observations and bricked grid would be lost. The check runs on the host, before any launch."""
    import torch

    from cryojax_amd import search

    n, n_img, n_off = 16, 2, 4
    grid = torch.zeros((n, n, n), dtype=torch.complex128)
    cand = torch.zeros((n_img, 64), dtype=torch.int32)
    with pytest.raises(ValueError, match="REUSE_PREPARED"):
        search.ncc_search_candidates(grid, torch.zeros((n_img, n, n // 2 + 1), dtype=torch.complex128),
                                     torch.zeros((n_img, _lib.N_PARAM), dtype=torch.float64),
                                     torch.zeros((8, 9), dtype=torch.float64), torch.zeros((n_off, 2)),
                                     None, cand, torch.zeros(n_img, dtype=torch.int32),
                                     best=(None, None, None), flags=search.REUSE_PREPARED)
