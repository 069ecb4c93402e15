"""cryomi_simulate_noalloc (allocation-free forward) against cryomi_simulate, through the C ABI.

Bar: bit-exact images (the same arithmetic per image, whatever the persistent grid size), on
consecutive calls that share the device's static scratch, the later ones smaller.
"""

import ctypes

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, ops
from tests.helpers import draw, fourier_grid, make_volume, packed_params

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FULL = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION


def _inputs(n, n_images, seed, bricked):
    vol = make_volume(n, max(40, n), n // 4, seed)
    g = torch.from_numpy(fourier_grid(vol).astype(np.complex64)).to(DEV)
    grid = ops.grid_to_bricks(g) if bricked else g
    prm = ops.params_to_device(packed_params(draw(n_images, seed + 1), 1.0), DEV)
    return grid, prm


def _noalloc(grid, n, prm, flags, out):
    lib = _lib.load()
    with torch.cuda.device(grid.device):
        st = lib.cryomi_simulate_noalloc(ctypes.c_void_p(grid.data_ptr()), n, int(prm.shape[0]),
                                         ctypes.c_void_p(prm.data_ptr()), flags, ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_void_p(ops._stream_ptr(grid.device)))
    _lib.check(st, "cryomi_simulate_noalloc")


# a large call first, then smaller ones on the same static scratch: box 256 (register-FFT kernel,
# grid capped at the scratch's 254 slots), box 128, box 96 (the Stockham kernel), box 64 with
# the fill-only corner rule, and the Fourier-space path (no scratch)
CASES = [(256, 2000, True, FULL), (128, 700, True, FULL), (96, 60, False, FULL),
         (64, 9, False, FULL | _lib.GATHER_FILL), (64, 5, True, FULL | _lib.OUTPUT_FOURIER)]


def test_noalloc_matches_simulate_on_consecutive_calls():
    outs = []
    for n, count, bricked, flags in CASES:
        print(f"[noalloc] box={n} images={count} bricked={bricked} flags={flags:#x}", flush=True)
        grid, prm = _inputs(n, count, 7 + n, bricked)
        f = flags | (_lib.GRID_BRICKED if bricked else 0)
        ref = ops.simulate_images(grid, prm, f, box=n)
        got = torch.full_like(ref, float("nan"))
        _noalloc(grid, n, prm, f, got)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (n, count)
        outs.append(got)
    assert len(outs) == len(CASES)


def test_noalloc_argument_errors():
    lib = _lib.load()
    assert lib.cryomi_simulate_noalloc(None, 42, 1, None, 0, None, None) == _lib.CRYOMI_ERR_UNSUPPORTED
    assert lib.cryomi_simulate_noalloc(None, 64, 1, None, 0x100, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate_noalloc(None, 64, 1, None, 0, None, None) == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    assert lib.cryomi_simulate_noalloc(None, 64, 0, None, 0, None, None) == _lib.CRYOMI_OK
