"""FP64 volume transform (from_real_voxel_grid) and observed-image spectra on
the GPU against the FP64 restatement."""

import numpy as np
import pytest
import torch

from cryojax_amd import ops, search
from cryojax_amd import simulator as cxs
from oracle import imaging as I
from oracle import search as S

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def close(got, want, rel):
    got = np.asarray(got)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < rel, err


# unpadded power-of-two volumes with w >= 32 take the centred FFT passes, the others the contractions
@pytest.mark.parametrize("shape,padded", [
    ((16, 16, 16), (16, 16, 16)),
    ((32, 32, 32), (32, 32, 32)),
    ((16, 64, 32), (16, 64, 32)),
    ((64, 16, 128), (64, 16, 128)),
    ((128, 128, 128), (128, 128, 128)),
    ((15, 17, 20), (20, 24, 30)),
    ((9, 9, 9), (13, 13, 13)),
    ((64, 64, 64), (96, 96, 96)),
    ((1, 3, 2), (1, 4, 2)),
])
@pytest.mark.parametrize("f32", [False, True])
def test_voxel_grid_matches_oracle(shape, padded, f32, rng):
    x = rng.standard_normal(shape)
    if f32:
        x = x.astype(np.float32)
    print(f"[fft] volume {shape} -> {padded} f32={f32}", flush=True)
    g128, g64 = ops.fourier_voxel_grid(torch.from_numpy(x).to(DEV), padded)
    want = I.fourier_grid_from_real(x.astype(np.float64), padded_shape=padded)
    close(g128.cpu().numpy(), want, 1e-12)
    close(g64.cpu().numpy(), want, 2e-7)


@pytest.mark.parametrize("shape", [(32, 32, 32), (12, 12, 12)])
def test_voxel_grid_single_outputs(shape, rng):
    """Only one output requested: the FFT path then keeps its complex128 passes in the workspace."""
    x = rng.standard_normal(shape)
    want = I.fourier_grid_from_real(x)
    xd = torch.from_numpy(x).to(DEV)
    g128, g64 = ops.fourier_voxel_grid(xd, complex64=False)
    assert g64 is None
    close(g128.cpu().numpy(), want, 1e-12)
    g128, g64 = ops.fourier_voxel_grid(xd, complex128=False)
    assert g128 is None
    close(g64.cpu().numpy(), want, 2e-7)


def test_volume_api_256(rng):
    x = rng.standard_normal((256, 256, 256)).astype(np.float32)
    print("[fft] from_real_voxel_grid 256^3", flush=True)
    vol = cxs.FourierVoxelGridVolume.from_real_voxel_grid(x)
    want = I.fourier_grid_from_real(x.astype(np.float64))
    close(vol.fourier_voxel_grid_f64.cpu().numpy(), want, 1e-12)
    close(vol.fourier_voxel_grid.cpu().numpy(), want, 2e-7)
    assert vol.fourier_voxel_grid.dtype == torch.complex64
    r = vol.rotate_to_pose(cxs.EulerAnglePose(phi_angle=10.0, theta_angle=20.0, psi_angle=30.0))
    assert r.fourier_voxel_grid_f64 is vol.fourier_voxel_grid_f64


def test_volume_api_padding(rng):
    x = rng.standard_normal((20, 20, 20))
    vol = cxs.FourierVoxelGridVolume.from_real_voxel_grid(x, pad_scale=1.5)
    assert vol.shape == (30, 30, 30)
    close(vol.fourier_voxel_grid_f64.cpu().numpy(), I.fourier_grid_from_real(x, pad_scale=1.5), 1e-12)
    with pytest.raises(NotImplementedError):
        cxs.FourierVoxelGridVolume.from_real_voxel_grid(x, pad_scale=1.5, pad_mode="edge")
    with pytest.raises(ValueError):
        cxs.FourierVoxelGridVolume.from_real_voxel_grid(x, pad_scale=0.5)


# power-of-two sides 16..512 (w >= 32) take the Stockham FFT kernels, the others the DFT contractions
@pytest.mark.parametrize("n,h,w", [(3, 64, 64), (2, 16, 32), (2, 512, 32), (1, 32, 512), (3, 256, 128), (2, 128, 256),
                                   (5, 33, 40), (2, 7, 9), (4, 1, 6), (2, 128, 1), (3, 64, 16), (2, 1024, 64)])
@pytest.mark.parametrize("f32", [False, True])
def test_rfft2_matches_oracle(n, h, w, f32, rng):
    y = rng.standard_normal((n, h, w))
    if f32:
        y = y.astype(np.float32)
    print(f"[fft] rfft2 {n}x{h}x{w} f32={f32}", flush=True)
    got = search.observed_spectrum(torch.from_numpy(y).to(DEV))
    assert got.dtype == torch.complex128 and tuple(got.shape) == (n, h, w // 2 + 1)
    close(got.cpu().numpy(), S.observed_spectrum(y.astype(np.float64)), 1e-12)


def test_rfft2_chunked(rng):
    """384^2 images (DFT contractions): 452 per workspace chunk, so 460 images take two chunks."""
    y = torch.randn(460, 384, 384, device=DEV, generator=torch.Generator(device=DEV).manual_seed(2))
    print("[fft] rfft2 chunked 460 x 384^2", flush=True)
    got = search.observed_spectrum(y)
    idx = [0, 451, 452, 459]
    want = S.observed_spectrum(y[idx].double().cpu().numpy())
    close(got[idx].cpu().numpy(), want, 1e-12)


def test_rfft2_fft_path_large_batch(rng):
    """1030 x 256^2 through the FFT kernels (row pass, then the in-place column pass)."""
    y = torch.randn(1030, 256, 256, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3))
    print("[fft] rfft2 FFT path 1030 x 256^2", flush=True)
    got = search.observed_spectrum(y)
    idx = [0, 517, 1029]
    want = S.observed_spectrum(y[idx].double().cpu().numpy())
    close(got[idx].cpu().numpy(), want, 1e-12)
    # linearity over the whole batch, in float64 (a float32 sum would round at 6e-8):
    # rfft2(y + 2 y') = rfft2(y) + 2 rfft2(y'), y' = y rolled by one image
    y = y.double()
    got = search.observed_spectrum(y)
    lin = search.observed_spectrum(y + 2 * torch.roll(y, 1, dims=0))
    ref = got + 2 * torch.roll(got, 1, dims=0)
    err = ((lin - ref).abs().amax() / ref.abs().amax()).item()
    assert err < 1e-12, err


def test_misaligned_buffers_rejected_or_realigned():
    """The FFT paths need 16-byte aligned buffers. Real device allocations offset by 8 bytes:
    the C ABI refuses them with an error code before any launch, and the Python wrappers copy a
    misaligned input to an aligned buffer and give the aligned result."""
    import ctypes

    from cryojax_amd import _lib

    lib = _lib.load()
    dev = "cuda:0"
    n_img, h = 2, 64
    base = torch.randn(n_img * h * h + 2, dtype=torch.float32, device=dev)
    img = base[2:].view(n_img, h, h)  # data_ptr 8 bytes past an aligned allocation
    assert img.data_ptr() % 16 == 8
    out = torch.empty((n_img, h, h // 2 + 1), dtype=torch.complex128, device=dev)
    ws = torch.empty(max(lib.cryomi_rfft2_f64_workspace_bytes(n_img, h, h), 1), dtype=torch.uint8, device=dev)
    st = lib.cryomi_rfft2_f64(ctypes.c_void_p(img.data_ptr()), 0, n_img, h, h, ctypes.c_void_p(out.data_ptr()),
                              ctypes.c_void_p(ws.data_ptr()), ws.numel(), None)
    assert st == _lib.CRYOMI_ERR_INVALID_ARGUMENT
    got = ops.rfft2_f64(img)
    want = ops.rfft2_f64(img.clone())
    assert torch.equal(got, want)
    vb = torch.randn(32 ** 3 + 2, dtype=torch.float32, device=dev)
    vol = vb[2:].view(32, 32, 32)
    g128, g64 = ops.fourier_voxel_grid(vol)
    r128, r64 = ops.fourier_voxel_grid(vol.clone())
    assert torch.equal(g128, r128) and torch.equal(g64, r64)
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,pad", [(32, 1.0), (24, 1.5)])
def test_from_real_voxel_grid_with_filter(n, pad):
    """FourierVoxelGridVolume.from_real_voxel_grid(filter=...) against the oracle: the filter
    multiplies fftn(x) with its zero frequency in the corner (voxel_volume.py:105-109), so both the
    complex64 and the complex128 grids are the centred transform times fftshift(filter)."""
    from cryojax_amd import transforms as T

    rng = np.random.default_rng(91)
    x = rng.standard_normal((n, n, n))
    p = int(n * pad)
    f = rng.uniform(0.1, 2.0, (p, p, p)) * np.exp(1j * rng.uniform(-1, 1, (p, p, p)))
    vol = cxs.FourierVoxelGridVolume.from_real_voxel_grid(x, pad_scale=pad, filter=T.CustomFilter(f))
    want = I.fourier_grid_from_real(x, pad_scale=pad, fourier_filter=f)
    got64 = vol.fourier_voxel_grid.cpu().numpy()
    got128 = vol._grid_f64.cpu().numpy()
    scale = np.abs(want).max()
    print(f"volume filter n={n} pad={pad}: c128 max err {np.abs(got128 - want).max() / scale:.2e}", flush=True)
    assert np.abs(got128 - want).max() <= 1e-12 * scale
    assert np.abs(got64 - want).max() <= 1e-6 * scale
