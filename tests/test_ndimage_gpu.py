"""Fourier-crop downsampling on the GPU against the FP64 restatement, and the
batched FP32 MFMA product it runs on against an FP64 product."""

import numpy as np
import pytest
import torch

from cryojax_amd import ndimage
from oracle import downsample as D

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-5  # relative L2 and max-abs (relative to the largest output) against FP64


def extent(batch, b_stride, nseg, s_stride, rows, ld, k):
    return (batch - 1) * b_stride + (nseg - 1) * s_stride + (rows - 1) * ld + k


@pytest.mark.parametrize("m,n,k,nseg,batch,pad", [
    (128, 256, 256, 1, 3, 0),   # 128 x 128 tiles, float4 loads
    (64, 64, 64, 2, 5, 0),      # 64 x 64 tiles, two segments
    (130, 70, 37, 2, 2, 3),     # ragged edges, K tail, scalar loads
    (1, 7, 5, 1, 4, 1),
    (200, 33, 48, 3, 1, 4),
])
def test_bgemm_matches_fp64(m, n, k, nseg, batch, pad, rng):
    lda, ldb, ldc = k + pad, k + 2 * pad, n + pad
    a_seg, b_seg = m * lda + pad, n * ldb
    a_batch, b_batch, c_batch = nseg * a_seg + 4, nseg * b_seg, m * ldc
    a = rng.standard_normal(extent(batch, a_batch, nseg, a_seg, m, lda, k)).astype(np.float32)
    b = rng.standard_normal(extent(batch, b_batch, nseg, b_seg, n, ldb, k)).astype(np.float32)
    c = torch.full((batch * c_batch,), np.nan, dtype=torch.float32, device=DEV)
    print(f"[bgemm] m={m} n={n} k={k} nseg={nseg} batch={batch}", flush=True)
    ndimage.bgemm_nt(torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV), c, m, n, k, lda, ldb, ldc,
                     a_batch, a_seg, b_batch, b_seg, c_batch, nseg, batch)
    got = c.cpu().numpy()
    for i in range(batch):
        want = np.zeros((m, n))
        for s in range(nseg):
            A = np.lib.stride_tricks.as_strided(a[i * a_batch + s * a_seg:].astype(np.float64), (m, k), (8 * lda, 8))
            B = np.lib.stride_tricks.as_strided(b[i * b_batch + s * b_seg:].astype(np.float64), (n, k), (8 * ldb, 8))
            want += A @ B.T
        blk = got[i * c_batch:(i + 1) * c_batch].reshape(m, ldc)
        np.testing.assert_allclose(blk[:, :n], want, rtol=0, atol=2e-5 * np.sqrt(k * nseg) * 4)
        if pad:
            assert np.isnan(blk[:, n:]).all(), "columns beyond n untouched"


def rel_err(got, want):
    d = np.asarray(got, dtype=np.complex128) - want
    return np.linalg.norm(d) / np.linalg.norm(want), np.abs(d).max() / np.abs(want).max()


# square 128^2 / 256^2 images to 32^2 / 64^2 take the fused FFT kernel (both output spaces)
CASES = [
    ((64, 64), (32, 32)),
    ((256, 256), (64, 64)),
    ((256, 256), (32, 32)),
    ((128, 128), (64, 64)),
    ((128, 128), (32, 32)),
    ((256, 256), (128, 128)),
    ((33, 40), (11, 14)),
    ((30, 30), (10, 10)),
    ((9, 9), (9, 9)),
    ((20, 30), (1, 7)),
    ((100, 72), (50, 36)),
    ((5, 1), (3, 1)),
    ((4, 2), (2, 2)),
    ((7, 3), (7, 3)),
]


@pytest.mark.parametrize("shape,out", CASES)
@pytest.mark.parametrize("real,pm", [(True, False), (True, True), (False, False), (False, True)])
def test_downsample_matches_oracle(shape, out, real, pm, rng):
    n = 3
    x = rng.standard_normal((n, *shape)).astype(np.float32)
    print(f"[downsample] {shape} -> {out} real={real} preserve_mean={pm}", flush=True)
    got = ndimage.fourier_crop_downsample_images(torch.from_numpy(x).to(DEV), out, real, pm).cpu().numpy()
    for i in range(n):
        want = D.fourier_crop_downsample_to_shape(x[i].astype(np.float64), out, real, pm)
        assert got[i].shape == want.shape
        l2, mx = rel_err(got[i], want)
        assert l2 < TOL and mx < TOL, (i, l2, mx)


def test_identity_and_constant():
    x = torch.randn(4, 48, 48, device=DEV)
    same = ndimage.fourier_crop_downsample_images(x, (48, 48))
    torch.testing.assert_close(same, x, rtol=0, atol=2e-5)
    c = torch.full((2, 64, 64), 0.25, device=DEV)
    y = ndimage.fourier_crop_downsample_images(c, (16, 16))
    torch.testing.assert_close(y, torch.full_like(y, 0.25 * 16), rtol=1e-5, atol=0)
    ym = ndimage.fourier_crop_downsample_images(c, (16, 16), preserve_mean=True)
    torch.testing.assert_close(ym, torch.full_like(ym, 0.25), rtol=1e-5, atol=0)


def test_reference_api_single_image(rng):
    x = rng.standard_normal((96, 96)).astype(np.float32)
    got = ndimage.fourier_crop_downsample(torch.from_numpy(x).to(DEV), 3.0).cpu().numpy()
    want = D.fourier_crop_downsample(x.astype(np.float64), 3.0)
    assert got.shape == (32, 32)
    assert rel_err(got, want)[0] < TOL
    with pytest.raises(ValueError):
        ndimage.fourier_crop_downsample(torch.from_numpy(x).to(DEV), 0.5)


@pytest.mark.parametrize("h,o", [(256, 64), (128, 32)])
@pytest.mark.parametrize("pm,real", [(False, True), (True, True), (False, False), (True, False)])
def test_fft_path_large_batch(h, o, pm, real):
    """600 images through the persistent FFT kernel (more images than compute units, so workgroups
    loop and prefetch the next image), in both output spaces: the Fourier output's last stage
    reuses a window the next image's first stage overwrites."""
    n = 600
    x = torch.randn(n, h, h, device=DEV, generator=torch.Generator(device=DEV).manual_seed(4))
    print(f"[downsample] FFT path 600 x {h}^2 -> {o}^2 preserve_mean={pm} real={real}", flush=True)
    y = ndimage.fourier_crop_downsample_images(x, (o, o), outputs_real_space=real, preserve_mean=pm).cpu().numpy()
    for i in (0, 255, 256, 511, 599):
        want = D.fourier_crop_downsample_to_shape(x[i].double().cpu().numpy(), (o, o), real, pm)
        l2, mx = rel_err(y[i], want)
        assert l2 < TOL and mx < TOL, (i, l2, mx)


def test_misaligned_input_takes_the_contractions():
    """A contiguous view 4 bytes into its storage cannot feed the FFT kernel's 16-byte loads; the
    call falls back to the contractions and still matches the oracle."""
    n, h, o = 3, 256, 64
    base = torch.randn(n * h * h + 1, device=DEV, generator=torch.Generator(device=DEV).manual_seed(6))
    x = base[1:].view(n, h, h)
    assert x.is_contiguous() and x.data_ptr() % 16 != 0
    print("[downsample] misaligned 3 x 256^2 -> 64^2", flush=True)
    y = ndimage.fourier_crop_downsample_images(x, (o, o)).cpu().numpy()
    for i in range(n):
        want = D.fourier_crop_downsample_to_shape(x[i].double().cpu().numpy(), (o, o))
        l2, mx = rel_err(y[i], want)
        assert l2 < TOL and mx < TOL, (i, l2, mx)


def test_chunked_batch(rng):
    """2048^2 -> 1024^2 holds 9 images per workspace chunk: 11 images take two chunks."""
    n = 11
    x = torch.randn(n, 2048, 2048, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3))
    print("[downsample] chunked 11 x 2048^2 -> 1024^2", flush=True)
    y = ndimage.fourier_crop_downsample_images(x, (1024, 1024)).cpu().numpy()
    for i in (0, 8, 9, 10):
        want = D.fourier_crop_downsample_to_shape(x[i].double().cpu().numpy(), (1024, 1024))
        l2, mx = rel_err(y[i], want)
        assert l2 < TOL and mx < TOL, (i, l2, mx)


@pytest.mark.parametrize("shape,out", [((32, 32, 32), (16, 16, 16)), ((20, 24, 18), (9, 12, 7)), ((16, 16, 16), (16, 16, 16))])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("real,pm", [(True, False), (False, False), (True, True)])
def test_volume_downsampling_matches_oracle(shape, out, dtype, real, pm):
    """The 3D branch (_downsample.py:135-147): FP64 centred transforms on device, the Hartley crop both ways."""
    x = torch.randn(shape, dtype=dtype, device=DEV, generator=torch.Generator(device=DEV).manual_seed(9))
    print(f"[downsample] volume {shape} -> {out} {dtype} real={real} preserve_mean={pm}", flush=True)
    y = ndimage.fourier_crop_downsample_to_shape(x, out, outputs_real_space=real, preserve_mean=pm)
    want = D.fourier_crop_downsample_to_shape(x.double().cpu().numpy(), out, real, pm)
    assert tuple(y.shape) == want.shape
    single = dtype == torch.float32
    assert y.dtype == ((torch.float32 if single else torch.float64) if real else
                       (torch.complex64 if single else torch.complex128))
    l2, mx = rel_err(y.cpu().numpy(), want)
    assert l2 < TOL and mx < TOL, (l2, mx)


@pytest.mark.parametrize("shape,out", [((24, 20), (12, 9)), ((64, 64), (32, 32)), ((16, 18, 20), (8, 9, 7))])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("real,pm", [(True, False), (False, False), (True, True)])
def test_complex_downsampling_matches_oracle(shape, out, dtype, real, pm):
    """The complex branch (_downsample.py:225-238): centred Fourier crop, then the centred inverse."""
    g = torch.Generator(device=DEV).manual_seed(10)
    x = torch.randn(shape, dtype=dtype, device=DEV, generator=g)
    print(f"[downsample] complex {shape} -> {out} {dtype} real={real} preserve_mean={pm}", flush=True)
    y = ndimage.fourier_crop_downsample_to_shape(x, out, outputs_real_space=real, preserve_mean=pm)
    want = D.fourier_crop_downsample_to_shape(x.cpu().numpy().astype(np.complex128), out, real, pm)
    assert tuple(y.shape) == want.shape and y.dtype == dtype
    l2, mx = rel_err(y.cpu().numpy(), want)
    assert l2 < TOL and mx < TOL, (l2, mx)


def test_fourier_crop_downsample_factor_on_volumes():
    x = torch.randn((30, 30, 30), dtype=torch.float32, device=DEV, generator=torch.Generator(device=DEV).manual_seed(1))
    y = ndimage.fourier_crop_downsample(x, 2.5)
    assert tuple(y.shape) == (12, 12, 12)
    want = D.fourier_crop_downsample(x.double().cpu().numpy(), 2.5)
    l2, mx = rel_err(y.cpu().numpy(), want)
    assert l2 < TOL and mx < TOL
