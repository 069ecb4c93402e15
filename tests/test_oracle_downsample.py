"""Pins of the FP64 Fourier-crop downsampling restatement against closed forms
and an independent Fourier-crop formulation (no GPU)."""

import numpy as np
import pytest

from oracle import downsample as D


def centred(n):
    return np.arange(n) - n // 2


@pytest.mark.parametrize("shape", [(8, 8), (9, 9), (12, 7), (16, 10)])
def test_same_shape_is_identity(shape, rng):
    x = rng.standard_normal(shape)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(x, shape), x, atol=1e-12)


@pytest.mark.parametrize("shape,out", [((8, 8), (4, 4)), ((9, 12), (5, 7)), ((64, 64), (16, 16))])
def test_constant_keeps_sum_or_mean(shape, out):
    x = np.full(shape, 1.5)
    y = D.fourier_crop_downsample_to_shape(x, out)
    np.testing.assert_allclose(y.sum(), x.sum(), rtol=1e-12)
    np.testing.assert_allclose(y, x.sum() / np.prod(out), rtol=1e-12)
    ym = D.fourier_crop_downsample_to_shape(x, out, preserve_mean=True)
    np.testing.assert_allclose(ym, 1.5, rtol=1e-12)


@pytest.mark.parametrize("shape,out,k", [((64, 64), (16, 16), (3, -5)), ((48, 40), (20, 15), (-7, 6)),
                                         ((33, 30), (11, 10), (4, 4))])
def test_band_limited_cosine_resamples(shape, out, k):
    h, w = shape
    oh, ow = out
    ph = 0.7
    x = np.cos(2 * np.pi * (k[0] * centred(h)[:, None] / h + k[1] * centred(w)[None, :] / w) + ph)
    want = np.cos(2 * np.pi * (k[0] * centred(oh)[:, None] / oh + k[1] * centred(ow)[None, :] / ow) + ph)
    y = D.fourier_crop_downsample_to_shape(x, out)
    np.testing.assert_allclose(y, want * (h * w) / (oh * ow), atol=1e-10)


@pytest.mark.parametrize("shape,out", [((32, 32), (15, 15)), ((27, 40), (9, 13)), ((64, 48), (21, 5))])
def test_odd_output_equals_fourier_crop(shape, out, rng):
    x = rng.standard_normal(shape)
    spec = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x)))
    crop = spec[tuple(slice(n // 2 - s // 2, n // 2 + s // 2 + 1) for n, s in zip(shape, out))]
    want = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(crop))).real
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(x, out), want, atol=1e-10)


@pytest.mark.parametrize("shape,out", [((32, 32), (8, 8)), ((24, 30), (12, 10))])
def test_even_output_unpaired_nyquist_bin(shape, out):
    h, w = shape
    oh, ow = out
    k = oh // 2  # the kept -oh/2 bin has no +oh/2 partner
    arg = 2 * np.pi * k * centred(h)[:, None] / h + np.zeros((1, w))
    sign = ((-1.0) ** centred(oh))[:, None] * np.ones((1, ow))
    amp = (h * w) / (2 * oh * ow)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(np.cos(arg), out), amp * sign, atol=1e-10)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(np.sin(arg), out), -amp * sign, atol=1e-10)


def test_fourier_output_inverts_to_real_space(rng):
    x = rng.standard_normal((40, 36))
    for pm in (False, True):
        y = D.fourier_crop_downsample_to_shape(x, (20, 18), preserve_mean=pm)
        f = D.fourier_crop_downsample_to_shape(x, (20, 18), outputs_real_space=False, preserve_mean=pm)
        assert f.shape == (20, 10)
        np.testing.assert_allclose(np.fft.fftshift(np.fft.irfftn(f, s=(20, 18), axes=(0, 1))), y, atol=1e-12)


def test_factor_and_shape_rules(rng):
    assert D.downsampled_shape((64, 64), 2) == (32, 32)
    assert D.downsampled_shape((65, 50), 1.5) == (43, 33)
    with pytest.raises(ValueError):
        D.downsampled_shape((8, 8), 0.5)
    x = rng.standard_normal((30, 30))
    np.testing.assert_allclose(D.fourier_crop_downsample(x, 3), D.fourier_crop_downsample_to_shape(x, (10, 10)))


def test_volume_constant_along_an_axis_reduces_to_images(rng):
    img = rng.standard_normal((20, 24))
    vol = np.broadcast_to(img, (16, 20, 24))
    y3 = D.fourier_crop_downsample_to_shape(vol, (8, 10, 12))
    y2 = D.fourier_crop_downsample_to_shape(img, (10, 12))
    np.testing.assert_allclose(y3, np.broadcast_to(y2 * 16 / 8, (8, 10, 12)), atol=1e-10)


@pytest.mark.parametrize("shape,out", [((12, 12), (12, 12)), ((9, 8, 10), (9, 8, 10))])
def test_complex_identity(shape, out, rng):
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(x, out), x, atol=1e-12)


@pytest.mark.parametrize("shape,out", [((24, 20), (9, 7)), ((16, 18, 20), (5, 9, 7))])
def test_complex_branch_of_real_input_equals_hartley_branch_for_odd_sizes(shape, out, rng):
    x = rng.standard_normal(shape)
    c = D.fourier_crop_downsample_to_shape(x.astype(np.complex128), out)
    r = D.fourier_crop_downsample_to_shape(x, out)
    np.testing.assert_allclose(c.real, r, atol=1e-12)
    np.testing.assert_allclose(c.imag, 0.0, atol=1e-12)


def test_complex_exponential_resamples_exactly():
    """exp(2 pi i (a y / h + b x / w)) with |a| < oh / 2, |b| < ow / 2 on centred coordinates comes back
    as the same exponential on the coarse grid, scaled by h w / (oh ow)."""
    h, w, oh, ow, a, b = 32, 24, 16, 12, 3, -5
    y = np.arange(h)[:, None] - h // 2
    x = np.arange(w)[None, :] - w // 2
    sig = np.exp(2j * np.pi * (a * y / h + b * x / w))
    yc = np.arange(oh)[:, None] - oh // 2
    xc = np.arange(ow)[None, :] - ow // 2
    want = np.exp(2j * np.pi * (a * yc / oh + b * xc / ow)) * (h * w) / (oh * ow)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(sig, (oh, ow)), want, atol=1e-10)
    np.testing.assert_allclose(D.fourier_crop_downsample_to_shape(sig, (oh, ow), preserve_mean=True),
                               want * (oh * ow) / (h * w), atol=1e-12)


def test_complex_fourier_output_inverts(rng):
    x = rng.standard_normal((10, 12, 14)) + 1j * rng.standard_normal((10, 12, 14))
    real = D.fourier_crop_downsample_to_shape(x, (6, 8, 7))
    four = D.fourier_crop_downsample_to_shape(x, (6, 8, 7), outputs_real_space=False)
    np.testing.assert_allclose(np.fft.fftshift(np.fft.ifftn(four)), real, atol=1e-12)
