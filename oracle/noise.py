"""FP64 restatement of the noise model and image normalisation.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows cryojax/ndimage/_normalize.py:11-98 (normalize_image /
rescale_image, real- and Fourier-space branches),
cryojax/simulator/_image_model/base_image_model.py:181-189 (_normalize_image)
and cryojax/simulator/_noise_model/gaussian_noise_model.py:171-251
(compute_noise, log_likelihood).

Pinning (tests/test_oracle_noise.py): the real-space branch by its closed
form (zero mean, unit population std over `where`); the Fourier-space branch
by agreement with the real-space branch after an inverse FFT; the
log-likelihood against scipy.stats.norm.logpdf; noise images by their exact
structure (zero mean, point symmetry about the centre, which a real-valued
half-plane spectrum forces) and their per-pixel variance over many draws.
The random stream itself (JAX threefry in the reference) is not restated:
parity unpinned for individual draws.
"""

from __future__ import annotations

import math

import numpy as np


def normalize_image(image, where=None):
    """rescale_image(image, 1, 0) for real-space input (_normalize.py:73-78)."""
    x = np.asarray(image, dtype=np.float64)
    sel = x if where is None else x[np.asarray(where, dtype=bool)]
    return (x - sel.mean()) / sel.std()


def normalize_image_fourier(spectrum, shape_in_real_space=None):
    """rescale_image(image, 1, 0, input_is_real_space=False, input_is_rfft=True) (_normalize.py:79-97)."""
    f = np.array(spectrum, dtype=np.complex128)
    n1, n2 = f.shape
    n_pixels = n1 * (2 * n2 - 1) if shape_in_real_space is None else math.prod(shape_in_real_space)
    f[0, 0] = 0.0
    std = np.sqrt(np.sum(np.abs(f[:, 0]) ** 2) + 2 * np.sum(np.abs(f[:, 1:]) ** 2)) / n_pixels
    out = f / std
    out[0, 0] = 0.0
    return out


def noise_image_from_draws(draws, variance):
    """compute_noise for given standard-normal draws on the (N, N/2+1) half-plane
    (gaussian_noise_model.py:192-205): fftshift(irfftn(sqrt(N^2 var) z, [0,0] = 0))."""
    z = np.array(draws, dtype=np.float64)
    n = z.shape[-2]
    spec = math.sqrt(n * n * variance) * z
    spec[..., 0, 0] = 0.0
    return np.fft.fftshift(np.fft.irfftn(spec.astype(np.complex128), s=(n, n), axes=(-2, -1)), axes=(-2, -1))


def log_likelihood(simulated, observed, variance):
    """gaussian_noise_model.py:239-251, summed over the last two axes."""
    r = np.asarray(simulated, dtype=np.float64) - np.asarray(observed, dtype=np.float64)
    per_pixel = -(np.abs(r) ** 2 / (2 * variance) + np.log(2 * np.pi * variance) / 2)
    return per_pixel.sum(axis=(-2, -1))
