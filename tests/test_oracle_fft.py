"""Pins of the FP64 volume transform and observed-image spectra restatements
against closed forms (no GPU)."""

import itertools

import numpy as np
import pytest

from oracle import imaging as I
from oracle import search as S


@pytest.mark.parametrize("p,s", [(8, 6), (9, 6), (8, 5), (9, 5), (7, 7), (10, 3)])
def test_pad_split_aligns_centres(p, s):
    lo, hi = I.left_right_pad(p, s)
    assert lo + hi == p - s
    assert lo == p // 2 - s // 2


def plane_wave(shape, padded, pos, natural_last=False):
    """exp(-2 pi i sum (q~)(m - s//2) / P) for a unit voxel at `pos`."""
    out = np.ones(padded, dtype=np.complex128)
    for ax, (s, p, m) in enumerate(zip(shape, padded, pos)):
        q = np.arange(p) - p // 2
        ph = np.exp(-2j * np.pi * q * (m - s // 2) / p)
        sh = [1] * len(padded)
        sh[ax] = p
        out = out * ph.reshape(sh)
    return out


@pytest.mark.parametrize("shape,padded", [((6, 6, 6), (6, 6, 6)), ((5, 6, 7), (8, 9, 10)), ((6, 5, 4), (7, 8, 8))])
def test_voxel_grid_of_single_voxel_is_plane_wave(shape, padded, rng):
    for _ in range(3):
        pos = tuple(int(rng.integers(0, s)) for s in shape)
        x = np.zeros(shape)
        x[pos] = 1.0
        got = I.fourier_grid_from_real(x, padded_shape=padded)
        np.testing.assert_allclose(got, plane_wave(shape, padded, pos), atol=1e-12)


@pytest.mark.parametrize("h,w", [(8, 8), (7, 10), (9, 5)])
def test_observed_spectrum_of_single_pixel(h, w):
    for pos in itertools.product([0, h // 2, h - 1], [0, w // 2, w - 1]):
        y = np.zeros((h, w))
        y[pos] = 1.0
        got = S.observed_spectrum(y[None])[0]
        qh = np.arange(h)[:, None]
        qw = np.arange(w // 2 + 1)[None, :]
        want = np.exp(-2j * np.pi * (qh * (pos[0] - h // 2) / h + qw * (pos[1] - w // 2) / w))
        np.testing.assert_allclose(got, want, atol=1e-12)
