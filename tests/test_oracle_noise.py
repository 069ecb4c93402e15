"""Pin the noise-model / normalisation restatement (CPU)."""

import numpy as np
from scipy.stats import norm

from oracle import noise as N


def test_normalize_closed_form(rng):
    x = rng.standard_normal((32, 32)) * 3 + 5
    y = N.normalize_image(x)
    assert abs(y.mean()) < 1e-12 and abs(y.std() - 1) < 1e-12
    where = np.zeros((32, 32), bool)
    where[8:24, 8:24] = True
    z = N.normalize_image(x, where)
    assert abs(z[where].mean()) < 1e-12 and abs(z[where].std() - 1) < 1e-12
    np.testing.assert_allclose(z, (x - x[where].mean()) / x[where].std())


def test_fourier_branch_matches_real_branch(rng):
    """Normalising the rfft then inverting equals normalising in real space, for odd
    sizes where N1 * (2 N2 - 1) is the pixel count the Fourier branch assumes."""
    x = rng.standard_normal((33, 33)) * 2 - 1
    f = np.fft.rfftn(x)
    y = np.fft.irfftn(N.normalize_image_fourier(f), s=x.shape)
    np.testing.assert_allclose(y, N.normalize_image(x), atol=1e-12)
    # even sizes: the branch weights the Nyquist column by 2 like every other
    # column, so the result is zero-mean with std = sqrt(S_true / S_branch),
    # S the weighted power sums (Parseval), slightly below 1
    x = rng.standard_normal((32, 32))
    f = np.fft.rfftn(x)
    y = np.fft.irfftn(N.normalize_image_fourier(f, shape_in_real_space=(32, 32)), s=x.shape)
    f0 = f.copy()
    f0[0, 0] = 0
    p = np.abs(f0) ** 2
    s_true = p[:, 0].sum() + p[:, -1].sum() + 2 * p[:, 1:-1].sum()
    s_branch = p[:, 0].sum() + 2 * p[:, 1:].sum()
    assert abs(y.mean()) < 1e-12
    np.testing.assert_allclose(y.std(), np.sqrt(s_true / s_branch), rtol=1e-12)
    np.testing.assert_allclose(y / y.std(), N.normalize_image(x), atol=1e-12)


def test_log_likelihood_matches_scipy(rng):
    sim = rng.standard_normal((3, 16, 16))
    obs = sim + rng.standard_normal((3, 16, 16)) * 0.7
    want = norm.logpdf(obs, loc=sim, scale=np.sqrt(0.49)).sum(axis=(1, 2))
    np.testing.assert_allclose(N.log_likelihood(sim, obs, 0.49), want, rtol=1e-12)


def test_noise_structure_and_variance(rng):
    n, m = 32, 17
    z = rng.standard_normal((4000, n, m))
    img = N.noise_image_from_draws(z, 2.5)
    assert np.max(np.abs(img.mean(axis=(1, 2)))) < 1e-12  # zero DC
    c = n // 2
    # real spectrum -> even function about the (shifted) origin
    np.testing.assert_allclose(img[:, c + 1:, c + 1:], img[:, c - 1:0:-1, c - 1:0:-1], atol=1e-12)
    v = img.var(axis=0).mean()
    assert abs(v - 2.5) < 0.1, v
