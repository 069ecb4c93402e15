"""Pin the grid-search and NCC oracle (CPU only)."""

import json
import os

import numpy as np
import pytest

from oracle import imaging as O
from oracle import search as S
from cryojax_amd.synthetic import gaussian_mixture_volume

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")


def cost_fn(point, args):
    # tests/test_jax_util.py:127-133 (vectorised over a leading batch axis)
    variance, (mu_x, mu_y) = args
    x, y = point
    return -np.exp(-((x - mu_x) ** 2 + (y - mu_y) ** 2) / (2 * variance)) / np.sqrt(2 * np.pi * variance)


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["grid_search_landscape"]["cases"])
def test_landscape_known_minimum(case):
    batch_size, dim, ox, oy, variance = case
    args = (variance, (ox, oy))
    c = np.fft.fftshift(np.fft.fftfreq(dim)) * dim
    xx, yy = np.meshgrid(c, c, indexing="ij")
    landscape = cost_fn((xx, yy), args)
    true_min = landscape.min()
    i, j = np.unravel_index(np.argmin(landscape), landscape.shape)
    st = S.run_grid_search_minimum(cost_fn, (c, c), args, batch_size)
    np.testing.assert_allclose(st.current_minimum_eval, true_min)
    xi, yi = S.tree_grid_unravel_index(st.current_best_raveled_index, (c, c))
    np.testing.assert_allclose((c[xi], c[yi]), (c[i], c[j]))


@pytest.mark.parametrize("batch_size", [None, 1, 3, 4, 7])
def test_ties_lowest_raveled_index_wins(batch_size):
    vals = np.array([5.0, 2.0, 3.0, 1.0, 4.0, 1.0, 1.0, 9.0, 1.0, 6.0])
    grid = (np.arange(10),)

    def fn(point, args):
        return vals[point[0]]

    st = S.run_grid_search_minimum(fn, grid, None, batch_size)
    assert st.current_best_raveled_index == 3 and st.current_minimum_eval == 1.0
    # per-element minimum over a vector output
    vals2 = np.stack([vals, vals[::-1]], -1)
    st2 = S.run_grid_search_minimum(lambda p, a: vals2[p[0]], grid, None, batch_size)
    np.testing.assert_array_equal(st2.current_best_raveled_index, [3, 1])


def test_unravel_clips_like_jnp():
    grid = (np.arange(3), np.arange(4))
    assert tuple(int(v) for v in S.tree_grid_unravel_index(11, grid)) == (2, 3)
    assert tuple(int(v) for v in S.tree_grid_unravel_index(17, grid)) == (2, 3)
    assert tuple(int(v) for v in S.tree_grid_unravel_index(5, grid)) == (1, 1)


def test_super_fibonacci_is_unit_and_covers(rng):
    q = S.super_fibonacci_quaternions(30000)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    probe = O.quat_normalize(rng.standard_normal((300, 4)))
    nearest = np.array([S.geodesic_degrees(p[None], q).min() for p in probe])
    assert nearest.max() < 9.0  # covering radius of a ~7.5 degree grid
    assert np.median(nearest) < 6.0


def test_geodesic_known_angles():
    q0 = np.array([1.0, 0, 0, 0])
    for deg in (5.0, 30.0, 179.0):
        q = O.so3_exp(np.array([0.0, 0.0, np.deg2rad(deg)]))
        assert abs(S.geodesic_degrees(q0, q) - deg) < 1e-9
    # q and -q are the same rotation
    assert S.geodesic_degrees(q0, -q0) < 1e-6


def test_ncc_fourier_matches_real_space(rng):
    n = 32
    y = rng.standard_normal((n, n))
    t = rng.standard_normal((n, n))
    w = S.rfft_weights(n)
    Y, T = O.rfftn(y), O.rfftn(t)
    mask = (w > 0).astype(float)
    yr = O.irfftn(Y * mask, s=(n, n))
    tr = O.irfftn(T * mask, s=(n, n))
    real = np.sum(yr * tr) / np.sqrt(np.sum(yr * yr) * np.sum(tr * tr))
    four = np.sum(w * np.real(Y * np.conj(T))) / np.sqrt(np.sum(w * np.abs(Y) ** 2) * np.sum(w * np.abs(T) ** 2))
    assert abs(real - four) < 1e-12
    assert abs(np.mean(yr)) < 1e-12  # zero mean: DC removed


@pytest.fixture(scope="module")
def small_problem():
    n = 32
    vol = gaussian_mixture_volume(n, 40, 8, seed=4)
    g = O.fourier_grid_from_real(vol)
    quats = S.super_fibonacci_quaternions(60)
    offs = S.offset_grid(3, 1.0)
    rng = np.random.default_rng(9)
    n_img = 4
    rot_i = rng.integers(0, 60, n_img)
    off_i = rng.integers(0, 9, n_img)
    ctf = {"defocus": rng.uniform(8000, 12000, n_img), "astigmatism": rng.uniform(0, 300, n_img),
           "astigmatism_angle": rng.uniform(0, 180, n_img)}
    imgs = O.simulate(g, quats[rot_i], offs[off_i], 1.0, defocus=ctf["defocus"], astigmatism=ctf["astigmatism"],
                      astigmatism_angle=ctf["astigmatism_angle"])
    return g, quats, offs, ctf, imgs, rot_i * 9 + off_i


def test_ncc_noise_free_hits_generating_pose(small_problem):
    g, quats, offs, ctf, imgs, truth = small_problem
    y = S.observed_spectrum(imgs)
    best, idx = S.ncc_exhaustive(g, y, ctf, quats, offs, 1.0)
    np.testing.assert_array_equal(idx, truth)
    np.testing.assert_allclose(best, -1.0, atol=1e-9)
    # affine changes of the observed intensities do not move the scores
    y2 = S.observed_spectrum(3.5 * imgs + 7.0)
    s1 = S.ncc_scores(g, y, ctf, quats[:5], offs, 1.0)
    s2 = S.ncc_scores(g, y2, ctf, quats[:5], offs, 1.0)
    np.testing.assert_allclose(s1, s2, atol=1e-12)


def test_ncc_grid_search_loop_matches_exhaustive(small_problem, rng):
    g, quats, offs, ctf, imgs, _ = small_problem
    noisy = imgs + rng.standard_normal(imgs.shape) * imgs.std() * 2.0
    y = S.observed_spectrum(noisy)
    q, o = quats[:12], offs
    ref_v, ref_i = S.ncc_exhaustive(g, y, ctf, q, o, 1.0)
    for bs in (None, 5, 7):
        v, i = S.ncc_grid_search(g, y, ctf, q, o, 1.0, batch_size=bs)
        np.testing.assert_array_equal(i, ref_i)
        np.testing.assert_allclose(v, ref_v, rtol=1e-12)


@pytest.mark.parametrize("wrap", [True, False])
def test_chunked_exhaustive_equals_exhaustive(wrap):
    """The parallel chunked oracle returns ncc_exhaustive's minimum and index, including the
    lowest-index rule for exact ties that span chunks (duplicated rotations)."""
    rng = np.random.default_rng(3)
    n = 16
    g = O.fourier_grid_from_real(gaussian_mixture_volume(n, 20, 4, 3))
    quats = S.super_fibonacci_quaternions(70)
    quats = np.concatenate([quats, quats[::-1]])  # every rotation twice, the copies in another chunk
    offs = S.offset_grid(3, 1.0)
    ctf = {"defocus": rng.uniform(8000, 12000, 3), "astigmatism": rng.uniform(0, 300, 3),
           "astigmatism_angle": rng.uniform(0, 180, 3)}
    imgs = O.simulate(g, quats[[5, 40, 66]], offs[[0, 4, 8]], 1.0, wrap_negative_indices=wrap, **ctf)
    imgs = imgs + 0.5 * imgs.std() * rng.standard_normal(imgs.shape)
    y = S.observed_spectrum(imgs)
    ref_s, ref_i = S.ncc_exhaustive(g, y, ctf, quats, offs, 1.0, wrap_negative_indices=wrap)
    got_s, got_i = S.ncc_exhaustive_chunked(g, y, ctf, quats, offs, 1.0, chunk=32, workers=2,
                                            wrap_negative_indices=wrap)
    np.testing.assert_array_equal(got_i, ref_i)
    np.testing.assert_allclose(got_s, ref_s, rtol=0, atol=1e-13)
    assert np.all(ref_i < 70 * 9)  # the first copy of the winning rotation
