"""Pin the oracle's multiplicative transforms and the host mirror of the Fourier operators (CPU only).

The filter (AbstractFilter.__call__, cryojax/ndimage/transforms/_filters.py:30-33), the mask
(AbstractMask.__call__, _masks.py:25-28), their places in postprocess (base_image_model.py:119-158)
and the CTF envelope (transfer_theory.py:117-118) are pinned by what the arithmetic fixes: unit
factors change nothing, a constant envelope scales the image, a filter that keeps only the zero
frequency leaves the constant F(0, 0) / P^2 (at the padded size, before the crop), and the mask
comes after the normalisation. FourierGaussian (_fourier_operator.py:226-234) is pinned by its
value at zero frequency and by the product rule of Gaussians; the host operators' arithmetic by
evaluating both sides.
"""

import numpy as np
import pytest

from cryojax_amd import transforms as T
from cryojax_amd.synthetic import gaussian_mixture_volume
from oracle import imaging as O

N = 24
CTF = dict(defocus=12000.0, astigmatism=300.0, astigmatism_angle=20.0)


@pytest.fixture(scope="module")
def setup():
    g = O.fourier_grid_from_real(gaussian_mixture_volume(N, 40, N // 5, 11))
    rng = np.random.default_rng(12)
    q = O.quat_normalize(rng.standard_normal((3, 4)))
    off = rng.uniform(-3, 3, (3, 2))
    return g, q, off


def sim(setup, **kw):
    g, q, off = setup
    return O.simulate(g, q, off, 1.5, **CTF, **kw)


@pytest.mark.parametrize("padded,norm", [(None, False), ((32, 32), False), (None, True), ((30, 30), True)])
def test_unit_factors_change_nothing(setup, padded, norm):
    kw = dict(padded_shape=padded, normalizes_signal=norm)
    ref = sim(setup, **kw)
    pr = ((N, N) if padded is None else padded)
    ones_f = np.ones((pr[0], pr[1] // 2 + 1))
    got = sim(setup, fourier_filter=ones_f, mask=np.ones((N, N)), envelope=lambda f: np.ones(f.shape[:-1]), **kw)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_constant_envelope_scales_the_image(setup):
    ref = sim(setup)
    got = sim(setup, envelope=lambda f: 0.37)
    np.testing.assert_allclose(got, 0.37 * ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("padded", [None, (32, 32), (30, 30)])
def test_zero_frequency_filter_leaves_the_mean(setup, padded):
    """A filter that keeps only F(0, 0) leaves the constant F(0, 0) / P^2 of the padded image, so
    the filter acts before the inverse transform at the padded size (and before the crop)."""
    pr = (N, N) if padded is None else padded
    spec = O.compute_fourier_image(setup[0], setup[1], setup[2], 1.5, padded_shape=pr, **CTF)
    f = np.zeros((pr[0], pr[1] // 2 + 1))
    f[0, 0] = 1.0
    got = sim(setup, padded_shape=padded, fourier_filter=f)
    want = (spec[:, 0, 0].real / (pr[0] * pr[1]))[:, None, None] * np.ones((1, N, N))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


def test_mask_comes_after_the_normalisation(setup):
    m = np.zeros((N, N))
    m[4:20, 6:18] = 1.0
    normed = sim(setup, normalizes_signal=True)
    got = sim(setup, normalizes_signal=True, mask=m)
    np.testing.assert_array_equal(got[:, m == 0], 0.0)
    np.testing.assert_allclose(got[:, m == 1], normed[:, m == 1], rtol=1e-13)
    # normalised over the whole crop, not over the window
    assert abs(got[:, m == 1].std() - 1.0) > 1e-3


def test_fourier_gaussian_closed_forms():
    f = O.make_frequency_grid((16, 16), outputs_rfftfreqs=True) / 1.3
    g = O.fourier_gaussian(f, amplitude=0.8, b_factor=40.0)
    assert g[0, 0] == 0.8 and np.all(g <= 0.8) and np.all(g > 0)
    # exp(-b1 k^2 / 4) exp(-b2 k^2 / 4) = exp(-(b1 + b2) k^2 / 4)
    np.testing.assert_allclose(O.fourier_gaussian(f, 0.5, 10.0) * O.fourier_gaussian(f, 3.0, 25.0),
                               O.fourier_gaussian(f, 1.5, 35.0), rtol=1e-14)


def test_host_operators_match_the_oracle_and_compose():
    f = O.make_frequency_grid((20, 20), outputs_rfftfreqs=True) / 0.9
    np.testing.assert_allclose(T.FourierGaussian(0.8, 40.0)(f), O.fourier_gaussian(f, 0.8, 40.0), rtol=1e-15)
    a, b = T.FourierGaussian(0.5, 10.0), T.FourierGaussian(3.0, 25.0)
    np.testing.assert_allclose((a * b)(f), T.FourierGaussian(1.5, 35.0)(f), rtol=1e-14)
    np.testing.assert_allclose((a + 2.0)(f), a(f) + 2.0, rtol=1e-15)
    np.testing.assert_allclose((1.0 - a)(f), 1.0 - a(f), rtol=1e-15)
    np.testing.assert_allclose((2.0 * a)(f), 2.0 * a(f), rtol=1e-15)
    np.testing.assert_array_equal(T.Empirical(np.arange(6.0))(f), np.arange(6.0))
    z = T.ZeroMode(2.5)(f)
    assert z.shape == f.shape[:-1] and z[0, 0] == 2.5 and z.sum() == 2.5
    ls = 3.0
    k2 = np.sum(f**2, axis=-1)
    np.testing.assert_allclose(T.Lorenzian(2.0, ls)(f) * (k2 + 1 / ls**2), 2.0 / ls**2, rtol=1e-13)
    np.testing.assert_allclose(T.FourierExp2D(2.0, ls)(f) * (k2 + 1 / ls**2) ** 1.5,
                               2.0 / (2 * np.pi * ls**3), rtol=1e-13)
    with pytest.raises(ValueError):
        T.FourierGaussian(1.0, 0.0)


def test_transform_arrays():
    a = np.arange(12.0).reshape(3, 4)
    assert T.transform_array(T.CustomFilter(a)) is not None
    np.testing.assert_array_equal(T.transform_array(T.CustomMask(a) * T.CustomMask(2 * a)), 2 * a * a)
    np.testing.assert_array_equal(T.transform_array(a), a)
    with pytest.raises(NotImplementedError):
        T.transform_array(lambda x: x)


def test_volume_filter_closed_forms():
    """from_real_voxel_grid's filter acts on fftn(x) with the zero frequency in the corner: a
    filter of ones changes nothing, and one that keeps only [0, 0, 0] leaves the sum of the
    volume at the centred grid's index (n // 2,) * 3 and zeros elsewhere."""
    rng = np.random.default_rng(81)
    for shape, pad in (((8, 8, 8), 1.0), ((7, 7, 7), 1.0), ((6, 6, 6), 1.5)):
        x = rng.standard_normal(shape)
        plain = O.fourier_grid_from_real(x, pad_scale=pad)
        p = plain.shape
        np.testing.assert_array_equal(O.fourier_grid_from_real(x, pad_scale=pad, fourier_filter=np.ones(p)), plain)
        dc = np.zeros(p)
        dc[0, 0, 0] = 1.0
        got = O.fourier_grid_from_real(x, pad_scale=pad, fourier_filter=dc)
        want = np.zeros(p, dtype=complex)
        want[p[0] // 2, p[1] // 2, p[2] // 2] = x.sum()
        np.testing.assert_allclose(got, want, atol=1e-12)


def test_batched_operators_evaluate_per_image():
    """A batch of operator parameters (one per particle, as a vmapped operator holds them) and a
    batch of pixel sizes each give one array per image, equal to the oracle's per-image value."""
    from cryojax_amd.simulator import BasicImageConfig, _envelope_factor

    amp, b = np.array([0.5, 1.0, 2.0]), np.array([0.0, 30.0, 90.0])
    cfg = BasicImageConfig((20, 20), 1.7, 300.0, pad_options=dict(shape=(24, 24)))
    f = cfg.padded_frequency_grid_in_angstroms
    env = T.FourierGaussian(1.0, 1.0)
    env.amplitude, env.b_factor = amp, b
    got = _envelope_factor(env, cfg, (3, 24, 13))
    assert got.shape == (3, 24, 13)
    for i in range(3):
        np.testing.assert_allclose(got[i], O.fourier_gaussian(f, amp[i], b[i]), rtol=1e-15)
    np.testing.assert_array_equal(_envelope_factor(T.Constant(amp), cfg, (3, 24, 13)),
                                  np.broadcast_to(amp[:, None, None], (3, 24, 13)))
    np.testing.assert_array_equal(_envelope_factor(T.Constant(0.3), cfg, (3, 24, 13)), np.full((24, 13), 0.3))
    with pytest.raises(ValueError):
        _envelope_factor(T.Constant(amp), cfg, (4, 24, 13))
    ps = np.array([1.0, 1.7, 2.5])
    cfgb = BasicImageConfig((20, 20), ps, 300.0, pad_options=dict(shape=(24, 24)))
    fb = cfgb.padded_frequency_grid_in_angstroms
    assert fb.shape == (3, 24, 13, 2)
    got = _envelope_factor(T.FourierGaussian(0.8, 40.0), cfgb, (3, 24, 13))
    for i in range(3):
        fi = O.make_frequency_grid((24, 24), outputs_rfftfreqs=True) / ps[i]
        np.testing.assert_allclose(got[i], O.fourier_gaussian(fi, 0.8, 40.0), rtol=1e-14)


# ---------------------------------------------------------------------------
# built-in filters, masks and operators (host arrays), pinned by closed forms and by agreement
# between the reference's different constructions of the same shape
# ---------------------------------------------------------------------------


def test_coordinate_and_frequency_grids():
    from cryojax_amd import coordinates as C

    for shape in ((10, 10), (9, 12), (7, 8)):
        for rfft in (True, False):
            np.testing.assert_array_equal(C.make_frequency_grid(shape, outputs_rfftfreqs=rfft),
                                          O.make_frequency_grid(shape, outputs_rfftfreqs=rfft))
        g = C.make_coordinate_grid(shape, 1.5)
        # (i - n // 2) dx exactly; the reference's fftshift(fftfreq(n, 1 / dx)) * n rounds
        np.testing.assert_allclose(g[0, :, 0], O.make_coordinate_grid_1d(shape[1], 1.5), rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(g[:, 0, 1], O.make_coordinate_grid_1d(shape[0], 1.5), rtol=1e-15, atol=1e-15)
    g3 = C.make_coordinate_grid((3, 4, 5))
    assert g3.shape == (3, 4, 5, 3)
    np.testing.assert_array_equal(g3[1, 2, :, 0], [-2, -1, 0, 1, 2])
    np.testing.assert_array_equal(g3[1, :, 3, 1], [-2, -1, 0, 1])
    np.testing.assert_array_equal(g3[:, 0, 0, 2], [-1, 0, 1])
    f3 = C.make_frequency_grid((4, 6, 8))
    assert f3.shape == (4, 6, 5, 3) and f3[0, 0, 4, 0] == 0.5 and f3[2, 0, 0, 2] == -0.5


def test_radial_masks_and_filters_closed_forms():
    from cryojax_amd import coordinates as C

    g = C.make_coordinate_grid((40, 40))
    r = np.linalg.norm(g, axis=-1)
    m = T.CircularCosineMask(g, 8.0, 4.0).get()
    np.testing.assert_array_equal(m[r <= 8.0], 1.0)
    np.testing.assert_array_equal(m[r > 12.0], 0.0)
    band = (r > 8.0) & (r <= 12.0)
    np.testing.assert_allclose(m[band], np.cos(np.pi * (r[band] - 8.0) / 8.0) ** 2, rtol=1e-12, atol=1e-15)
    # at the middle of the band the edge is one half (pixel (0, 10): r = 10)
    assert m[20, 30] == pytest.approx(0.5)
    off = T.CircularCosineMask(g, 8.0, 4.0, xy_offset=(3.0, -2.0)).get()
    np.testing.assert_array_equal(off[18:38, 23:], m[20:, 20:37])  # shifted by (+3 in x, -2 in y)
    g3 = C.make_coordinate_grid((9, 40, 40))
    np.testing.assert_array_equal(T.SphericalCosineMask(g3, 8.0, 4.0).get()[4], m)
    f = C.make_frequency_grid((32, 32))
    lo = T.LowpassFilter(f, 1.0, 0.5, 0.2).get()
    k = np.linalg.norm(f, axis=-1)
    np.testing.assert_array_equal(lo[k <= 0.25], 1.0)
    np.testing.assert_array_equal(lo[k > 0.35], 0.0)
    np.testing.assert_allclose(lo + T.HighpassFilter(f, 1.0, 0.5, 0.2).get(), 1.0, rtol=1e-15)
    # grid spacing scales the Nyquist frequency: the same filter on a grid in inverse Angstroms
    np.testing.assert_allclose(T.LowpassFilter(f / 2.5, 2.5, 0.5, 0.2).get(), lo, rtol=1e-13, atol=1e-15)
    f3 = C.make_frequency_grid((16, 16, 16))
    lo3 = T.LowpassFilter(f3).get()
    assert lo3.shape == (16, 16, 9) and lo3[0, 0, 0] == 1.0 and lo3[8, 8, 8] == 0.0


def test_box_masks_agree_with_each_other():
    from cryojax_amd import coordinates as C

    g = C.make_coordinate_grid((36, 36))
    sq = T.SquareCosineMask(g, 10.0, 2.5).get()
    # along y = 0 the square is the 1D cosine profile of |x|
    x = np.abs(g[18, :, 0])
    prof = np.where(x <= 5, 1.0, np.where(x < 7.5, np.cos(np.pi * (x - 5) / 5) ** 2, 0.0))
    np.testing.assert_allclose(sq[18], prof, atol=1e-15)
    # in the corner band it is the product of the two edges
    np.testing.assert_allclose(sq[18 + 6, 18 + 7], prof[18 + 6] * prof[18 + 7], rtol=1e-14)
    np.testing.assert_allclose(T.Rectangular2DCosineMask(g, 10.0, 10.0, 2.5).get(), sq, atol=1e-15)
    # a point exactly on the grown square lies in no edge band; the reference's selection then
    # takes the y edge, 0.5 (1 + cos(pi (|y| - 5) / 3)) = 0.75 at (x, y) = (8, 0)
    assert T.SquareCosineMask(g, 10.0, 3.0).get()[18, 26] == pytest.approx(0.75)
    # a rectangle rotated by 90 degrees is the rectangle with its widths swapped
    np.testing.assert_allclose(T.Rectangular2DCosineMask(g, 12.0, 6.0, 2.5, in_plane_rotation_angle=90.0).get(),
                               T.Rectangular2DCosineMask(g, 6.0, 12.0, 2.5).get(), atol=1e-12)
    # an infinite cylinder at angle 0 is a function of y alone; a very long one equals it
    cyl = T.Cylindrical2DCosineMask(g, 4.3, 2.0).get()
    y = np.abs(g[:, 0, 1])
    want = np.where(y < 4.3, 1.0, np.where(y < 6.3, np.cos(np.pi * (y - 4.3) / 4.0) ** 2, 0.0))
    np.testing.assert_allclose(cyl, np.repeat(want[:, None], 36, axis=1), atol=1e-15)
    np.testing.assert_allclose(T.Cylindrical2DCosineMask(g, 4.3, 2.0, length=1000.0).get(), cyl, atol=1e-15)
    g3 = C.make_coordinate_grid((7, 36, 36))
    box = T.Rectangular3DCosineMask(g3, 10.0, 10.0, 1000.0, 2.5).get()
    np.testing.assert_allclose(box[3], sq, atol=1e-15)
    z = T.Rectangular3DCosineMask(g3, 1000.0, 1000.0, 2.0, 2.0).get()[:, 18, 18]
    zz = np.abs(g3[:, 0, 0, 2])
    np.testing.assert_allclose(z, np.where(zz <= 1, 1.0, np.where(zz <= 3, np.cos(np.pi * (zz - 1) / 4) ** 2, 0.0)),
                               atol=1e-15)


def test_inverse_sinc_mask():
    from cryojax_amd import coordinates as C

    g = C.make_coordinate_grid((8, 12))
    m = T.InverseSincMask(g).get()
    assert m[4, 6] == 1.0  # the centre
    # x = -6 = -n_x / 2: 1 / sinc(1 / 2) = pi / 2; y = -4 = -n_y / 2 likewise
    assert m[4, 0] == pytest.approx(np.pi / 2) and m[0, 6] == pytest.approx(np.pi / 2)
    np.testing.assert_allclose(m, np.outer(m[:, 6], m[4, :]), rtol=1e-14)  # separable
    g3 = C.make_coordinate_grid((4, 8, 12))
    np.testing.assert_allclose(T.InverseSincMask(g3).get()[2], m, rtol=1e-15)


def test_more_fourier_operators():
    from cryojax_amd import coordinates as C

    f = C.make_frequency_grid((24, 24)) / 1.2
    np.testing.assert_allclose(T.FourierGaussianWithRadialOffset(0.7, 30.0, 0.0)(f),
                               T.FourierGaussian(0.7, 30.0)(f), rtol=1e-14)
    k = np.linalg.norm(f, axis=-1)
    ro = T.FourierGaussianWithRadialOffset(0.7, 30.0, 0.2)(f)
    np.testing.assert_allclose(ro[np.isclose(k, 0.2)], 0.7, rtol=1e-14)
    with pytest.raises(ValueError):
        T.FourierGaussianWithRadialOffset(1.0, 1.0, -0.1)
    s = T.FourierSinc(2.0)(f)
    assert s[0, 0] == 1.0
    np.testing.assert_allclose(s, np.sinc(2 * f[..., 0]) * np.sinc(2 * f[..., 1]), rtol=1e-15)
    assert abs(T.FourierSinc(2.5)(np.array([[[0.4, 0.1]]])))[0, 0] < 1e-15  # sinc(1) = 0
    g = C.make_coordinate_grid((31, 31), 0.5)
    gauss = T.Gaussian2D(2.0, 1.5, offset=(1.0, -0.5))(g)
    assert gauss[15 - 1, 15 + 2] == pytest.approx(2.0 / np.sqrt(2 * np.pi * 1.5))  # peak at the offset
    with pytest.raises(ValueError):
        T.Gaussian2D(1.0, 0.0)


@pytest.mark.parametrize("shape", [None, (9, 9), (11, 11), (12, 12)])
@pytest.mark.parametrize("mode", ["linear", "nearest"])
@pytest.mark.parametrize("square", [False, True])
def test_whitening_filter_of_a_flat_spectrum(shape, mode, square):
    """A flat power spectrum c (white images) whitens to 1 / sqrt(c) (or 1 / c) everywhere but the
    zero frequency and, for even sizes, the Nyquist corner; resizing in real space keeps a flat
    spectrum flat (its inverse transform is c times a centred delta)."""
    from cryojax_amd.transforms import _whitening_filter

    c = 0.37
    out = _whitening_filter(np.full((10, 6), c), (10, 10), shape, mode, square)
    sh = (10, 10) if shape is None else shape
    assert out.shape == (sh[0], sh[1] // 2 + 1)
    want = np.full(out.shape, 1 / c if square else 1 / np.sqrt(c))
    want[0, 0] = 0.0
    if sh[0] % 2 == 0 and sh[1] % 2 == 0:
        want[sh[0] // 2, sh[1] // 2] = 0.0
    np.testing.assert_allclose(out, want, rtol=1e-12)
    with pytest.raises(ValueError):
        _whitening_filter(np.full((10, 6), c), (10, 10), None, "cubic", False)


def test_resize_with_crop_or_pad():
    a = np.arange(30.0).reshape(5, 6)
    np.testing.assert_array_equal(T.resize_with_crop_or_pad(a, (3, 4)), a[1:4, 1:5])
    p = T.resize_with_crop_or_pad(a, (8, 7))
    assert p.shape == (8, 7)
    # 5 -> 8 (odd to even): the extra row goes left; 6 -> 7 (even to odd): the extra column right
    np.testing.assert_array_equal(p[2:7, 0:6], a)
    e = T.resize_with_crop_or_pad(a, (7, 4), mode="edge")
    np.testing.assert_array_equal(e[0], a[0, 1:5])
    np.testing.assert_array_equal(e[1:6], a[:, 1:5])


@pytest.mark.parametrize("square", [False, True])
def test_whitening_filter_of_a_binwise_constant_spectrum(square):
    """A power spectrum that is constant on each radial bin (bins of width 1 / max(rfft shape) up
    to sqrt(2) / 2, a frequency in bin b when b_{i-1} < |k| <= b_i) has bin means equal to those
    constants, so the nearest-bin filter whitens it exactly: 1 / sqrt(P) (or 1 / P), with the zero
    frequency and the even Nyquist corner set to 0 (_filters.py:228-297)."""
    from cryojax_amd import coordinates as C
    from cryojax_amd.transforms import _whitening_filter

    h = w = 24
    k = C.make_radial_frequency_grid((h, w))
    n_bins = 1 + int((np.sqrt(2) / 2) * max(h, w // 2 + 1))
    edges = np.linspace(0.0, np.sqrt(2) / 2, n_bins)
    level = np.random.default_rng(5).uniform(0.5, 3.0, n_bins)
    power = level[np.digitize(k, edges, right=True)]
    out = _whitening_filter(power, (h, w), None, "nearest", square)
    want = 1 / power if square else 1 / np.sqrt(power)
    want[0, 0] = 0.0
    want[h // 2, w // 2] = 0.0
    np.testing.assert_allclose(out, want, rtol=1e-13)
