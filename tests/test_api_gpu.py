"""The cryoJAX-API mirror (cryojax_amd.simulator) against the FP64 oracle,
written the way the reference's own tests drive make_image_model/simulate."""

import numpy as np
import pytest
import torch

import cryojax_amd.simulator as cxs
from oracle import imaging as O
from tests.helpers import fourier_grid, make_volume, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def volume64():
    vol = make_volume(64, 200, 20, 7)
    return vol, fourier_grid(vol)


def test_make_image_model_single_image(volume64):
    vol, g = volume64
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    config = cxs.BasicImageConfig((64, 64), pixel_size=2.0, voltage_in_kilovolts=300.0)
    pose = cxs.EulerAnglePose(offset_x_in_angstroms=3.0, offset_y_in_angstroms=-2.0, phi_angle=20.0,
                              theta_angle=50.0, psi_angle=-30.0)
    ctf = cxs.AstigmaticCTF(defocus_in_angstroms=11000.0, astigmatism_in_angstroms=300.0, astigmatism_angle=25.0)
    model = cxs.make_image_model(volume, config, pose, cxs.ContrastTransferTheory(ctf, amplitude_contrast_ratio=0.1))
    img = model.simulate().cpu().numpy()
    q = O.euler_to_quat(20.0, 50.0, -30.0)
    ref = O.simulate(g, q, np.array([3.0, -2.0]), 2.0, defocus=11000.0, astigmatism=300.0, astigmatism_angle=25.0)
    assert img.shape == (64, 64)
    assert rel_l2(img, ref) <= 1e-5


def test_batched_leaves_are_vmap(volume64, rng):
    vol, g = volume64
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    config = cxs.BasicImageConfig((64, 64), 2.0, 300.0)
    b = 6
    wxyz = O.quat_normalize(rng.standard_normal((b, 4)))
    off = rng.uniform(-5, 5, (b, 2))
    df = rng.uniform(8000, 20000, b)
    pose = cxs.QuaternionPose(off[:, 0], off[:, 1], wxyz)
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(df, 100.0, 10.0))
    imgs = cxs.make_image_model(volume, config, pose, tt).simulate().cpu().numpy()
    ref = O.simulate(g, wxyz, off, 2.0, defocus=df, astigmatism=100.0, astigmatism_angle=10.0)
    assert imgs.shape == (b, 64, 64)
    assert np.all(rel_l2(imgs, ref) <= 1e-5)


def test_projection_model_and_fourier_output(volume64):
    vol, g = volume64
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    config = cxs.BasicImageConfig((64, 64), 2.0, 300.0)
    pose = cxs.EulerAnglePose(phi_angle=10.0, theta_angle=20.0, psi_angle=30.0)
    model = cxs.make_image_model(volume, config, pose)  # no transfer theory
    assert isinstance(model, cxs.ProjectionImageModel)
    q = O.euler_to_quat(10.0, 20.0, 30.0)
    ref = O.simulate(g, q, np.zeros(2), 2.0, applies_ctf=False)
    assert rel_l2(model.simulate().cpu().numpy(), ref) <= 1e-5
    four = model.simulate(outputs_real_space=False)
    ref_f = O.compute_fourier_image(g, q, np.zeros(2), 2.0, applies_ctf=False)
    assert rel_l2(four.cpu().numpy(), ref_f) <= 1e-5
    # postprocess(compute_fourier_image()) == simulate()
    np.testing.assert_allclose(model.postprocess(four).cpu().numpy(), model.simulate().cpu().numpy(),
                               rtol=0, atol=1e-5 * np.abs(ref).max())


def test_integrator_outputs_integral(volume64):
    vol, g = volume64
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    config = cxs.BasicImageConfig((64, 64), 1.5, 300.0)
    pose = cxs.EulerAnglePose(phi_angle=5.0, theta_angle=70.0)
    proj = cxs.FourierSliceExtraction().integrate(volume.rotate_to_pose(pose), config)
    ref = O.compute_fourier_image(g, O.euler_to_quat(5.0, 70.0, 0.0), np.zeros(2), 1.5, applies_ctf=False,
                                  applies_translation=False, outputs_integral=True)
    assert rel_l2(proj.cpu().numpy(), ref) <= 1e-5


def test_api_errors(volume64):
    vol, _ = volume64
    with pytest.raises(ValueError):
        cxs.BasicImageConfig((64, 64), pixel_size=0.0, voltage_in_kilovolts=300.0)
    with pytest.raises(ValueError):
        cxs.ContrastTransferTheory(cxs.AstigmaticCTF(), amplitude_contrast_ratio=1.5)
    with pytest.raises(ValueError):
        cxs.AstigmaticCTF(spherical_aberration_in_mm=-1.0)
    volume = cxs.FourierVoxelGridVolume(torch.zeros((8, 8, 6), dtype=torch.complex64))
    with pytest.raises(AttributeError):
        cxs.FourierSliceExtraction().integrate(volume, cxs.BasicImageConfig((8, 8), 1.0, 300.0))
    with pytest.raises(ValueError):
        cxs.make_image_model(vol, cxs.BasicImageConfig((8, 8), 1.0, 300.0), cxs.EulerAnglePose())


def _padded_problem(n_vol, pad_scale, b, seed):
    rng = np.random.default_rng(seed)
    real = make_volume(n_vol, 4 * n_vol, n_vol // 4, seed)
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(real, pad_scale=pad_scale)
    g = O.fourier_grid_from_real(real, pad_scale=pad_scale)
    wxyz = O.quat_normalize(rng.standard_normal((b, 4)))
    off = rng.uniform(-5, 5, (b, 2))
    df = rng.uniform(8000, 20000, b)
    ast = rng.uniform(-200, 200, b)
    ang = rng.uniform(0, 180, b)
    pose = cxs.QuaternionPose(off[:, 0], off[:, 1], wxyz)
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(df, ast, ang), amplitude_contrast_ratio=0.1)
    ref_kw = dict(defocus=df, astigmatism=ast, astigmatism_angle=ang)
    return volume, g, wxyz, off, tt, pose, ref_kw


@pytest.mark.parametrize("wrap", [True, False])
def test_padded_volume_cropped_normalised_images(wrap):
    """The reference's published setup, shape for shape: an 80^3 volume with pad_scale=2 (a 160^3
    grid), 80x80 images padded to the grid's 160x160, normalizes_signal=True; against the oracle's
    simulate with the same shape and padded_shape, per image."""
    volume, g, wxyz, off, tt, pose, kw = _padded_problem(80, 2, 8, 31)
    assert volume.shape == (160, 160, 160)
    config = cxs.BasicImageConfig((80, 80), 1.2, 300.0, pad_options=dict(shape=volume.shape[0:2]))
    integ = cxs.FourierSliceExtraction(outputs_integral=False, wrap_negative_indices=wrap)
    model = cxs.make_image_model(volume, config, pose, tt, integ, normalizes_signal=True)
    img = model.simulate().cpu().numpy()
    ref = O.simulate(g, wxyz, off, 1.2, normalizes_signal=True, shape=(80, 80), padded_shape=(160, 160),
                     wrap_negative_indices=wrap, **kw)
    err = rel_l2(img, ref)
    print(f"80^3 pad 2 -> 160^3, 80^2 images in 160^2, normalised, wrap={wrap}: rel L2 max {err.max():.2e}")
    assert img.shape == (8, 80, 80) and np.all(err <= 1e-5), err
    # Fourier output of the cropped, normalised image; the padded image; postprocess of the
    # padded spectrum
    four = model.simulate(outputs_real_space=False).cpu().numpy()
    ref_f = O.simulate(g, wxyz, off, 1.2, normalizes_signal=True, shape=(80, 80), padded_shape=(160, 160),
                       outputs_real_space=False, wrap_negative_indices=wrap, **kw)
    assert four.shape == (8, 80, 41) and np.all(rel_l2(four, ref_f) <= 1e-5)
    big = model.simulate(removes_padding=False).cpu().numpy()
    ref_b = O.simulate(g, wxyz, off, 1.2, shape=(80, 80), padded_shape=(160, 160), removes_padding=False,
                       wrap_negative_indices=wrap, **kw)
    assert big.shape == (8, 160, 160) and np.all(rel_l2(big, ref_b) <= 1e-5)
    post = model.postprocess(model.compute_fourier_image()).cpu().numpy()
    assert np.all(rel_l2(post, ref) <= 1e-5)


@pytest.mark.parametrize("shape,padded", [((64, 64), (96, 96)), ((48, 48), (48, 48)), ((40, 40), (80, 80)),
                                          ((64, 64), (64, 64))])
def test_integrator_resize_branch(shape, padded):
    """padded_shape different from the volume's 64^3 box: the integrator resizes the slice in real
    space (irfftn, crop or zero pad, rfftn) before the CTF and shift at padded_shape; then the crop
    to shape. outputs_integral scales after the resize."""
    volume, g, wxyz, off, tt, pose, kw = _padded_problem(64, 1, 6, 41)
    config = cxs.BasicImageConfig(shape, 1.5, 300.0, pad_options=dict(shape=padded))
    for integral in (False, True):
        model = cxs.make_image_model(volume, config, pose, tt,
                                     cxs.FourierSliceExtraction(outputs_integral=integral))
        img = model.simulate().cpu().numpy()
        ref = O.simulate(g, wxyz, off, 1.5, shape=shape, padded_shape=padded, outputs_integral=integral, **kw)
        err = rel_l2(img, ref)
        print(f"resize branch box 64 shape {shape} padded {padded} integral={integral}: rel L2 max {err.max():.2e}")
        assert img.shape == (6,) + tuple(shape) and np.all(err <= 1e-5), err
    spec = cxs.FourierSliceExtraction().integrate(volume.rotate_to_pose(cxs.QuaternionPose(0.0, 0.0, wxyz[0])),
                                                  config)
    ref_s = O.compute_fourier_image(g, wxyz[0], np.zeros(2), 1.5, applies_ctf=False, applies_translation=False,
                                    outputs_integral=True, padded_shape=padded)
    assert rel_l2(spec.cpu().numpy(), ref_s) <= 1e-5


def test_noise_model_on_padded_config():
    volume, g, wxyz, off, tt, pose, kw = _padded_problem(40, 2, 3, 51)
    config = cxs.BasicImageConfig((40, 40), 1.0, 300.0, pad_options=dict(shape=(80, 80)))
    model = cxs.make_image_model(volume, config, pose, tt, normalizes_signal=True)
    nm = cxs.GaussianWhiteNoiseModel(model, variance=1.0, signal_scale_factor=0.5)
    noise = nm.compute_noise(7)
    assert noise.shape == (3, 40, 40)
    x = noise.cpu().numpy()
    np.testing.assert_allclose(x.mean(axis=(1, 2)), 0.0, atol=1e-5)
    np.testing.assert_allclose(x.std(axis=(1, 2)), 1.0, atol=1e-5)
    assert nm.sample(8).shape == (3, 40, 40)


@pytest.mark.parametrize("shape,padded,norm", [((64, 64), (64, 64), False), ((64, 64), (64, 64), True),
                                               ((48, 48), (64, 64), True), ((40, 40), (80, 80), False)])
def test_envelope_filter_and_mask(shape, padded, norm):
    """ContrastTransferTheory(envelope=FourierGaussian) times the contrast spectrum
    (transfer_theory.py:117-118); simulate(filter=..., mask=...): the filter on the padded rfft
    image before the inverse transform, the mask on the cropped, normalised image
    (base_image_model.py:119-158); in both output spaces, against the oracle's same steps."""
    from cryojax_amd import transforms as T

    volume, g, wxyz, off, _, pose, kw = _padded_problem(64, 1, 5, 61)
    rng = np.random.default_rng(62)
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(kw["defocus"], kw["astigmatism"], kw["astigmatism_angle"]),
                                    envelope=T.FourierGaussian(amplitude=0.8, b_factor=60.0),
                                    amplitude_contrast_ratio=0.1)
    config = cxs.BasicImageConfig(shape, 1.5, 300.0, pad_options=dict(shape=padded))
    model = cxs.make_image_model(volume, config, pose, tt, normalizes_signal=norm)
    filt = rng.uniform(0.2, 1.5, (padded[0], padded[1] // 2 + 1))
    mask = (rng.uniform(0, 1, shape) > 0.3).astype(np.float64) * rng.uniform(0.5, 1.0, shape)
    env = lambda f: O.fourier_gaussian(f, 0.8, 60.0)  # noqa: E731
    common = dict(shape=shape, padded_shape=padded, normalizes_signal=norm, envelope=env, **kw)
    print(f"case: shape {shape} padded {padded} normalizes_signal={norm}", flush=True)
    for real in (True, False):
        for f, m in ((None, None), (filt, None), (None, mask), (filt, mask)):
            got = model.simulate(outputs_real_space=real, filter=None if f is None else T.CustomFilter(f),
                                 mask=None if m is None else T.CustomMask(m)).cpu().numpy()
            ref = O.simulate(g, wxyz, off, 1.5, outputs_real_space=real, fourier_filter=f, mask=m, **common)
            err = rel_l2(got, ref)
            print(f"  real={real} filter={f is not None} mask={m is not None}: rel L2 max {err.max():.2e}")
            assert got.shape == ref.shape and np.all(err <= 1e-5), err
    # postprocess of compute_fourier_image with the same transforms equals simulate
    post = model.postprocess(model.compute_fourier_image(), filter=T.CustomFilter(filt), mask=T.CustomMask(mask))
    ref = O.simulate(g, wxyz, off, 1.5, fourier_filter=filt, mask=mask, **common)
    assert np.all(rel_l2(post.cpu().numpy(), ref) <= 1e-5)


def test_filter_shape_error():
    from cryojax_amd import transforms as T

    volume, _, _, _, tt, pose, _ = _padded_problem(32, 1, 2, 71)
    config = cxs.BasicImageConfig((32, 32), 1.5, 300.0, pad_options=dict(shape=(48, 48)))
    model = cxs.make_image_model(volume, config, pose, tt)
    with pytest.raises(ValueError, match="padded_shape"):
        model.simulate(filter=T.CustomFilter(np.ones((32, 17))))
    with pytest.raises(NotImplementedError):
        model.simulate(mask=lambda x: x)


def test_builtin_filter_and_mask_in_simulate():
    """The reference's built-in transforms (LowpassFilter on the padded frequency grid,
    CircularCosineMask on the image's coordinates) through simulate, against the oracle with the
    same arrays (_filters.py:58-96, _masks.py:49-81)."""
    from cryojax_amd import coordinates as C
    from cryojax_amd import transforms as T

    volume, g, wxyz, off, _, pose, kw = _padded_problem(64, 1, 4, 63)
    config = cxs.BasicImageConfig((56, 56), 1.5, 300.0, pad_options=dict(shape=(64, 64)))
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(kw["defocus"], kw["astigmatism"], kw["astigmatism_angle"]))
    model = cxs.make_image_model(volume, config, pose, tt, normalizes_signal=True)
    lp = T.LowpassFilter(config.padded_frequency_grid_in_angstroms, 1.5, 0.6, 0.1)
    mk = T.CircularCosineMask(C.make_coordinate_grid((56, 56), 1.5), 30.0, 6.0)
    got = model.simulate(filter=lp, mask=mk).cpu().numpy()
    ref = O.simulate(g, wxyz, off, 1.5, shape=(56, 56), padded_shape=(64, 64), normalizes_signal=True,
                     fourier_filter=lp.get(), mask=mk.get(), **kw)
    err = rel_l2(got, ref)
    print(f"[api] LowpassFilter + CircularCosineMask: rel L2 max {err.max():.2e}", flush=True)
    assert np.all(err <= 1e-5)
    # calling a transform returns a new tensor and leaves its input alone
    x = torch.ones((2, 56, 56), device="cuda")
    y = mk(x)
    assert torch.all(x == 1) and np.allclose(y.cpu().numpy(), np.broadcast_to(mk.get(), (2, 56, 56)), rtol=1e-7)


@pytest.mark.parametrize("shape", [None, (40, 40), (47, 47)])
def test_whitening_filter_of_deltas(shape):
    """Images that are deltas (any position: a shift is a phase) have flat power spectra
    |F|^2 / P = a^2 / P; a stack of deltas of heights a_i whitens to sqrt(P / mean(a_i^2)) away
    from the zero frequency and the even Nyquist corner (_filters.py:140-297). The transforms run
    on the device (rfft2 in FP64)."""
    from cryojax_amd import transforms as T

    h = w = 44
    amps = np.array([1.0, 2.0, 0.5])
    stack = np.zeros((3, h, w))
    for i, (a, (y, x)) in enumerate(zip(amps, [(22, 22), (3, 40), (30, 7)])):
        stack[i, y, x] = a
    f = T.WhiteningFilter(torch.as_tensor(stack, device="cuda"), shape=shape).get()
    sh = (h, w) if shape is None else shape
    want = np.full((sh[0], sh[1] // 2 + 1), np.sqrt(h * w / np.mean(amps**2)))
    want[0, 0] = 0.0
    if sh[0] % 2 == 0 and sh[1] % 2 == 0:
        want[sh[0] // 2, sh[1] // 2] = 0.0
    print(f"[api] WhiteningFilter shape={shape}: max rel err {np.max(np.abs(f - want) / want.max()):.2e}", flush=True)
    np.testing.assert_allclose(f, want, rtol=1e-10, atol=1e-10 * want.max())
    sq = T.WhiteningFilter(stack[0], outputs_squared=True).get()  # a host image goes to the device
    assert sq[1, 1] == pytest.approx(h * w / 1.0, rel=1e-10)
