"""Noise model and normalisation on device against the FP64 restatement
(oracle/noise.py), and through the cryoJAX-style API."""

import numpy as np
import pytest
import torch

import cryojax_amd.simulator as cxs
from cryojax_amd import ops
from oracle import imaging as O
from oracle import noise as N
from tests.helpers import fourier_grid, make_volume, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_normalize_images_matches_oracle(rng):
    x = (rng.standard_normal((5, 48, 48)) * 3 + 2).astype(np.float32)
    region = np.zeros((48, 48), bool)
    region[10:30, 5:40] = True
    for where in (None, region):
        t = torch.from_numpy(x.copy()).to(DEV)
        ops.normalize_images(t, where)
        want = np.stack([N.normalize_image(xi, where) for xi in x])
        assert np.max(np.abs(t.cpu().numpy() - want)) <= 1e-5


def test_white_noise_spectrum_and_image():
    n, box, var = 400, 64, 2.0
    spec = ops.white_noise_spectrum(n, box, var, 1234, DEV)
    s = spec.cpu().numpy()
    assert np.all(s.imag == 0) and np.all(s[:, 0, 0] == 0)
    z = s.real / np.sqrt(box * box * var)
    vals = z.reshape(n, -1)[:, 1:].ravel()
    assert abs(vals.mean()) < 0.01 and abs(vals.std() - 1) < 0.01
    assert abs(np.mean(vals**4) - 3) < 0.05
    again = ops.white_noise_spectrum(n, box, var, 1234, DEV)
    other = ops.white_noise_spectrum(n, box, var, 1235, DEV)
    assert torch.equal(spec, again) and not torch.equal(spec, other)
    img = ops.irfft2(spec, box).cpu().numpy()
    want = N.noise_image_from_draws(z, var)
    assert np.max(np.abs(img - want)) <= 1e-5 * np.abs(want).max()
    assert abs(img.var(axis=0).mean() - var) < 0.1


def test_noisy_images_and_log_likelihood(rng):
    sig = rng.standard_normal((3, 32, 32)).astype(np.float32)
    nz = rng.standard_normal((3, 32, 32)).astype(np.float32)
    s, z = torch.from_numpy(sig).to(DEV), torch.from_numpy(nz).to(DEV)
    out = ops.noisy_images(s, z, 1.7, -0.3).cpu().numpy()
    np.testing.assert_allclose(out, 1.7 * sig - 0.3 + nz, rtol=1e-6, atol=1e-6)
    ll = ops.gaussian_log_likelihood(s, torch.from_numpy(out).to(DEV), 0.8).cpu().numpy()
    np.testing.assert_allclose(ll, N.log_likelihood(sig, out, 0.8), rtol=1e-10)


@pytest.fixture(scope="module")
def model():
    vol = make_volume(64, 200, 20, 7)
    g = fourier_grid(vol)
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    config = cxs.BasicImageConfig((64, 64), 2.0, 300.0)
    b = 4
    r = np.random.default_rng(3)
    wxyz = O.quat_normalize(r.standard_normal((b, 4)))
    off = r.uniform(-4, 4, (b, 2))
    df = r.uniform(9000, 15000, b)
    pose = cxs.QuaternionPose(off[:, 0], off[:, 1], wxyz)
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(df, 100.0, 10.0))
    ref = O.simulate(g, wxyz, off, 2.0, defocus=df, astigmatism=100.0, astigmatism_angle=10.0)
    return volume, config, pose, tt, ref


def test_normalizes_signal_matches_oracle(model):
    volume, config, pose, tt, ref = model
    region = np.zeros((64, 64), bool)
    region[16:48, 16:48] = True
    for where in (None, region):
        m = cxs.make_image_model(volume, config, pose, tt, normalizes_signal=True, signal_region=where)
        img = m.simulate().cpu().numpy()
        want = np.stack([N.normalize_image(r, where) for r in ref])
        assert np.all(rel_l2(img, want) <= 1e-5)
    # Fourier output with normalizes_signal: rfftn of the normalised image (base_image_model.py:158)
    four = cxs.make_image_model(volume, config, pose, tt, normalizes_signal=True).simulate(outputs_real_space=False)
    want = np.stack([N.normalize_image(r, None) for r in ref])
    assert np.all(rel_l2(four.cpu().numpy(), O.rfftn(want)) <= 1e-5)


def test_gaussian_white_noise_model(model):
    volume, config, pose, tt, ref = model
    im = cxs.make_image_model(volume, config, pose, tt)
    nm = cxs.GaussianWhiteNoiseModel(im, variance=0.3, signal_scale_factor=2.0, signal_offset=0.5)
    sig = nm.compute_signal().cpu().numpy()
    assert np.all(rel_l2(sig, 2.0 * ref + 0.5) <= 1e-5)
    noise = nm.compute_noise(7)
    sample = nm.sample(7)
    torch.testing.assert_close(sample, nm.compute_signal() + noise, rtol=1e-6, atol=1e-6)
    assert not torch.equal(nm.compute_noise(8), noise)
    obs = sample.cpu().numpy()
    ll = nm.log_likelihood(torch.from_numpy(obs)).cpu().numpy()
    np.testing.assert_allclose(ll, N.log_likelihood(2.0 * ref + 0.5, obs, 0.3), rtol=1e-5)
    with pytest.raises(ValueError):
        cxs.GaussianWhiteNoiseModel(im, variance=0.0)
