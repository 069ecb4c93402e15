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
