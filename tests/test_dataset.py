"""RELION parameter files and particle stacks on the CPU: reading the
reference's STAR / MRCS fixtures, the STAR <-> parameter conventions, image
naming, and simulate_particle_stack with a host-side image function."""

import os
import shutil

import numpy as np
import pytest

from cryojax_amd import dataset as D
from cryojax_amd import io
from cryojax_amd import simulator as cxs
from cryojax_amd import transforms as T

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture()
def project(tmp_path):
    shutil.copy(os.path.join(GOLDEN, "relion_particles.star"), tmp_path / "particles.star")
    shutil.copy(os.path.join(GOLDEN, "relion_000000.mrcs"), tmp_path / "000000.mrcs")
    return tmp_path


def test_read_parameters_follow_relion_conventions(project):
    pf = D.RelionParticleParameterFile(project / "particles.star")
    assert len(pf) == 3
    p = pf[0]
    pose, ctf, cfg = p["pose"], p["transfer_theory"].ctf, p["image_config"]
    assert float(pose.phi_angle) == pytest.approx(-50.195320)
    assert float(pose.theta_angle) == pytest.approx(70.588104)
    assert float(pose.psi_angle) == pytest.approx(81.394234)
    np.testing.assert_allclose(pose.offset_in_angstroms, [-32.056030, 21.564449])
    assert float(ctf.defocus_in_angstroms) == pytest.approx((10061.054688 + 10045.544922) / 2)
    assert float(ctf.astigmatism_in_angstroms) == pytest.approx(10061.054688 - 10045.544922)
    assert float(ctf.astigmatism_angle) == pytest.approx(1.324032)
    assert float(ctf.spherical_aberration_in_mm) == pytest.approx(2.7)
    assert float(p["transfer_theory"].amplitude_contrast_ratio) == pytest.approx(0.1)
    assert cfg.shape == (16, 16) and float(cfg.pixel_size) == 12.0 and float(cfg.voltage_in_kilovolts) == 300.0
    batch = pf[np.array([2, 0])]
    np.testing.assert_allclose(batch["pose"].psi_angle, [161.981888, 81.394234])
    assert batch["image_config"].pixel_size.shape == (2,)
    sel = D.RelionParticleParameterFile(project / "particles.star",
                                        selection_filter={"rlnDefocusU": lambda x: x > 11000})
    assert len(sel) == 2
    with pytest.raises(IndexError):
        pf[3]


def test_load_envelope_params(project):
    """The fixture's rlnCtfBfactor (1.0) and rlnCtfScalefactor (2.0) become a FourierGaussian
    per particle (relion.py:1016-1029, 1156-1176); without loads_envelope there is none."""
    pf = D.RelionParticleParameterFile(project / "particles.star", loads_envelope=True)
    assert pf.loads_envelope is True
    one = pf[0]["transfer_theory"].envelope
    assert isinstance(one, T.FourierGaussian)
    assert float(one.b_factor) == 1.0 and float(one.amplitude) == 2.0
    env = pf[np.arange(3)]["transfer_theory"].envelope
    data = pf.starfile_data["particles"]
    np.testing.assert_allclose(env.b_factor, data["rlnCtfBfactor"], rtol=1e-12)
    np.testing.assert_allclose(env.amplitude, data["rlnCtfScalefactor"], rtol=1e-12)
    assert D.RelionParticleParameterFile(project / "particles.star")[0]["transfer_theory"].envelope is None


def test_envelope_round_trips_and_warns(tmp_path, rng):
    prm = make_parameters(4, rng)
    b, amp = rng.uniform(0, 80, 4), rng.uniform(0.5, 2, 4)
    b[0] = 0.0  # RELION's "no B-factor", kept as the reference keeps it
    for env in (T.FourierGaussian(1.0, 1.0), T.Constant(amp)):
        if isinstance(env, T.FourierGaussian):
            env.amplitude, env.b_factor = amp, b
        prm["transfer_theory"].envelope = env
        path = tmp_path / f"{type(env).__name__}.star"
        pf = D.RelionParticleParameterFile(path, mode="w")
        pf.append(prm)
        pf.save()
        back = D.RelionParticleParameterFile(path, loads_envelope=True)[np.arange(4)]["transfer_theory"].envelope
        assert type(back) is type(env)
        if isinstance(env, T.FourierGaussian):
            np.testing.assert_allclose(back.b_factor, b, rtol=1e-5, atol=1e-6)  # STAR keeps 6 decimals
            np.testing.assert_allclose(back.amplitude, amp, rtol=1e-5)
        else:
            np.testing.assert_allclose(back.value, amp, rtol=1e-5)
    prm["transfer_theory"].envelope = T.Lorenzian()
    with pytest.raises(ValueError):
        D.RelionParticleParameterFile(tmp_path / "bad.star", mode="w").append(prm)
    prm["transfer_theory"].envelope = None
    pf = D.RelionParticleParameterFile(tmp_path / "none.star", mode="w")
    pf.append(prm)
    pf.save()
    with pytest.warns(UserWarning, match="no envelope parameters"):
        got = D.RelionParticleParameterFile(tmp_path / "none.star", loads_envelope=True)[0]
    assert got["transfer_theory"].envelope is None


def test_read_stack_images(project):
    pf = D.RelionParticleParameterFile(project / "particles.star")
    ds = D.RelionParticleStackDataset(pf, project)
    _, ref = io.read_mrc(project / "000000.mrcs")
    got = ds[1]
    np.testing.assert_array_equal(got["images"], ref[1].astype(np.float64))
    assert float(got["parameters"]["pose"].phi_angle) == pytest.approx(-125.329025)
    np.testing.assert_array_equal(ds[0:3]["images"], ref.astype(np.float64))
    ds.loads_parameters = False
    assert ds[2]["parameters"] is None


def make_parameters(n, rng, size=16, pixel_size=2.0):
    pose = cxs.EulerAnglePose(rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), rng.uniform(-180, 180, n),
                              rng.uniform(0, 180, n), rng.uniform(-180, 180, n))
    ctf = cxs.AstigmaticCTF(rng.uniform(8000, 12000, n), rng.uniform(-200, 200, n), rng.uniform(0, 180, n), 2.7)
    theory = cxs.ContrastTransferTheory(ctf, amplitude_contrast_ratio=0.07, phase_shift=np.zeros(n))
    config = cxs.BasicImageConfig((size, size), pixel_size, 300.0)
    return dict(image_config=config, pose=pose, transfer_theory=theory)


def test_write_then_read_parameters_and_images(tmp_path, rng):
    pf = D.RelionParticleParameterFile(tmp_path / "out.star", mode="w")
    ds = D.RelionParticleStackDataset(pf, tmp_path, mode="w",
                                      mrcfile_settings=dict(prefix="f", output_folder="stacks"))
    prm = make_parameters(5, rng)
    images = rng.standard_normal((5, 16, 16)).astype(np.float32)
    ds.append(dict(parameters=prm, images=images))
    prm2 = make_parameters(2, rng)
    images2 = rng.standard_normal((2, 16, 16)).astype(np.float32)
    ds.append(dict(parameters=prm2, images=images2))
    pf.save()
    names = pf.starfile_data["particles"]["rlnImageName"].tolist()
    assert names[:2] == ["000001@stacks/f_000000.mrcs", "000002@stacks/f_000000.mrcs"]
    assert names[5:] == ["000001@stacks/f_000001.mrcs", "000002@stacks/f_000001.mrcs"]
    assert pf.starfile_data["optics"]["rlnOpticsGroup"].tolist() == [1, 2]
    back = D.RelionParticleParameterFile(tmp_path / "out.star")
    got = back[np.arange(5)]
    np.testing.assert_allclose(got["pose"].phi_angle, prm["pose"].phi_angle, atol=1e-5)
    np.testing.assert_allclose(got["pose"].offset_in_angstroms, prm["pose"].offset_in_angstroms, atol=1e-5)
    np.testing.assert_allclose(got["transfer_theory"].ctf.defocus_in_angstroms,
                               prm["transfer_theory"].ctf.defocus_in_angstroms, atol=1e-5)
    np.testing.assert_allclose(got["transfer_theory"].ctf.astigmatism_in_angstroms,
                               prm["transfer_theory"].ctf.astigmatism_in_angstroms, atol=2e-5)
    assert float(got["transfer_theory"].amplitude_contrast_ratio[0]) == pytest.approx(0.07)
    read = D.RelionParticleStackDataset(back, tmp_path)
    np.testing.assert_array_equal(read[np.arange(5)]["images"], images.astype(np.float64))
    np.testing.assert_array_equal(read[np.arange(5, 7)]["images"], images2.astype(np.float64))
    with pytest.raises(NotImplementedError):  # one optics group per read, as the reference
        read[np.arange(7)]
    with pytest.raises(FileExistsError):
        pf.save()
    with pytest.raises(FileExistsError):
        D.RelionParticleParameterFile(tmp_path / "out.star", mode="w")


def test_inverted_rotation_round_trips(tmp_path, rng):
    pf = D.RelionParticleParameterFile(tmp_path / "inv.star", mode="w", inverts_rotation=True)
    prm = make_parameters(3, rng)
    pf.append(prm)
    pf.save()
    back = D.RelionParticleParameterFile(tmp_path / "inv.star", inverts_rotation=True)[np.arange(3)]
    for a in ("phi_angle", "theta_angle", "psi_angle"):
        d = (getattr(back["pose"], a) - getattr(prm["pose"], a) + 180) % 360 - 180
        np.testing.assert_allclose(d, 0, atol=1e-4)


def fake_image(parameters, constant_args, per_particle):
    """A host-side stand-in for an image model: the image encodes its particle's defocus and scale."""
    d = np.atleast_1d(parameters["transfer_theory"].ctf.defocus_in_angstroms)
    s = np.atleast_1d(per_particle)
    img = (d * s)[:, None, None] * np.ones((1, 16, 16)) + constant_args
    return img if np.ndim(parameters["transfer_theory"].ctf.defocus_in_angstroms) else img[0]


@pytest.mark.parametrize("batch_size,per_file", [(None, None), (2, 3), (4, 7)])
def test_simulate_particle_stack(tmp_path, rng, batch_size, per_file):
    pf = D.RelionParticleParameterFile(tmp_path / "sim.star", mode="w")
    pf.append(make_parameters(7, rng))
    ds = D.RelionParticleStackDataset(pf, tmp_path, mode="w")
    scale = rng.uniform(1, 2, 7)
    D.simulate_particle_stack(ds, fake_image, constant_args=1.0, per_particle_args=scale, batch_size=batch_size,
                              images_per_file=per_file)
    back = D.RelionParticleParameterFile(tmp_path / "sim.star")
    imgs = D.RelionParticleStackDataset(back, tmp_path)[np.arange(7)]["images"]
    defocus = back[np.arange(7)]["transfer_theory"].ctf.defocus_in_angstroms
    np.testing.assert_allclose(imgs[:, 0, 0], (defocus * scale + 1.0).astype(np.float32), rtol=1e-6)
    n_files = len({n.split("@")[1] for n in back.starfile_data["particles"]["rlnImageName"]})
    assert n_files == (1 if per_file is None else -(-7 // per_file))
    with pytest.raises(ValueError):
        D.simulate_particle_stack(D.RelionParticleStackDataset(back, tmp_path), fake_image)
