"""simulate_particle_stack with the device image model: the MRCS files it
writes hold the fused forward's images (checked against the FP64 oracle), and
the STAR file holds their parameters."""

import numpy as np
import pytest

import cryojax_amd.simulator as cxs
from cryojax_amd import dataset as D
from oracle import imaging as O
from tests.helpers import fourier_grid, make_volume, rel_l2

pytestmark = pytest.mark.gpu


def compute_image(parameters, volume, _):
    model = cxs.make_image_model(volume, parameters["image_config"], parameters["pose"],
                                 parameters["transfer_theory"])
    return model.simulate()


@pytest.mark.parametrize("batch_size", [None, 4])
def test_simulated_stack_on_disk_matches_oracle(tmp_path, rng, batch_size):
    n, box, ps = 9, 64, 2.0
    vol = make_volume(box, 150, 16, 5)
    g = fourier_grid(vol)
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    pose = cxs.EulerAnglePose(rng.uniform(-4, 4, n), rng.uniform(-4, 4, n), rng.uniform(-180, 180, n),
                              rng.uniform(0, 180, n), rng.uniform(-180, 180, n))
    ctf = cxs.AstigmaticCTF(rng.uniform(9000, 15000, n), rng.uniform(-300, 300, n), rng.uniform(0, 180, n), 2.7)
    theory = cxs.ContrastTransferTheory(ctf, amplitude_contrast_ratio=0.1, phase_shift=np.zeros(n))
    pf = D.RelionParticleParameterFile(tmp_path / "sim.star", mode="w")
    pf.append(dict(image_config=cxs.BasicImageConfig((box, box), ps, 300.0), pose=pose, transfer_theory=theory))
    ds = D.RelionParticleStackDataset(pf, tmp_path, mode="w", mrcfile_settings=dict(output_folder="stacks"))
    print(f"[dataset] simulate_particle_stack batch_size={batch_size}", flush=True)
    D.simulate_particle_stack(ds, compute_image, constant_args=volume, batch_size=batch_size, images_per_file=5)
    back = D.RelionParticleParameterFile(tmp_path / "sim.star")
    got = D.RelionParticleStackDataset(back, tmp_path)[np.arange(n)]
    assert got["images"].shape == (n, box, box)
    for i in (0, 4, 5, 8):
        q = O.euler_to_quat(pose.phi_angle[i], pose.theta_angle[i], pose.psi_angle[i])
        ref = O.simulate(g, q, pose.offset_in_angstroms[i], ps, defocus=ctf.defocus_in_angstroms[i],
                         astigmatism=ctf.astigmatism_in_angstroms[i], astigmatism_angle=ctf.astigmatism_angle[i])
        assert rel_l2(got["images"][i], ref) <= 1e-5, i
    # the parameters read back from the STAR file re-simulate the same images
    again = compute_image(back[np.arange(n)], volume, None).cpu().numpy()
    assert np.all(rel_l2(again, got["images"]) <= 1e-5)


@pytest.mark.parametrize("kind", ["gaussian", "constant"])
def test_stack_with_per_particle_envelopes_matches_oracle(tmp_path, rng, kind):
    """Envelope parameters written to the STAR file (rlnCtfBfactor, rlnCtfScalefactor) are read
    back per particle (loads_envelope=True) and each image carries its own FourierGaussian, in
    batches of the fused forward (relion.py:1016-1029, 1115-1118; transfer_theory.py:117-118)."""
    from cryojax_amd import transforms as T

    n, box, ps = 7, 64, 2.0
    vol = make_volume(box, 150, 16, 9)
    g = fourier_grid(vol)
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
    pose = cxs.EulerAnglePose(rng.uniform(-4, 4, n), rng.uniform(-4, 4, n), rng.uniform(-180, 180, n),
                              rng.uniform(0, 180, n), rng.uniform(-180, 180, n))
    ctf = cxs.AstigmaticCTF(rng.uniform(9000, 15000, n), rng.uniform(-300, 300, n), rng.uniform(0, 180, n), 2.7)
    b, amp = rng.uniform(20, 200, n), rng.uniform(0.5, 2.0, n)
    b[1] = 0.0
    if kind == "gaussian":
        env = T.FourierGaussian(1.0, 1.0)
        env.amplitude, env.b_factor = amp, b
    else:
        env = T.Constant(amp)
    theory = cxs.ContrastTransferTheory(ctf, envelope=env, amplitude_contrast_ratio=0.1, phase_shift=np.zeros(n))
    pf = D.RelionParticleParameterFile(tmp_path / "env.star", mode="w", loads_envelope=True)
    pf.append(dict(image_config=cxs.BasicImageConfig((box, box), ps, 300.0), pose=pose, transfer_theory=theory))
    ds = D.RelionParticleStackDataset(pf, tmp_path, mode="w", mrcfile_settings=dict(output_folder="stacks"))
    print(f"[dataset] simulate_particle_stack with {kind} envelopes, batch_size=3", flush=True)
    D.simulate_particle_stack(ds, compute_image, constant_args=volume, batch_size=3)
    back = D.RelionParticleParameterFile(tmp_path / "env.star", loads_envelope=True)
    got = D.RelionParticleStackDataset(back, tmp_path)[np.arange(n)]["images"]
    read = back[np.arange(n)]["transfer_theory"].envelope
    for i in range(n):
        q = O.euler_to_quat(pose.phi_angle[i], pose.theta_angle[i], pose.psi_angle[i])
        if kind == "gaussian":
            a_i, b_i = float(read.amplitude[i]), float(read.b_factor[i])
            fn = lambda f, a_i=a_i, b_i=b_i: O.fourier_gaussian(f, a_i, b_i)  # noqa: E731
        else:
            a_i, b_i = float(read.value[i]), 0.0
            fn = lambda f, a_i=a_i: a_i  # noqa: E731
        ref = O.simulate(g, q, pose.offset_in_angstroms[i], ps, defocus=ctf.defocus_in_angstroms[i],
                         astigmatism=ctf.astigmatism_in_angstroms[i], astigmatism_angle=ctf.astigmatism_angle[i],
                         envelope=fn)
        err = rel_l2(got[i], ref)
        print(f"  image {i}: amplitude {a_i:.3f} b_factor {b_i:.1f} rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5, i
