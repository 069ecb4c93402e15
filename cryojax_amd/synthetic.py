"""Seeded synthetic workloads: Gaussian-mixture volumes and per-image parameter draws.

These generate benchmark and test inputs on the host (NumPy PCG64); they are
not part of the image-formation path itself.
"""

from __future__ import annotations

import math

import numpy as np


def _centred_coordinates(n):
    # Real-space voxel coordinates with index n // 2 at zero
    # (cryojax/coordinates/_make_coordinate_grids.py:308-309).
    return np.fft.fftshift(np.fft.fftfreq(n, 1.0)) * n


def gaussian_mixture_volume(n, n_gaussians, radius, seed, sigma=(1.5, 3.0), amp=(0.5, 1.5)):
    """Sum of isotropic 3D Gaussians on an (n, n, n) grid (float64 accumulate, float32 out).

    Centres are uniform in a ball of `radius` voxels around the grid centre,
    sigma ~ U(sigma), amplitude ~ U(amp). Each Gaussian is evaluated in a local
    box of +-7 sigma (truncation below exp(-24.5) ~ 2e-11 of its peak, so the
    volume is smooth to that level).
    """
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((n_gaussians, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = radius * rng.uniform(0, 1, n_gaussians) ** (1.0 / 3.0)
    centres = u * r[:, None]
    sig = rng.uniform(sigma[0], sigma[1], n_gaussians)
    amps = rng.uniform(amp[0], amp[1], n_gaussians)
    vol = np.zeros((n, n, n))
    c1 = _centred_coordinates(n)
    for (cx, cy, cz), s, a in zip(centres, sig, amps):
        h = int(math.ceil(7 * s)) + 1
        sl = []
        for cc in (cz, cy, cx):
            i0 = int(math.floor(cc)) + n // 2
            sl.append((max(i0 - h, 0), min(i0 + h + 1, n)))
        z = c1[sl[0][0] : sl[0][1]] - cz
        y = c1[sl[1][0] : sl[1][1]] - cy
        x = c1[sl[2][0] : sl[2][1]] - cx
        g = (
            np.exp(-(z**2) / (2 * s * s))[:, None, None]
            * np.exp(-(y**2) / (2 * s * s))[None, :, None]
            * np.exp(-(x**2) / (2 * s * s))[None, None, :]
        )
        vol[sl[0][0] : sl[0][1], sl[1][0] : sl[1][1], sl[2][0] : sl[2][1]] += a * g
    return vol.astype(np.float32)


def uniform_quaternions(n, rng):
    """Uniform draws on SO(3), same map as cryojax/rotations/_lie_group.py:319-341."""
    u1 = rng.uniform(0.0, 1.0, n)
    u2 = rng.uniform(0.0, 2 * np.pi, n)
    u3 = rng.uniform(0.0, 2 * np.pi, n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(u2), a * np.cos(u2), b * np.sin(u3), b * np.cos(u3)], -1)


def draw_image_parameters(
    n_images,
    seed,
    offset_max=10.0,
    defocus=(8000.0, 20000.0),
    astigmatism=(0.0, 800.0),
    astigmatism_angle=(0.0, 180.0),
):
    """Per-image pose and CTF draws used by the batched workloads.

    Returns a dict of float64 arrays: wxyz (n, 4), offset (n, 2) in Angstroms,
    defocus, astigmatism (Angstroms) and astigmatism_angle (degrees).
    """
    rng = np.random.default_rng(seed)
    return {
        "wxyz": uniform_quaternions(n_images, rng),
        "offset": rng.uniform(-offset_max, offset_max, (n_images, 2)),
        "defocus": rng.uniform(defocus[0], defocus[1], n_images),
        "astigmatism": rng.uniform(astigmatism[0], astigmatism[1], n_images),
        "astigmatism_angle": rng.uniform(astigmatism_angle[0], astigmatism_angle[1], n_images),
    }


def pack_drawn_parameters(d, pixel_size, voltage_in_kilovolts=300.0, amplitude_contrast_ratio=0.1,
                          spherical_aberration_in_mm=2.7, amplitude=1.0):
    """cryomi_image_params table (n, 20) float64 for the draws of draw_image_parameters."""
    from . import _lib
    from .rotations import SO3
    from .simulator import wavelength_from_kilovolts

    n = d["wxyz"].shape[0]
    ac = amplitude_contrast_ratio
    mats = SO3(d["wxyz"]).normalize().sandwich_matrix()
    return _lib.pack_params(mats.reshape(n, 9), d["offset"], d["defocus"], d["astigmatism"], d["astigmatism_angle"],
                            spherical_aberration_in_mm, wavelength_from_kilovolts(voltage_in_kilovolts), pixel_size,
                            math.atan(ac / math.sqrt(1 - ac * ac)), amplitude, n)
