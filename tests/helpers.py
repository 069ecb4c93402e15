"""Shared builders for parity tests: one set of host parameters feeds both the
oracle (float64) and the HIP path (through the C ABI)."""

import numpy as np

from cryojax_amd import _lib
from cryojax_amd.rotations import SO3
from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume
from cryojax_amd.simulator import wavelength_from_kilovolts


def make_volume(n, n_gauss, radius, seed):
    return gaussian_mixture_volume(n, n_gauss, radius, seed)


def fourier_grid(vol):
    """complex128 centred grid (the oracle's and, rounded once, the device's)."""
    x = np.asarray(vol, dtype=np.float64)
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x)))


def draw(n_images, seed, offset_max=10.0, defocus=(8000.0, 20000.0), astig=(0.0, 800.0)):
    return draw_image_parameters(n_images, seed, offset_max=offset_max, defocus=defocus, astigmatism=astig)


def packed_params(d, pixel_size, voltage=300.0, ac=0.1, cs=2.7, amplitude=1.0):
    n = d["wxyz"].shape[0]
    mats = SO3(d["wxyz"]).normalize().sandwich_matrix()
    return _lib.pack_params(
        mats.reshape(n, 9),
        d["offset"],
        d["defocus"],
        d["astigmatism"],
        d["astigmatism_angle"],
        cs,
        wavelength_from_kilovolts(voltage),
        pixel_size,
        np.arctan(ac / np.sqrt(1 - ac * ac)),
        amplitude,
        n,
    )


def oracle_images(grid128, d, pixel_size, voltage=300.0, ac=0.1, cs=2.7, fourier=False, **kw):
    from oracle import imaging as O

    q = O.quat_normalize(d["wxyz"])
    args = dict(
        voltage_kv=voltage,
        defocus=d["defocus"],
        astigmatism=d["astigmatism"],
        astigmatism_angle=d["astigmatism_angle"],
        cs_mm=cs,
        amplitude_contrast_ratio=ac,
    )
    args.update(kw)
    if fourier:
        return O.compute_fourier_image(grid128, q, d["offset"], pixel_size, **args)
    return O.simulate(grid128, q, d["offset"], pixel_size, **args)


def rel_l2(a, b):
    """Per-image relative L2 error ||a - b|| / ||b|| over the last two axes."""
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b)
    num = np.sqrt(np.sum(np.abs(a - b) ** 2, axis=(-2, -1)))
    den = np.sqrt(np.sum(np.abs(b) ** 2, axis=(-2, -1)))
    return num / den


def convention_report(label, gpu, ref, ref_other):
    """Per-image GPU error against the oracle of the same out-of-range rule, and the
    rule's sensitivity |oracle_wrap - oracle_fill| / |oracle| on the same images.
    Prints both per image (the stated error per result) and returns the errors."""
    err = rel_l2(gpu, ref)
    sens = rel_l2(ref_other, ref)
    print(f"{label}: images={err.size} gpu rel L2 max {err.max():.2e} median {np.median(err):.2e}; "
          f"wrap-vs-fill sensitivity max {sens.max():.2e} median {np.median(sens):.2e}")
    print(f"{label}: per image gpu rel L2 {np.array2string(err, precision=2, max_line_width=200)}")
    print(f"{label}: per image sensitivity {np.array2string(sens, precision=2, max_line_width=200)}", flush=True)
    return err, sens
