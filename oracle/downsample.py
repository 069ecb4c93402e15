"""FP64 restatement of Fourier-crop (Hartley) downsampling.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows cryojax/ndimage/_downsample.py:91-155 (fourier_crop_downsample:
new shape int(size / factor) per axis), :158-199
(fourier_crop_downsample_to_shape: real input takes the Hartley branch,
preserve_mean scales by the pixel-count ratio) and :202-222
(_fft_ds_real_signal_to_shape), with crop_to_shape from
cryojax/ndimage/_edges.py:38-56 and cryoJAX's centred transforms
fftn(x) = fft(ifftshift(x)) and rfftn(x) = rfft(ifftshift(x))
(cryojax/ndimage/_fft.py:34-54, 80-99).

Pinning (tests/test_oracle_downsample.py), each against something other than
this code:
  * out shape == in shape returns the input (the Hartley transform is its own
    inverse up to the pixel count), even and odd sizes;
  * a constant image keeps its sum (its mean with preserve_mean);
  * a cosine below the coarse Nyquist frequency comes back as the same cosine
    on the coarse grid, times h w / (oh ow);
  * for odd output sizes the Hartley crop equals a plain centred Fourier crop
    (numpy ifft of the cropped spectrum), since every kept frequency keeps its
    mirror;
  * for even output sizes the unpaired -oh/2 bin: a cosine at that frequency
    comes back at half amplitude and a sine as minus a cosine of half
    amplitude (closed forms of cas at pi p);
  * outputs_real_space=False inverts with numpy's irfft2 back to the
    real-space result;
  * a 3D volume reduces to the 2D case for inputs constant along one axis;
  * complex input (the Fourier branch): the identity at equal shapes; for odd
    output sizes a real array given as complex equals the real (Hartley)
    branch, both being the plain centred Fourier crop; a band-limited complex
    exponential comes back as the same exponential on the coarse grid;
    outputs_real_space=False inverts (numpy ifftn, centred) to the real-space
    result.
"""

from __future__ import annotations

import numpy as np


def _fftn(x):
    """cryojax.ndimage.fftn (_fft.py:34-54): fft of the ifftshifted array."""
    return np.fft.fftn(np.fft.ifftshift(x))


def _rfftn(x):
    """cryojax.ndimage.rfftn (_fft.py:80-99)."""
    return np.fft.rfftn(np.fft.ifftshift(x))


def crop_to_shape(x, shape):
    """crop_to_shape (_edges.py:38-56): centred crop, centre index size // 2."""
    sl = tuple(slice(n // 2 - s // 2, n // 2 + s // 2 + s % 2) for n, s in zip(x.shape, shape))
    return x[sl]


def _fft_ds_complex(x, shape, outputs_real_space):
    """_fft_ds_complex_signal_to_shape (_downsample.py:225-238): centred Fourier crop, then
    cryoJAX's ifftn(ifftshift(c)) = fftshift(ifftn(ifftshift(c))) (_fft.py:11-31)."""
    f = np.fft.fftshift(_fftn(x))
    c = crop_to_shape(f, shape)
    if outputs_real_space:
        return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(c)))
    return np.fft.ifftshift(c)


def fourier_crop_downsample_to_shape(x, shape, outputs_real_space=True, preserve_mean=False):
    """_downsample.py:158-238: the Hartley branch for real input, the Fourier branch for complex."""
    shape = tuple(int(s) for s in shape)
    if np.iscomplexobj(x):
        x = np.asarray(x, dtype=np.complex128)
        if x.ndim not in (2, 3) or len(shape) != x.ndim:
            raise ValueError("images (2D) and volumes (3D) only, with a target shape of the same rank")
        out = _fft_ds_complex(x, shape, outputs_real_space)
        return (np.prod(shape) / np.prod(x.shape)) * out if preserve_mean else out
    x = np.asarray(x, dtype=np.float64)
    if x.ndim not in (2, 3) or len(shape) != x.ndim:
        raise ValueError("images (2D) and volumes (3D) only, with a target shape of the same rank")
    hart = np.fft.fftshift(_fftn(x))
    hart = hart.real - hart.imag
    ds = crop_to_shape(hart, shape)
    ds = np.fft.fftshift(_fftn(ds))
    ds = ds / ds.size
    ds = ds.real - ds.imag
    out = ds if outputs_real_space else _rfftn(ds)
    if preserve_mean:
        out = (np.prod(shape) / np.prod(x.shape)) * out
    return out


def downsampled_shape(shape, downsample_factor):
    """_downsample.py:118-147: int(size / factor) per axis, factor >= 1."""
    f = float(downsample_factor)
    if f < 1.0:
        raise ValueError("downsample_factor must be at least 1")
    return tuple(int(s / f) for s in shape)


def fourier_crop_downsample(x, downsample_factor, outputs_real_space=True, preserve_mean=False):
    """_downsample.py:91-155."""
    x = np.asarray(x)
    return fourier_crop_downsample_to_shape(
        x, downsampled_shape(x.shape, downsample_factor), outputs_real_space, preserve_mean
    )
