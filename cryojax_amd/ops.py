"""Device-side calls into libcryomi on torch-owned HBM buffers.

torch supplies device memory and the current HIP stream; all arithmetic runs
in the HIP kernels behind the C ABI (include/cryomi.h).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream


def _require_cuda(t, name, dtype=None):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{name} must be a torch tensor on a ROCm device")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class Workspace:
    """A grow-only device scratch buffer reused across calls on one device."""

    def __init__(self):
        self._buf = {}

    def get(self, device, nbytes):
        key = str(device)
        buf = self._buf.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
            self._buf[key] = buf
        return buf

    def holds(self, device, nbytes):
        """True when get(device, nbytes) would return the current buffer (contents kept)."""
        buf = self._buf.get(str(device))
        return buf is not None and buf.numel() >= nbytes


_default_ws = Workspace()


def params_to_device(params_host, device):
    """(n, 20) float64 host table -> contiguous device tensor."""
    t = torch.as_tensor(params_host, dtype=torch.float64)
    if t.ndim != 2 or t.shape[1] != _lib.N_PARAM:
        raise ValueError(f"params must have shape (n, {_lib.N_PARAM})")
    return t.contiguous().to(device)


def simulate_images(grid, params_dev, flags, out=None, workspace=None, box=None):
    """cryomi_simulate on device buffers.

    grid      : complex64 (N, N, N) cuda tensor, zero frequency at N//2, or the
                flat output of grid_to_bricks with GRID_BRICKED (then box = N)
    params_dev: float64 (n, 20) cuda tensor (cryomi_image_params)
    returns   : float32 (n, N, N), or complex64 (n, N, N//2+1) with OUTPUT_FOURIER
    """
    lib = _lib.load()
    _require_cuda(grid, "grid", torch.complex64)
    _require_cuda(params_dev, "params", torch.float64)
    if flags & _lib.GRID_BRICKED:
        if box is None:
            raise ValueError("box is required with a bricked grid")
        n = int(box)
        if grid.numel() * 8 < lib.cryomi_bricked_grid_bytes(n):
            raise ValueError("bricked grid is smaller than cryomi_bricked_grid_bytes(box)")
    else:
        n = int(grid.shape[0])
        if tuple(grid.shape) != (n, n, n):
            raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
    n_images = int(params_dev.shape[0])
    if out is None:
        if flags & _lib.OUTPUT_FOURIER:
            out = torch.empty((n_images, n, n // 2 + 1), dtype=torch.complex64, device=grid.device)
        else:
            out = torch.empty((n_images, n, n), dtype=torch.float32, device=grid.device)
    else:
        want = torch.complex64 if flags & _lib.OUTPUT_FOURIER else torch.float32
        _require_cuda(out, "out", want)
    nbytes = lib.cryomi_simulate_workspace_bytes(n, max(n_images, 1), flags)
    if workspace is None:
        ws = _default_ws.get(grid.device, nbytes) if nbytes else None
    else:
        ws = workspace
    with torch.cuda.device(grid.device):
        st = lib.cryomi_simulate(
            grid.data_ptr(),
            n,
            n_images,
            params_dev.data_ptr(),
            flags,
            ws.data_ptr() if ws is not None else None,
            ws.numel() if ws is not None else 0,
            out.data_ptr(),
            _stream_ptr(grid.device),
        )
    _lib.check(st, "cryomi_simulate")
    return out


def simulate_band(grid, params_dev, flags, band, out=None):
    """cryomi_simulate_band: the (n, band + 1, band/2 + 1) complex64 low-frequency band of the
    images' rfft half-planes (row r = frequency r - band/2), the frequencies a Fourier crop of the
    box images to band x band reads."""
    lib = _lib.load()
    _require_cuda(grid, "grid", torch.complex64)
    _require_cuda(params_dev, "params", torch.float64)
    n = int(grid.shape[0])
    if tuple(grid.shape) != (n, n, n):
        raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
    n_images = int(params_dev.shape[0])
    shape = (n_images, band + 1, band // 2 + 1)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=grid.device)
    else:
        _require_cuda(out, "out", torch.complex64)
        if tuple(out.shape) != shape:
            raise ValueError(f"out must have shape {shape}")
    with torch.cuda.device(grid.device):
        st = lib.cryomi_simulate_band(grid.data_ptr(), n, n_images, params_dev.data_ptr(), flags, int(band),
                                      out.data_ptr(), _stream_ptr(grid.device))
    _lib.check(st, "cryomi_simulate_band")
    return out


def grid_to_bricks(grid):
    """The kernels' bricked copy of a centred complex64 (N, N, N) grid (cryomi_grid_to_bricks).

    Returns a flat complex64 tensor of cryomi_bricked_grid_bytes(N) / 8 elements;
    pass it to simulate_images with GRID_BRICKED set.
    """
    lib = _lib.load()
    _require_cuda(grid, "grid", torch.complex64)
    n = int(grid.shape[0])
    if tuple(grid.shape) != (n, n, n):
        raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
    nbytes = lib.cryomi_bricked_grid_bytes(n)
    if nbytes == 0:
        raise NotImplementedError(f"no bricked layout for box {n} (needs a supported box, multiple of 4)")
    out = torch.empty(nbytes // 8, dtype=torch.complex64, device=grid.device)
    with torch.cuda.device(grid.device):
        st = lib.cryomi_grid_to_bricks(grid.data_ptr(), n, out.data_ptr(), _stream_ptr(grid.device))
    _lib.check(st, "cryomi_grid_to_bricks")
    return out


def irfft2(spectrum, box):
    """fftshift(irfft2(spectrum, s=(box, box))) over a leading batch (cryomi_irfft2)."""
    lib = _lib.load()
    _require_cuda(spectrum, "spectrum", torch.complex64)
    lead = spectrum.shape[:-2]
    if tuple(spectrum.shape[-2:]) != (box, box // 2 + 1):
        raise ValueError(f"spectrum must end in ({box}, {box // 2 + 1})")
    flat = spectrum.reshape(-1, box, box // 2 + 1)
    n = int(flat.shape[0])
    out = torch.empty((n, box, box), dtype=torch.float32, device=spectrum.device)
    nbytes = lib.cryomi_irfft2_workspace_bytes(box, max(n, 1))
    ws = _default_ws.get(spectrum.device, nbytes)
    with torch.cuda.device(spectrum.device):
        st = lib.cryomi_irfft2(
            flat.data_ptr(), box, n, ws.data_ptr(), ws.numel(), out.data_ptr(),
            _stream_ptr(spectrum.device),
        )
    _lib.check(st, "cryomi_irfft2")
    return out.reshape(*lead, box, box)


def normalize_images(images, region=None):
    """In place: each image (..., N, N) minus its mean over `region`, divided by
    its population std over `region` (cryomi_normalize_images)."""
    lib = _lib.load()
    _require_cuda(images, "images", torch.float32)
    npix = int(images.shape[-1] * images.shape[-2])
    n = int(images.numel() // npix)
    reg = None
    if region is not None:
        reg = torch.as_tensor(np.asarray(region, dtype=np.uint8).reshape(-1)).to(images.device)
        if reg.numel() != npix:
            raise ValueError("signal_region must have the image's shape")
    with torch.cuda.device(images.device):
        st = lib.cryomi_normalize_images(images.data_ptr(), n, npix, None if reg is None else reg.data_ptr(),
                                         _stream_ptr(images.device))
    _lib.check(st, "cryomi_normalize_images")
    return images


_MB_TYPES = {torch.float32: 0, torch.complex64: 1, torch.float64: 2, torch.complex128: 3}  # CRYOMI_F32 ...


def multiply_broadcast(data, factor):
    """In place: every trailing block of `data` ((..., *factor.shape), float32, complex64, float64 or
    complex128) times `factor` (a real or complex array of that shape, used at the data's
    precision), on device (cryomi_multiply_broadcast). The multiplicative image transforms:
    filters on rfft images and Fourier grids, masks on real images, CTF envelopes."""
    lib = _lib.load()
    if not isinstance(data, torch.Tensor) or not data.is_cuda or data.dtype not in _MB_TYPES:
        raise TypeError("data must be a CUDA float32, complex64, float64 or complex128 tensor")
    if not data.is_contiguous():
        raise ValueError("data must be contiguous")
    f = factor if isinstance(factor, torch.Tensor) else torch.from_numpy(np.array(factor))
    fshape = tuple(f.shape)
    if tuple(data.shape[data.dim() - len(fshape):]) != fshape:
        raise ValueError(f"factor of shape {fshape} does not match the trailing dimensions of {tuple(data.shape)}")
    fc, dc = f.is_complex(), data.is_complex()
    if fc and not dc:
        raise ValueError("a complex factor needs complex data")
    double = data.dtype in (torch.float64, torch.complex128)
    fdt = (torch.complex128 if fc else torch.float64) if double else (torch.complex64 if fc else torch.float32)
    f = f.to(device=data.device, dtype=fdt).contiguous()
    npix = int(f.numel())
    n = int(data.numel() // npix) if npix else 0
    with torch.cuda.device(data.device):
        st = lib.cryomi_multiply_broadcast(data.data_ptr(), n, npix, _MB_TYPES[data.dtype], f.data_ptr(),
                                           _MB_TYPES[fdt], _stream_ptr(data.device))
    _lib.check(st, "cryomi_multiply_broadcast")
    return data


def resize_images(images, shape, normalize=False, region=None):
    """Centred crop / zero pad of real images (..., h, w) to (..., oh, ow), then, with
    normalize, mean 0 and population std 1 over `region` (cryomi_resize_images)."""
    lib = _lib.load()
    _require_cuda(images, "images", torch.float32)
    h, w = int(images.shape[-2]), int(images.shape[-1])
    oh, ow = (int(s) for s in shape)
    lead = tuple(images.shape[:-2])
    n = int(images.numel() // max(h * w, 1))
    out = torch.empty((*lead, oh, ow), dtype=torch.float32, device=images.device)
    reg = None
    if region is not None:
        reg = torch.as_tensor(np.asarray(region, dtype=np.uint8).reshape(-1)).to(images.device)
        if reg.numel() != oh * ow:
            raise ValueError("signal_region must have the output image's shape")
    flags = _lib.RESIZE_NORMALIZE if normalize else 0
    with torch.cuda.device(images.device):
        st = lib.cryomi_resize_images(images.data_ptr(), n, h, w, oh, ow, flags,
                                      None if reg is None else reg.data_ptr(), out.data_ptr(),
                                      _stream_ptr(images.device))
    _lib.check(st, "cryomi_resize_images")
    return out


def complex128_to_complex64(x):
    """complex64 copy of a complex128 device tensor (cryomi_complex128_to_complex64)."""
    lib = _lib.load()
    _require_cuda(x, "x", torch.complex128)
    out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
    with torch.cuda.device(x.device):
        st = lib.cryomi_complex128_to_complex64(x.data_ptr(), x.numel(), out.data_ptr(), _stream_ptr(x.device))
    _lib.check(st, "cryomi_complex128_to_complex64")
    return out


def rfft2(images):
    """rfftn(ifftshift(y)) of real images as complex64 (the FP64 transform, rounded once)."""
    return complex128_to_complex64(rfft2_f64(images))


def simulate_from_spectrum(spectrum, params_dev, flags, out=None):
    """cryomi_simulate_from_spectrum: fix-ups, CTF and shift of complex64 half-planes (n, P, P/2+1),
    then fftshift(irfft2) to float32 (n, P, P), or the half-planes with OUTPUT_FOURIER."""
    lib = _lib.load()
    _require_cuda(spectrum, "spectrum", torch.complex64)
    _require_cuda(params_dev, "params", torch.float64)
    n, box = int(spectrum.shape[0]), int(spectrum.shape[1])
    if tuple(spectrum.shape[1:]) != (box, box // 2 + 1) or int(params_dev.shape[0]) != n:
        raise ValueError("spectrum must be (n, P, P//2+1) with one parameter row per image")
    fourier = bool(flags & _lib.OUTPUT_FOURIER)
    if out is None:
        shp = (n, box, box // 2 + 1) if fourier else (n, box, box)
        out = torch.empty(shp, dtype=torch.complex64 if fourier else torch.float32, device=spectrum.device)
    ws = None if fourier else _default_ws.get(spectrum.device, lib.cryomi_irfft2_workspace_bytes(box, max(n, 1)))
    with torch.cuda.device(spectrum.device):
        st = lib.cryomi_simulate_from_spectrum(spectrum.data_ptr(), box, n, params_dev.data_ptr(), flags,
                                               None if ws is None else ws.data_ptr(), 0 if ws is None else ws.numel(),
                                               out.data_ptr(), _stream_ptr(spectrum.device))
    _lib.check(st, "cryomi_simulate_from_spectrum")
    return out


def white_noise_spectrum(n, box, variance, seed, device):
    """(n, box, box/2+1) complex64 real-valued Gaussian spectra of std sqrt(box^2 variance), zero DC."""
    lib = _lib.load()
    out = torch.empty((n, box, box // 2 + 1), dtype=torch.complex64, device=device)
    with torch.cuda.device(device):
        st = lib.cryomi_white_noise_spectrum(int(n), int(box), float(variance), int(seed) & (2**64 - 1),
                                             out.data_ptr(), _stream_ptr(device))
    _lib.check(st, "cryomi_white_noise_spectrum")
    return out


def noisy_images(signal, noise=None, scale=1.0, offset=0.0, out=None):
    """scale * signal + offset + noise (cryomi_noisy_images)."""
    lib = _lib.load()
    _require_cuda(signal, "signal", torch.float32)
    if noise is not None:
        _require_cuda(noise, "noise", torch.float32)
        if noise.shape != signal.shape:
            raise ValueError("noise and signal shapes differ")
    out = torch.empty_like(signal) if out is None else out
    with torch.cuda.device(signal.device):
        st = lib.cryomi_noisy_images(signal.data_ptr(), None if noise is None else noise.data_ptr(), signal.numel(),
                                     float(scale), float(offset), out.data_ptr(), _stream_ptr(signal.device))
    _lib.check(st, "cryomi_noisy_images")
    return out


def gaussian_log_likelihood(simulated, observed, variance):
    """Per-image float64 log-likelihood (cryomi_gaussian_log_likelihood)."""
    lib = _lib.load()
    _require_cuda(simulated, "simulated", torch.float32)
    _require_cuda(observed, "observed", torch.float32)
    if simulated.shape != observed.shape:
        raise ValueError("simulated and observed shapes differ")
    npix = int(simulated.shape[-1] * simulated.shape[-2])
    n = int(simulated.numel() // npix)
    out = torch.empty(n, dtype=torch.float64, device=simulated.device)
    with torch.cuda.device(simulated.device):
        st = lib.cryomi_gaussian_log_likelihood(simulated.data_ptr(), observed.data_ptr(), n, npix, float(variance),
                                                out.data_ptr(), _stream_ptr(simulated.device))
    _lib.check(st, "cryomi_gaussian_log_likelihood")
    return out.reshape(simulated.shape[:-2])


def _real_input(x, name):
    _require_cuda(x, name)
    if x.dtype == torch.float32:
        return _lib.REAL_F32
    if x.dtype == torch.float64:
        return _lib.REAL_F64
    raise ValueError(f"{name} must be float32 or float64, got {x.dtype}")


def _aligned16(x):
    """x, or an aligned copy when its data does not start on a 16-byte boundary (a contiguous view
    at an offset): the FFT paths of the C ABI refuse such buffers."""
    return x if x.data_ptr() % 16 == 0 else x.clone()


def fourier_voxel_grid(real_grid, padded_shape=None, complex128=True, complex64=True):
    """fftshift(fftn(ifftshift(pad_to_shape(x, padded_shape)))) of a real (d, h, w) volume with zero
    padding, in FP64 on device (cryomi_fourier_voxel_grid). Returns (complex128 or None, complex64 or None)."""
    lib = _lib.load()
    kind = _real_input(real_grid, "real_voxel_grid")
    real_grid = _aligned16(real_grid)
    if real_grid.ndim != 3:
        raise ValueError("real_voxel_grid must have shape (d, h, w)")
    d, h, w = (int(s) for s in real_grid.shape)
    pd, ph, pw = (d, h, w) if padded_shape is None else (int(s) for s in padded_shape)
    dev = real_grid.device
    nbytes = lib.cryomi_fourier_voxel_grid_workspace_bytes(d, h, w, pd, ph, pw)
    if nbytes == 0:
        raise ValueError(f"unsupported volume {d}x{h}x{w} padded to {pd}x{ph}x{pw} (sizes up to 1024)")
    ws = _default_ws.get(dev, nbytes)
    g128 = torch.empty((pd, ph, pw), dtype=torch.complex128, device=dev) if complex128 else None
    g64 = torch.empty((pd, ph, pw), dtype=torch.complex64, device=dev) if complex64 else None
    with torch.cuda.device(dev):
        st = lib.cryomi_fourier_voxel_grid(real_grid.data_ptr(), kind, d, h, w, pd, ph, pw,
                                           None if g128 is None else g128.data_ptr(),
                                           None if g64 is None else g64.data_ptr(), ws.data_ptr(), ws.numel(),
                                           _stream_ptr(dev))
    _lib.check(st, "cryomi_fourier_voxel_grid")
    return g128, g64


def rfft2_f64(images):
    """rfftn(ifftshift(y)) over the last two axes of real images (..., h, w), complex128 on device
    (cryomi_rfft2_f64)."""
    lib = _lib.load()
    kind = _real_input(images, "images")
    images = _aligned16(images)
    if images.ndim < 2:
        raise ValueError("images must have shape (..., h, w)")
    h, w = int(images.shape[-2]), int(images.shape[-1])
    lead = tuple(images.shape[:-2])
    n = int(images.numel() // max(h * w, 1))
    dev = images.device
    out = torch.empty((*lead, h, w // 2 + 1), dtype=torch.complex128, device=dev)
    nbytes = lib.cryomi_rfft2_f64_workspace_bytes(n, h, w)
    if nbytes == 0:
        raise ValueError(f"unsupported image size {h}x{w} (sizes up to 4096)")
    ws = _default_ws.get(dev, nbytes)
    with torch.cuda.device(dev):
        st = lib.cryomi_rfft2_f64(images.data_ptr(), kind, n, h, w, out.data_ptr(), ws.data_ptr(), ws.numel(),
                                  _stream_ptr(dev))
    _lib.check(st, "cryomi_rfft2_f64")
    return out
