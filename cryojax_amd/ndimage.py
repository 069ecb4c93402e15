"""Image operations beside the image-formation path, on device.

Mirrors cryojax.ndimage.fourier_crop_downsample and
fourier_crop_downsample_to_shape (cryojax/ndimage/_downsample.py:91-238):
the Hartley-transform crop of real arrays, the Fourier crop of complex ones,
`preserve_mean`, and `outputs_real_space=False` (rfftn of a real result, the
uncentred crop of a complex one). Real images run through the fused FFT kernel
or the batched FP32 MFMA contractions behind include/cryomi_ndimage.h (the
operators are built on device once per shape and kept); volumes and complex
arrays through FP64 centred transforms composed with the crop and
recombination kernels of the same header.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .ops import Workspace, _require_cuda, _stream_ptr

_ws = Workspace()
_ops_cache: dict = {}


def downsampled_shape(shape, downsample_factor):
    """int(size / factor) per axis (_downsample.py:118-147); factor must be >= 1."""
    f = float(downsample_factor)
    if f < 1.0:
        raise ValueError("Called `fourier_crop_downsample` with `downsample_factor` less than 1.")
    return tuple(int(s / f) for s in shape)


def _flags(outputs_real_space, preserve_mean):
    return (0 if outputs_real_space else _lib.DS_OUTPUT_FOURIER) | (_lib.DS_PRESERVE_MEAN if preserve_mean else 0)


def crop_operators(h, w, oh, ow, flags, device):
    """The device operator buffer for one (h, w) -> (oh, ow) map (cached per device)."""
    key = (str(device), h, w, oh, ow, flags)
    ops = _ops_cache.get(key)
    if ops is None:
        lib = _lib.load()
        n = lib.cryomi_fourier_crop_operator_floats(h, w, oh, ow, flags)
        if n == 0:
            raise ValueError(f"unsupported downsampling {h}x{w} -> {oh}x{ow} (sizes 1 <= out <= in <= 4096)")
        ops = torch.empty(n, dtype=torch.float32, device=device)
        with torch.cuda.device(device):
            st = lib.cryomi_fourier_crop_operators(h, w, oh, ow, flags, ops.data_ptr(), _stream_ptr(device))
        _lib.check(st, "cryomi_fourier_crop_operators")
        _ops_cache[key] = ops
    return ops


def fourier_crop_downsample_images(images, downsampled_shape, outputs_real_space=True, preserve_mean=False,
                                   out=None):
    """Downsample a batch of real images (..., h, w) float32 to (..., oh, ow).

    Returns float32 (..., oh, ow), or complex64 (..., oh, ow // 2 + 1) with
    outputs_real_space=False.
    """
    lib = _lib.load()
    _require_cuda(images, "images", torch.float32)
    if images.ndim < 2:
        raise ValueError("images must have shape (..., h, w)")
    h, w = int(images.shape[-2]), int(images.shape[-1])
    oh, ow = (int(s) for s in downsampled_shape)
    lead = tuple(images.shape[:-2])
    n = 1
    for s in lead:
        n *= int(s)
    flags = _flags(outputs_real_space, preserve_mean)
    dev = images.device
    ops = crop_operators(h, w, oh, ow, flags, dev)
    if outputs_real_space:
        shape, dtype = (*lead, oh, ow), torch.float32
    else:
        shape, dtype = (*lead, oh, ow // 2 + 1), torch.complex64
    if out is None:
        out = torch.empty(shape, dtype=dtype, device=dev)
    else:
        _require_cuda(out, "out", dtype)
        if tuple(out.shape) != shape:
            raise ValueError(f"out must have shape {shape}")
    nbytes = lib.cryomi_fourier_crop_downsample_workspace_bytes(n, h, w, oh, ow, flags)
    ws = _ws.get(dev, nbytes)
    with torch.cuda.device(dev):
        st = lib.cryomi_fourier_crop_downsample(images.data_ptr(), n, h, w, oh, ow, flags, ops.data_ptr(),
                                                ws.data_ptr(), ws.numel(), out.data_ptr(), _stream_ptr(dev))
    _lib.check(st, "cryomi_fourier_crop_downsample")
    return out


def _transform(x):
    """T(x) = fftshift(fftn(ifftshift(x))) of a real (d, h, w) array, complex128, FP64 on device."""
    from . import ops

    g128, _ = ops.fourier_voxel_grid(x.contiguous(), complex128=True, complex64=False)
    return g128


def _spectrum_crop(g, out_shape, flags, scale, dtype):
    lib = _lib.load()
    d, h, w = (int(s) for s in g.shape)
    out = torch.empty(tuple(out_shape), dtype=dtype, device=g.device)
    with torch.cuda.device(g.device):
        st = lib.cryomi_spectrum_crop(g.data_ptr(), d, h, w, *(int(s) for s in out_shape), flags, float(scale),
                                      out.data_ptr(), _stream_ptr(g.device))
    _lib.check(st, "cryomi_spectrum_crop")
    return out


def _split(x):
    lib = _lib.load()
    re = torch.empty(x.shape, dtype=torch.float64, device=x.device)
    im = torch.empty(x.shape, dtype=torch.float64, device=x.device)
    with torch.cuda.device(x.device):
        st = lib.cryomi_complex_split(x.data_ptr(), 1 if x.dtype == torch.complex128 else 0, x.numel(), re.data_ptr(),
                                      im.data_ptr(), _stream_ptr(x.device))
    _lib.check(st, "cryomi_complex_split")
    return re, im


def _combine(a, b, inverse, scale, dtype):
    lib = _lib.load()
    out = torch.empty(a.shape, dtype=dtype, device=a.device)
    with torch.cuda.device(a.device):
        st = lib.cryomi_complex_combine(a.data_ptr(), b.data_ptr(), a.numel(), 1 if inverse else 0, float(scale),
                                        1 if dtype == torch.complex64 else 0, out.data_ptr(), _stream_ptr(a.device))
    _lib.check(st, "cryomi_complex_combine")
    return out


def _downsample_by_transforms(x, shape, outputs_real_space, preserve_mean):
    """Volumes and complex arrays (_downsample.py:202-238) through FP64 centred transforms of real
    arrays: the Hartley crop both ways for real input, the centred Fourier crop and its inverse for
    complex input (T(a + i b) = T(a) + i T(b)). 32-bit input gives float32 / complex64 output."""
    x = x.contiguous()
    vol = x.reshape((1,) + tuple(x.shape)) if x.ndim == 2 else x
    out3 = (1,) + tuple(shape) if x.ndim == 2 else tuple(shape)
    n_in, n_out = vol.numel(), 1
    for s in out3:
        n_out *= int(s)
    factor = n_out / n_in if preserve_mean else 1.0
    single = x.dtype in (torch.float32, torch.complex64)
    f32 = _lib.SPEC_OUT_F32 if single else 0
    if not torch.is_complex(x):
        hc = _spectrum_crop(_transform(vol), out3, _lib.SPEC_HARTLEY, 1.0, torch.float64)
        g2 = _transform(hc)
        if outputs_real_space:
            y = _spectrum_crop(g2, out3, _lib.SPEC_HARTLEY | f32, factor / n_out,
                               torch.float32 if single else torch.float64)
        else:
            y64 = _spectrum_crop(g2, out3, _lib.SPEC_HARTLEY, 1.0 / n_out, torch.float64)
            half = out3[:-1] + (out3[-1] // 2 + 1,)
            y = _spectrum_crop(_transform(y64), half, _lib.SPEC_UNCENTRE | f32, factor,
                               torch.complex64 if single else torch.complex128)
    else:
        cdt = torch.complex64 if single else torch.complex128
        re, im = _split(vol)
        full = _combine(_transform(re), _transform(im), False, 1.0, torch.complex128)
        c = _spectrum_crop(full, out3, 0, 1.0, torch.complex128)
        if outputs_real_space:
            re2, im2 = _split(c)
            y = _combine(_transform(re2), _transform(im2), True, factor / n_out, cdt)
        else:
            y = _spectrum_crop(c, out3, _lib.SPEC_UNCENTRE | f32, factor, cdt)
    return y[0] if x.ndim == 2 else y


def fourier_crop_downsample_to_shape(image_or_volume, downsampled_shape, outputs_real_space=True,
                                     preserve_mean=False):
    """cryojax.ndimage.fourier_crop_downsample_to_shape (_downsample.py:158-199): real images take
    the batched image path; volumes and complex input the composed FP64 transforms."""
    x = image_or_volume
    if not isinstance(x, torch.Tensor):
        raise ValueError("expected a torch tensor on a ROCm device")
    if x.ndim not in (2, 3):
        raise ValueError(f"`fourier_crop_downsample` was passed an array with `ndim = {x.ndim}`, but this function "
                         "only supports images and volumes as input.")
    if len(downsampled_shape) != x.ndim:
        raise ValueError("downsampled_shape must have the array's rank")
    if any(int(o) > int(s) or int(o) < 1 for o, s in zip(downsampled_shape, x.shape)):
        raise ValueError("downsampled_shape must lie between 1 and the input shape on every axis")
    if x.ndim == 2 and x.dtype == torch.float32:
        return fourier_crop_downsample_images(x.contiguous(), downsampled_shape, outputs_real_space, preserve_mean)
    if not (torch.is_complex(x) or x.dtype in (torch.float32, torch.float64)):
        raise ValueError(f"unsupported dtype {x.dtype}")
    _require_cuda(x, "image_or_volume")
    return _downsample_by_transforms(x, tuple(int(s) for s in downsampled_shape), outputs_real_space, preserve_mean)


def fourier_crop_downsample(image_or_volume, downsample_factor, outputs_real_space=True, preserve_mean=False):
    """cryojax.ndimage.fourier_crop_downsample (_downsample.py:91-155)."""
    shape = downsampled_shape(tuple(image_or_volume.shape), downsample_factor)
    return fourier_crop_downsample_to_shape(image_or_volume, shape, outputs_real_space, preserve_mean)


def bgemm_nt(a, b, c, m, n, k, lda, ldb, ldc, a_batch=0, a_seg=0, b_batch=0, b_seg=0, c_batch=0, nseg=1, batch=1):
    """cryomi_bgemm_nt on float32 device buffers (strides in floats)."""
    lib = _lib.load()
    for t, name in ((a, "a"), (b, "b"), (c, "c")):
        _require_cuda(t, name, torch.float32)
    args = _lib.BgemmStruct(a.data_ptr(), a_batch, a_seg, lda, b.data_ptr(), b_batch, b_seg, ldb, c.data_ptr(),
                            c_batch, ldc, m, n, k, nseg, batch)
    with torch.cuda.device(c.device):
        st = lib.cryomi_bgemm_nt(ctypes.addressof(args), _stream_ptr(c.device))
    _lib.check(st, "cryomi_bgemm_nt")
    return c
