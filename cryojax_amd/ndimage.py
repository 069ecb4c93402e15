"""Image operations beside the image-formation path, on device.

Mirrors cryojax.ndimage.fourier_crop_downsample and
fourier_crop_downsample_to_shape (cryojax/ndimage/_downsample.py:91-222) for
real images: the Hartley-transform crop, `preserve_mean`, and
`outputs_real_space=False` (rfftn of the result, zero frequency in the
corner). Every image runs through the batched FP32 MFMA contractions behind
include/cryomi_ndimage.h; the operators are built on device once per shape
and kept.
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


def fourier_crop_downsample_to_shape(image_or_volume, downsampled_shape, outputs_real_space=True,
                                     preserve_mean=False):
    """cryojax.ndimage.fourier_crop_downsample_to_shape (_downsample.py:158-199) for a real image."""
    x = image_or_volume
    if not isinstance(x, torch.Tensor):
        raise ValueError("expected a torch tensor on a ROCm device")
    if torch.is_complex(x):
        raise NotImplementedError("complex input (the Fourier-crop branch, _downsample.py:225-238) is not built")
    if x.ndim == 3:
        raise NotImplementedError("volumes (3D) are not built; pass images, or batches through "
                                  "fourier_crop_downsample_images")
    if x.ndim != 2:
        raise ValueError(f"`fourier_crop_downsample` was passed an array with `ndim = {x.ndim}`, but this function "
                         "only supports images and volumes as input.")
    if len(downsampled_shape) != 2:
        raise ValueError("downsampled_shape must have the image's rank")
    return fourier_crop_downsample_images(x.contiguous(), downsampled_shape, outputs_real_space, preserve_mean)


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
