"""FP32 NumPy/SciPy restatement of the same forward path, timed as the CPU baseline.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py's `cpu_baseline`
leg times it on the host cores; it is never the measured GPU path.

It follows cryoJAX's float32 (x64 off) arithmetic: rotated frequency plane,
logical coordinates c = R k N + N//2 in float32 (fourier_voxel_extract.py:
519-544), trilinear fill gather (ndimage/_map_coordinates.py:85-123,168-175),
real-part fix-ups (_fourier_utils.py:6-132), CTF (transfer_function.py:27-169),
zero-then-shift (_pose.py:68-114), fftshift(irfft2) (ndimage/_fft.py:57-77).
Images are split over a thread pool; scipy.fft runs with one worker per thread.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft

from . import imaging as O


def _host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _simulate_chunk(grid64, mats, offsets, defocus, astig, angle, pixel_size, voltage, ac, cs):
    n = grid64.shape[0]
    m = n // 2
    f = np.float32
    kx = (np.fft.fftfreq(n)[: m + 1]).astype(f)
    ky = np.fft.fftfreq(n).astype(f)
    kxx, kyy = np.meshgrid(kx, ky, indexing="xy")
    b = mats.shape[0]
    # rotated plane (float32), logical coordinates
    c = np.empty((b, n, m + 1, 3), dtype=f)
    for a in range(3):
        c[..., a] = (mats[:, a, 0, None, None] * kxx + mats[:, a, 1, None, None] * kyy) * f(n) + f(n // 2)
    lower = np.floor(c)
    upper_w = c - lower
    idx = lower.astype(np.int32)
    out = np.zeros((b, n, m + 1), dtype=np.complex64)
    for a0 in (0, 1):
        i0 = idx[..., 2] + a0
        w0 = upper_w[..., 2] if a0 else 1 - upper_w[..., 2]
        i0 = np.where(i0 < 0, i0 + n, i0)
        v0 = (i0 >= 0) & (i0 < n)
        for a1 in (0, 1):
            i1 = idx[..., 1] + a1
            w1 = upper_w[..., 1] if a1 else 1 - upper_w[..., 1]
            i1 = np.where(i1 < 0, i1 + n, i1)
            v1 = v0 & (i1 >= 0) & (i1 < n)
            for a2 in (0, 1):
                i2 = idx[..., 0] + a2
                w2 = upper_w[..., 0] if a2 else 1 - upper_w[..., 0]
                i2 = np.where(i2 < 0, i2 + n, i2)
                v = v1 & (i2 >= 0) & (i2 < n)
                val = grid64[np.where(v, i0, 0), np.where(v, i1, 0), np.where(v, i2, 0)]
                out += np.where(v, w0 * w1 * w2, f(0)) * val
    out[:, m, 0] = 0
    out[:, 0, m] = 0
    out[:, m, m] = 0
    freq = O.make_frequency_grid((n, n), outputs_rfftfreqs=True).astype(f) / f(pixel_size)
    ctf = O.astigmatic_ctf(freq, voltage, defocus, astig, angle, cs, ac).astype(f)
    shift = O.translation_operator(freq, offsets).astype(np.complex64)
    spec = out * ctf * shift
    img = sfft.irfftn(spec, s=(n, n), axes=(-2, -1), workers=1)
    return np.fft.fftshift(img, axes=(-2, -1)).astype(f)


def simulate_fp32(grid_c64, mats, offsets, defocus, astig, angle, pixel_size, voltage=300.0, ac=0.1, cs=2.7,
                  threads=None, chunk=4):
    """Batched float32 forward on the CPU; returns (images, threads_used)."""
    threads = threads or _host_cores()
    b = mats.shape[0]
    starts = list(range(0, b, chunk))

    def work(s):
        e = min(s + chunk, b)
        return _simulate_chunk(grid_c64, mats[s:e], offsets[s:e], defocus[s:e], astig[s:e], angle[s:e],
                               pixel_size, voltage, ac, cs)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        parts = list(ex.map(work, starts))
    return np.concatenate(parts, axis=0), threads
