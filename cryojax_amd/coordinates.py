"""Coordinate and frequency grids of images and volumes (host arrays, float64).

Mirrors cryojax/coordinates/_make_coordinate_grids.py:10-317 by name, argument meaning and
layout: a 2D grid is (y, x, 2) with [..., 0] = x and [..., 1] = y; a 3D grid is (z, y, x, 3) with
[..., 0] = x, [..., 1] = y, [..., 2] = z. Real-space coordinates are centred, (i - n // 2) times
the grid spacing along each axis (:301-317, fftshift(fftfreq(n, 1 / dx)) * n). Frequencies have
the zero frequency in the corner (fftfreq / rfftfreq of the axis with spacing `grid_spacing`);
with `outputs_rfftfreqs` the last axis holds rfftfreq (:259-298).
"""

from __future__ import annotations

import numpy as np


def _axis(n, dx, real, rfft):
    if real:
        return (np.arange(n, dtype=np.float64) - n // 2) * dx
    return (np.fft.rfftfreq if rfft else np.fft.fftfreq)(n, dx)


def _grid(shape, grid_spacing, real, rfft):
    shape = tuple(int(s) for s in shape)
    if len(shape) not in (2, 3):
        raise ValueError(f"Only 2D and 3D coordinate grids are supported. Tried to create a grid of shape {shape}.")
    dx = float(grid_spacing)
    axes = [_axis(n, dx, real, rfft and i == len(shape) - 1) for i, n in enumerate(shape)]
    mesh = np.meshgrid(*axes, indexing="ij")  # (z,) y, x order
    return np.stack(mesh[::-1], axis=-1)  # x first


def make_coordinate_grid(shape, grid_spacing=1.0):
    """Centred real-space coordinates, (*shape, ndim) (_make_coordinate_grids.py:10-32)."""
    return _grid(shape, grid_spacing, True, False)


def make_frequency_grid(shape, grid_spacing=1.0, outputs_rfftfreqs=True):
    """Frequencies with the zero frequency in the corner (_make_coordinate_grids.py:67-97); the
    last axis is halved (rfftfreq) with `outputs_rfftfreqs`."""
    return _grid(shape, grid_spacing, False, outputs_rfftfreqs)


def make_radial_frequency_grid(shape, grid_spacing=1.0, outputs_rfftfreqs=True):
    """|make_frequency_grid| (_make_coordinate_grids.py:100-136)."""
    return np.linalg.norm(make_frequency_grid(shape, grid_spacing, outputs_rfftfreqs), axis=-1)
