"""Multiplicative image transforms and Fourier operators around the image model.

Mirrors, by name and meaning:
- the image transforms of cryojax/ndimage/transforms/ (`_base_transform.py:12-57`,
  `_filters.py:22-56`, `_masks.py:18-46`): a filter multiplies an rfft image by `get()`
  (`_filters.py:30-33`), a mask multiplies a real image by `get()` (`_masks.py:25-28`), and `*`
  composes two transforms (`ProductImageTransform`);
- the Fourier operators of cryojax/ndimage/operators/ (`_base_operator.py:15-160`,
  `_fourier_operator.py:64-234`): functions of a frequency grid (..., 2) in inverse Angstroms
  or inverse pixels, composed with `+`, `-` and `*`; `ContrastTransferTheory(envelope=...)`
  multiplies the contrast spectrum by `envelope(frequency_grid)`
  (`simulator/_transfer_theory/transfer_theory.py:117-118`).

A transform's array and an operator's values on the image's frequency grid are formed once per
call (they are the same for every image of a batch); the images are multiplied on the device by
`cryomi_multiply_broadcast` (`ops.multiply_broadcast`).
"""

from __future__ import annotations

import numpy as np

from . import ops
from .coordinates import make_radial_frequency_grid


# ---------------------------------------------------------------------------
# image transforms (cryojax/ndimage/transforms)
# ---------------------------------------------------------------------------


class AbstractImageTransform:
    """An array the image is multiplied by (`_base_transform.py:12-32`)."""

    def get(self):
        raise NotImplementedError

    def __call__(self, image):
        """image * get() for a CUDA tensor of images (..., *get().shape), as a new tensor (the
        multiply runs on the device, cryomi_multiply_broadcast)."""
        return ops.multiply_broadcast(image.clone(), self.get())

    def __mul__(self, other):
        return ProductImageTransform(self, other)

    def __rmul__(self, other):
        return ProductImageTransform(other, self)


class ProductImageTransform(AbstractImageTransform):
    """transform1(transform2(image)) (`_base_transform.py:35-57`): the product of the arrays."""

    def __init__(self, transform1, transform2):
        self.transform1 = transform1
        self.transform2 = transform2

    def get(self):
        return np.asarray(transform_array(self.transform1)) * np.asarray(transform_array(self.transform2))

    def __repr__(self):
        return f"{self.transform1!r} * {self.transform2!r}"


class AbstractFilter(AbstractImageTransform):
    """A filter: an array over the rfft image (`_filters.py:22-33`)."""


class AbstractMask(AbstractImageTransform):
    """A mask: an array over the real image (`_masks.py:18-28`)."""


class CustomFilter(AbstractFilter):
    """A filter given as an array (`_filters.py:39-56`)."""

    def __init__(self, filter):
        self.array = np.asarray(filter)

    def get(self):
        return self.array


class CustomMask(AbstractMask):
    """A mask given as an array (`_masks.py:34-46`)."""

    def __init__(self, mask_array):
        self.array = np.asarray(mask_array)

    def get(self):
        return self.array


def _cosine_edge(a, edge_start, width):
    """0.5 (1 + cos(pi (a - edge_start) / width)): the soft edge of the reference's filters and
    masks, evaluated everywhere (np.where picks where it applies)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return 0.5 * (1.0 + np.cos(np.pi * (a - edge_start) / width))


def _radial_cosine(r, radius, width):
    """1 inside `radius`, a cosine fall-off to 0 over `width`, 0 beyond (_filters.py:193-225,
    _masks.py:325-349): r <= radius -> 1; r > radius + width -> 0."""
    return np.where(r <= radius, 1.0, np.where(r > radius + width, 0.0, _cosine_edge(r, radius, width)))


def _box_cosine(abs_coords, widths, width):
    """The soft-edged box of the square / rectangle masks (_masks.py:352-399, 437-500, 503-580):
    1 inside the box, 0 outside the box grown by `width` on every side, and in between the product
    of the cosine edges of the axes whose coordinate lies strictly inside its edge band; where no
    axis does (a point exactly on the grown box), the reference's selection falls through to the
    edge of the last axis, and so does this one."""
    inside = np.logical_and.reduce([a <= w / 2 for a, w in zip(abs_coords, widths)])
    grown = np.logical_and.reduce([a <= (w + 2 * width) / 2 for a, w in zip(abs_coords, widths)])
    in_edge = [np.logical_and(a > w / 2, a < w / 2 + width) for a, w in zip(abs_coords, widths)]
    edge = [_cosine_edge(a, w / 2, width) for a, w in zip(abs_coords, widths)]
    if len(abs_coords) == 2:
        (ex, ey), (ix, iy) = edge, in_edge
        soft = np.where(ix & iy, ex * ey, np.where(ix, ex, ey))
    else:
        (ex, ey, ez), (ix, iy, iz) = edge, in_edge
        soft = np.where(ix & iy & iz, ex * ey * ez,
                        np.where(ix & iy, ex * ey,
                                 np.where(ix & iz, ex * ez,
                                          np.where(iy & iz, ey * ez, np.where(ix, ex, np.where(iy, ey, ez))))))
    return np.where(inside, 1.0, np.where(grown, soft, 0.0))


def _f64(x):
    return np.asarray(x, dtype=np.float64)


def _lowpass(frequency_grid, grid_spacing, cutoff_fraction, rolloff_fraction):
    """_filters.py:193-225: a cosine-edged disc (ball) of radius cutoff * k_max, k_max the
    Nyquist frequency 1 / (2 grid_spacing)."""
    k_max = 1.0 / (2.0 * _f64(grid_spacing))
    k = np.linalg.norm(_f64(frequency_grid), axis=-1)
    return _radial_cosine(k, _f64(cutoff_fraction) * k_max, _f64(rolloff_fraction) * k_max)


class LowpassFilter(AbstractFilter):
    """A low-pass filter with a cosine soft edge (_filters.py:58-96)."""

    def __init__(self, frequency_grid_in_angstroms_or_pixels, grid_spacing=1.0, frequency_cutoff_fraction=0.95,
                 rolloff_width_fraction=0.05):
        self.array = _lowpass(frequency_grid_in_angstroms_or_pixels, grid_spacing, frequency_cutoff_fraction,
                              rolloff_width_fraction)

    def get(self):
        return self.array


class HighpassFilter(AbstractFilter):
    """1 minus the low-pass filter (_filters.py:99-137)."""

    def __init__(self, frequency_grid_in_angstroms_or_pixels, grid_spacing=1.0, frequency_cutoff_fraction=0.95,
                 rolloff_width_fraction=0.05):
        self.array = 1.0 - _lowpass(frequency_grid_in_angstroms_or_pixels, grid_spacing, frequency_cutoff_fraction,
                                    rolloff_width_fraction)

    def get(self):
        return self.array


def _pad_split(target, size):
    """Left / right padding of _edges.py:175-185: the extra pixel of an odd difference goes right
    when the target is odd and the array even, left in the opposite case."""
    pad = target - size
    if target % 2 == 1 and size % 2 == 0:
        return pad // 2, pad // 2 + 1
    if target % 2 == 0 and size % 2 == 1:
        return pad // 2 + 1, pad // 2
    return pad // 2, pad // 2


def resize_with_crop_or_pad(image, shape, **pad_kwargs):
    """Centred crop and/or pad of a host 2D array (_edges.py:15-62, 110-172): a crop keeps
    [c - h // 2, c + h // 2 + h % 2) around c = n // 2; a pad splits as `_pad_split`; extra keyword
    arguments go to np.pad (mode="edge" and the like)."""
    image = np.asarray(image)
    if image.ndim != 2 or len(shape) != 2:
        raise ValueError(f"resize_with_crop_or_pad can only resize images. Got array shape of {image.shape} "
                         f"and desired shape {shape}.")
    for axis, m in enumerate(shape):
        n = image.shape[axis]
        if n > m:
            c = n // 2
            image = np.take(image, np.arange(c - m // 2, c + m // 2 + m % 2), axis=axis)
        elif n < m:
            width = [(0, 0), (0, 0)]
            width[axis] = _pad_split(m, n)
            image = np.pad(image, width, **pad_kwargs)
    return image


def _whitening_filter(power, image_shape, shape, interpolation_mode, outputs_squared):
    """_filters.py:228-297 on the stack's mean power spectrum (rfft layout, already divided by the
    pixel count): radial average over bins of width 1 / max(rfft shape) up to sqrt(2) / 2
    (_fourier_statistics.py:13-89, 276-280; digitize with right=True, bin means as
    _radial_average.py:9-40), interpolated back onto the grid (:43-86), optionally resized in real
    space with edge padding, then 1 / sqrt (or 1 / x) where the power is not ~0, with the zero
    frequency and, for even sizes, the Nyquist corner set to 0."""
    k = make_radial_frequency_grid(image_shape)
    q_max = np.sqrt(2.0) / 2
    n_bins = 1 + int(q_max / (1.0 / max(*power.shape)))
    bins = np.linspace(0.0, q_max, n_bins)
    idx = np.digitize(k.ravel(), bins, right=True)
    keep = idx < n_bins  # JAX's bincount(length=n_bins) drops larger indices
    with np.errstate(divide="ignore", invalid="ignore"):
        radial = (np.bincount(idx[keep], weights=power.ravel()[keep], minlength=n_bins)
                  / np.bincount(idx[keep], minlength=n_bins))
    if interpolation_mode == "nearest":
        on_grid = np.take(radial, np.digitize(k, bins, right=True), mode="clip")
    elif interpolation_mode == "linear":
        on_grid = np.interp(k.ravel(), bins, radial).reshape(k.shape)
    else:
        raise ValueError(f"`interpolation_mode` = {interpolation_mode} not supported. Supported interpolation "
                         "modes are 'nearest' or 'linear'.")
    if shape is not None:
        new_shape = tuple(shape)
        real = np.fft.fftshift(np.fft.irfftn(on_grid, s=image_shape, axes=(0, 1)))
        on_grid = np.fft.rfftn(np.fft.ifftshift(resize_with_crop_or_pad(real, new_shape, mode="edge"))).real
        on_grid = np.where(on_grid < 0, 0.0, on_grid)
    else:
        new_shape = tuple(image_shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / on_grid if outputs_squared else 1.0 / np.sqrt(on_grid)
    out = np.where(np.isclose(on_grid, 0.0), 0.0, inv)
    out[0, 0] = 0.0
    if new_shape[0] % 2 == 0 and new_shape[1] % 2 == 0:
        out[new_shape[0] // 2, new_shape[1] // 2] = 0.0
    return out


class WhiteningFilter(AbstractFilter):
    """The inverse square root (or, with `outputs_squared`, the inverse) of the radially averaged
    power spectrum of an image or image stack, cisTEM's whitening filter (_filters.py:140-185).

    The stack's transforms and its mean power spectrum are formed on the device
    (cryomi_rfft2_f64); the radial profile of that one (h, w // 2 + 1) array is formed on the host.
    """

    def __init__(self, image_or_image_stack, shape=None, *, interpolation_mode="linear", outputs_squared=False):
        import torch

        x = image_or_image_stack
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x, dtype=np.float64), device="cuda")
        if x.dtype != torch.float64:
            x = x.to(torch.float64)
        if x.dim() == 2:
            x = x[None]
        if x.dim() != 3:
            raise ValueError("WhiteningFilter takes an image (h, w) or an image stack (n, h, w)")
        image_shape = tuple(int(s) for s in x.shape[1:])
        f = ops.rfft2_f64(x.contiguous())
        power = (f.real.square() + f.imag.square()).mean(dim=0) / float(np.prod(image_shape))
        self.array = _whitening_filter(power.cpu().numpy(), image_shape, shape, interpolation_mode, outputs_squared)

    def get(self):
        return self.array


class CircularCosineMask(AbstractMask):
    """A disc with a cosine soft edge, optionally offset (_masks.py:49-81)."""

    def __init__(self, coordinate_grid, radius, rolloff_width, xy_offset=(0.0, 0.0)):
        r = np.linalg.norm(_f64(coordinate_grid) - _f64(xy_offset), axis=-1)
        self.array = _radial_cosine(r, _f64(radius), _f64(rolloff_width))

    def get(self):
        return self.array


class SphericalCosineMask(AbstractMask):
    """A ball with a cosine soft edge (_masks.py:84-115)."""

    def __init__(self, coordinate_grid, radius, rolloff_width):
        r = np.linalg.norm(_f64(coordinate_grid), axis=-1)
        self.array = _radial_cosine(r, _f64(radius), _f64(rolloff_width))

    def get(self):
        return self.array


class SquareCosineMask(AbstractMask):
    """A square of side `side_length` with a cosine soft edge (_masks.py:118-146)."""

    def __init__(self, coordinate_grid, side_length, rolloff_width):
        g = _f64(coordinate_grid)
        s = _f64(side_length)
        self.array = _box_cosine((np.abs(g[..., 0]), np.abs(g[..., 1])), (s, s), _f64(rolloff_width))

    def get(self):
        return self.array


def _rotated(coordinate_grid, angle):
    """(x cos + y sin, -x sin + y cos) for an in-plane angle in degrees."""
    g = _f64(coordinate_grid)
    t = np.deg2rad(_f64(angle))
    c, s = np.cos(t), np.sin(t)
    x, y = g[..., 0], g[..., 1]
    return x * c + y * s, -x * s + y * c


def _cylinder_2d(coordinate_grid, radius, length, angle, rolloff_width):
    """_masks.py:437-500: a band of width 2 radius along the rotated x axis, `length` long,
    soft-edged like the boxes."""
    xr, yr = _rotated(coordinate_grid, angle)
    return _box_cosine((np.abs(yr), np.abs(xr)), (2 * _f64(radius), _f64(length)), _f64(rolloff_width))


class Cylindrical2DCosineMask(AbstractMask):
    """An in-plane cylinder (a band) rotated by `in_plane_rotation_angle` degrees, infinite or
    `length` long, with a cosine soft edge (_masks.py:149-198, 402-434)."""

    def __init__(self, coordinate_grid, radius, rolloff_width, in_plane_rotation_angle=0.0, length=None):
        if length is None:
            _, yr = _rotated(coordinate_grid, in_plane_rotation_angle)
            r, rad, w = np.abs(yr), _f64(radius), _f64(rolloff_width)
            self.array = np.where(r < rad, 1.0, np.where(r < rad + w, _cosine_edge(r, rad, w), 0.0))
        else:
            self.array = _cylinder_2d(coordinate_grid, radius, length, in_plane_rotation_angle, rolloff_width)

    def get(self):
        return self.array


class Rectangular2DCosineMask(AbstractMask):
    """An x_width by y_width rectangle rotated in plane, with a cosine soft edge
    (_masks.py:201-240)."""

    def __init__(self, coordinate_grid, x_width, y_width, rolloff_width, in_plane_rotation_angle=0.0):
        self.array = _cylinder_2d(coordinate_grid, _f64(y_width) / 2, x_width, in_plane_rotation_angle,
                                  rolloff_width)

    def get(self):
        return self.array


class Rectangular3DCosineMask(AbstractMask):
    """A box of widths (x, y, z) with a cosine soft edge (_masks.py:243-281)."""

    def __init__(self, coordinate_grid, x_width, y_width, z_width, rolloff_width):
        g = _f64(coordinate_grid)
        self.array = _box_cosine(tuple(np.abs(g[..., i]) for i in range(3)),
                                 (_f64(x_width), _f64(y_width), _f64(z_width)), _f64(rolloff_width))

    def get(self):
        return self.array


class InverseSincMask(AbstractMask):
    """1 / prod_i sinc(x_i / n_i) on the pixel coordinates of an image or volume, n_i the box
    size along the axis of x_i (_masks.py:284-318)."""

    def __init__(self, coordinate_grid_in_pixels):
        g = _f64(coordinate_grid_in_pixels)
        ndim = g.ndim - 1
        box = g.shape[:ndim][::-1]  # x, y(, z) sizes
        prod = np.ones(g.shape[:-1])
        for i in range(ndim):
            prod = prod * np.sinc(g[..., i] / box[i])
        with np.errstate(divide="ignore"):
            self.array = 1.0 / prod

    def get(self):
        return self.array


def transform_array(transform):
    """The array a filter or mask multiplies by: `get()`, an `array` attribute, or an array."""
    if hasattr(transform, "get"):
        return transform.get()
    if hasattr(transform, "array"):
        return transform.array
    if isinstance(transform, np.ndarray) or hasattr(transform, "__array__"):
        return transform
    raise NotImplementedError(
        f"{type(transform).__name__}: only transforms that multiply by an array (get() / array) run on the device"
    )


# ---------------------------------------------------------------------------
# Fourier operators (cryojax/ndimage/operators)
# ---------------------------------------------------------------------------


class AbstractImageOperator:
    """A function of a frequency grid (`_base_operator.py:15-50`)."""

    def __call__(self, frequency_grid):
        raise NotImplementedError

    def __add__(self, other):
        return SumImageOperator(self, other if isinstance(other, AbstractImageOperator) else Constant(other))

    def __radd__(self, other):
        return SumImageOperator(other if isinstance(other, AbstractImageOperator) else Constant(other), self)

    def __sub__(self, other):
        return DiffImageOperator(self, other if isinstance(other, AbstractImageOperator) else Constant(other))

    def __rsub__(self, other):
        return DiffImageOperator(other if isinstance(other, AbstractImageOperator) else Constant(other), self)

    def __mul__(self, other):
        return ProductImageOperator(self, other if isinstance(other, AbstractImageOperator) else Constant(other))

    def __rmul__(self, other):
        return ProductImageOperator(other if isinstance(other, AbstractImageOperator) else Constant(other), self)


AbstractFourierOperator = AbstractImageOperator


class Constant(AbstractImageOperator):
    """A constant value (`_base_operator.py:53-67`)."""

    def __init__(self, value):
        self.value = np.asarray(value)

    def __call__(self, *args, **kwargs):
        return self.value


class CustomOperator(AbstractImageOperator):
    """fn(frequency_grid, **kwargs) (`_base_operator.py:70-94`), evaluated on the host."""

    def __init__(self, fn, **kwargs):
        self.fn = fn
        self.kwargs = kwargs

    def __call__(self, *args, **kwargs):
        return self.fn(*args, **kwargs, **self.kwargs)


class Empirical(AbstractImageOperator):
    """A stored array (`_base_operator.py:97-125`)."""

    def __init__(self, array):
        self.array = np.asarray(array)

    def __call__(self, *args, **kwargs):
        return self.array


class SumImageOperator(AbstractImageOperator):
    def __init__(self, operator1, operator2):
        self.operator1, self.operator2 = operator1, operator2

    def __call__(self, frequency_grid):
        return self.operator1(frequency_grid) + self.operator2(frequency_grid)


class DiffImageOperator(AbstractImageOperator):
    def __init__(self, operator1, operator2):
        self.operator1, self.operator2 = operator1, operator2

    def __call__(self, frequency_grid):
        return self.operator1(frequency_grid) - self.operator2(frequency_grid)


class ProductImageOperator(AbstractImageOperator):
    def __init__(self, operator1, operator2):
        self.operator1, self.operator2 = operator1, operator2

    def __call__(self, frequency_grid):
        return self.operator1(frequency_grid) * self.operator2(frequency_grid)


def _k_sqr(frequency_grid):
    return np.sum(np.asarray(frequency_grid, dtype=np.float64) ** 2, axis=-1)


def _per_image(param, frequency_grid):
    """A parameter with a batch shape (one value per particle, as a vmapped operator holds it)
    gets one trailing axis per grid dimension, so it broadcasts over each image's frequencies."""
    p = np.asarray(param)
    return p.reshape(p.shape + (1,) * np.shape(frequency_grid)[-1]) if p.ndim else p


class FourierGaussian(AbstractImageOperator):
    """amplitude * exp(-b_factor k^2 / 4) (`_fourier_operator.py:183-234`)."""

    def __init__(self, amplitude=1.0, b_factor=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        b = np.asarray(b_factor, dtype=np.float64)
        if np.any(b <= 0):  # error_if_not_positive
            raise ValueError("`b_factor` must be positive")
        self.b_factor = b

    def __call__(self, frequency_grid):
        a, b = _per_image(self.amplitude, frequency_grid), _per_image(self.b_factor, frequency_grid)
        return a * np.exp(-0.25 * b * _k_sqr(frequency_grid))


class Lorenzian(AbstractImageOperator):
    """amplitude / length_scale^2 / (k^2 + 1 / length_scale^2) (`_fourier_operator.py:132-180`)."""

    def __init__(self, amplitude=1.0, length_scale=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        self.length_scale = np.asarray(length_scale, dtype=np.float64)

    def __call__(self, frequency_grid):
        a, ls = _per_image(self.amplitude, frequency_grid), _per_image(self.length_scale, frequency_grid)
        return a / ls**2 / (_k_sqr(frequency_grid) + 1.0 / ls**2)


class FourierExp2D(AbstractImageOperator):
    """amplitude / (2 pi length_scale^3) / (k^2 + 1 / length_scale^2)^1.5 (`_fourier_operator.py:84-129`)."""

    def __init__(self, amplitude=1.0, length_scale=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        self.length_scale = np.asarray(length_scale, dtype=np.float64)

    def __call__(self, frequency_grid):
        a, ls = _per_image(self.amplitude, frequency_grid), _per_image(self.length_scale, frequency_grid)
        return a / (2.0 * np.pi * ls**3) / (_k_sqr(frequency_grid) + 1.0 / ls**2) ** 1.5


class FourierGaussianWithRadialOffset(AbstractImageOperator):
    """amplitude * exp(-b_factor (|k| - offset)^2 / 4) (`_fourier_operator.py:237-293`)."""

    def __init__(self, amplitude=1.0, b_factor=1.0, offset=0.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        b = np.asarray(b_factor, dtype=np.float64)
        if np.any(b <= 0):  # error_if_not_positive
            raise ValueError("`b_factor` must be positive")
        m = np.asarray(offset, dtype=np.float64)
        if np.any(m < 0):  # error_if_negative
            raise ValueError("`offset` must be non-negative")
        self.b_factor, self.offset = b, m

    def __call__(self, frequency_grid):
        a, b, m = (_per_image(v, frequency_grid) for v in (self.amplitude, self.b_factor, self.offset))
        k = np.sqrt(_k_sqr(frequency_grid))
        return a * np.exp(-0.25 * b * (k - m) ** 2)


class FourierSinc(AbstractImageOperator):
    """prod_i sinc(q_i box_width), the transform of a box of width `box_width`
    (`_fourier_operator.py:296-336`)."""

    def __init__(self, box_width=1.0):
        self.box_width = np.asarray(box_width, dtype=np.float64)

    def __call__(self, frequency_grid):
        f = np.asarray(frequency_grid, dtype=np.float64)
        w = _per_image(self.box_width, f)
        out = np.ones(f.shape[:-1])
        for i in range(f.shape[-1]):
            out = out * np.sinc(f[..., i] * w)
        return out


class Gaussian2D(AbstractImageOperator):
    """amplitude / sqrt(2 pi variance) * exp(-|r - offset|^2 / (2 variance)) on a real-space
    coordinate grid (`_real_operator.py:56-108`)."""

    def __init__(self, amplitude=1.0, variance=1.0, offset=(0.0, 0.0)):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        v = np.asarray(variance, dtype=np.float64)
        if np.any(v <= 0):  # error_if_not_positive
            raise ValueError("`variance` must be positive")
        self.variance = v
        self.offset = np.asarray(offset, dtype=np.float64)

    def __call__(self, coordinate_grid):
        r_sqr = np.sum((np.asarray(coordinate_grid, dtype=np.float64) - self.offset) ** 2, axis=-1)
        return self.amplitude / np.sqrt(2 * np.pi * self.variance) * np.exp(-0.5 * r_sqr / self.variance)


class ZeroMode(AbstractImageOperator):
    """`value` at the zero frequency, zero elsewhere (`_fourier_operator.py:64-81`)."""

    def __init__(self, value=0.0):
        self.value = np.asarray(value, dtype=np.float64)

    def __call__(self, frequency_grid):
        out = np.zeros(np.asarray(frequency_grid).shape[:-1])
        out[0, 0] = self.value
        return out
