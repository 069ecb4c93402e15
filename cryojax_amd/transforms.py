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


# ---------------------------------------------------------------------------
# image transforms (cryojax/ndimage/transforms)
# ---------------------------------------------------------------------------


class AbstractImageTransform:
    """An array the image is multiplied by (`_base_transform.py:12-32`)."""

    def get(self):
        raise NotImplementedError

    def __call__(self, image):
        """image * get() for a CUDA tensor of images (..., *get().shape), in place."""
        return ops.multiply_broadcast(image, self.get())

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


class FourierGaussian(AbstractImageOperator):
    """amplitude * exp(-b_factor k^2 / 4) (`_fourier_operator.py:183-234`)."""

    def __init__(self, amplitude=1.0, b_factor=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        b = np.asarray(b_factor, dtype=np.float64)
        if np.any(b <= 0):  # error_if_not_positive
            raise ValueError("`b_factor` must be positive")
        self.b_factor = b

    def __call__(self, frequency_grid):
        return self.amplitude * np.exp(-0.25 * self.b_factor * _k_sqr(frequency_grid))


class Lorenzian(AbstractImageOperator):
    """amplitude / length_scale^2 / (k^2 + 1 / length_scale^2) (`_fourier_operator.py:132-180`)."""

    def __init__(self, amplitude=1.0, length_scale=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        self.length_scale = np.asarray(length_scale, dtype=np.float64)

    def __call__(self, frequency_grid):
        ls = self.length_scale
        return self.amplitude / ls**2 / (_k_sqr(frequency_grid) + 1.0 / ls**2)


class FourierExp2D(AbstractImageOperator):
    """amplitude / (2 pi length_scale^3) / (k^2 + 1 / length_scale^2)^1.5 (`_fourier_operator.py:84-129`)."""

    def __init__(self, amplitude=1.0, length_scale=1.0):
        self.amplitude = np.asarray(amplitude, dtype=np.float64)
        self.length_scale = np.asarray(length_scale, dtype=np.float64)

    def __call__(self, frequency_grid):
        ls = self.length_scale
        return self.amplitude / (2.0 * np.pi * ls**3) / (_k_sqr(frequency_grid) + 1.0 / ls**2) ** 1.5


class ZeroMode(AbstractImageOperator):
    """`value` at the zero frequency, zero elsewhere (`_fourier_operator.py:64-81`)."""

    def __init__(self, value=0.0):
        self.value = np.asarray(value, dtype=np.float64)

    def __call__(self, frequency_grid):
        out = np.zeros(np.asarray(frequency_grid).shape[:-1])
        out[0, 0] = self.value
        return out
