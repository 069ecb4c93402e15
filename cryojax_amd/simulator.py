"""Drop-in mirror of cryoJAX's Fourier-voxel image-formation API on MI355X.

Same class names, constructor arguments, defaults and error behaviour as
`cryojax.simulator` for the Fourier-slice path; every image is formed by the
HIP kernels behind include/cryomi.h. JAX's `vmap` over pose / CTF leaves is
expressed by passing arrays with one leading batch axis: `simulate()` then
returns a batch of images from one fused launch.
"""

from __future__ import annotations

import math
from typing import Literal

import numpy as np
import torch

from . import _lib, ops
from .transforms import transform_array as _transform_array
from .rotations import SO3, convert_quaternion_to_euler_angles

__all__ = [
    "AstigmaticCTF",
    "BasicImageConfig",
    "ContrastTransferTheory",
    "EulerAnglePose",
    "FourierSliceExtraction",
    "FourierVoxelGridVolume",
    "LinearImageModel",
    "ProjectionImageModel",
    "QuaternionPose",
    "make_image_model",
    "wavelength_from_kilovolts",
]


def wavelength_from_kilovolts(voltage_in_kilovolts):
    """cryojax/constants/_physical_constants.py:10-29."""
    v = 1000.0 * np.asarray(voltage_in_kilovolts, dtype=np.float64)
    return 12.2639 / (v + 0.97845e-6 * v**2) ** 0.5


def _f64(x):
    return np.asarray(x, dtype=np.float64)


def _batch_size(*arrays):
    shapes = [np.shape(a) for a in arrays if a is not None]
    lead = [s[0] for s in shapes if len(s) > 0]
    if not lead:
        return None
    n = max(lead)
    for s in lead:
        if s not in (1, n):
            raise ValueError(f"inconsistent batch sizes {sorted(set(lead))}")
    return n


# ---------------------------------------------------------------------------
# poses, cryojax/simulator/_pose.py
# ---------------------------------------------------------------------------


class AbstractPose:
    offset_in_angstroms: np.ndarray

    @property
    def rotation(self) -> SO3:
        raise NotImplementedError

    @property
    def offset_x_in_angstroms(self):
        return self.offset_in_angstroms[..., 0]

    @property
    def offset_y_in_angstroms(self):
        return self.offset_in_angstroms[..., 1]

    @property
    def offset_z_in_angstroms(self):
        """_pose.py:132-142: None for 2D offsets."""
        return None if self.offset_in_angstroms.shape[-1] == 2 else self.offset_in_angstroms[..., 2]

    def rotate_coordinates(self, coordinate_grid_or_list, inverse=False):
        """_pose.py:29-66 on host arrays (..., 3)."""
        rot = self.rotation.inverse() if inverse else self.rotation
        c = _f64(coordinate_grid_or_list)
        if c.ndim not in (2, 4) or c.shape[-1] != 3:
            raise ValueError(
                "Coordinates must be an array either of shape (N, 3) or "
                f"(N1, N2, N3, 3). Instead, got {c.shape}."
            )
        return rot.apply(c)

    def compute_translation_operator(self, frequency_grid_in_angstroms):
        """_pose.py:97-114 (host)."""
        xy = self.offset_in_angstroms[..., 0:2]
        f = _f64(frequency_grid_in_angstroms)
        return np.exp(-1.0j * (2 * np.pi * (f @ xy)))

    @classmethod
    def from_rotation(cls, rotation: SO3):
        raise NotImplementedError

    @classmethod
    def from_rotation_and_translation(cls, rotation: SO3, offset_in_angstroms):
        off = _f64(offset_in_angstroms)
        if off.shape[-1:] not in [(2,), (3,)]:
            raise ValueError(
                "Array `offset_in_angstroms` given to constructor "
                f"`{cls.__name__}.from_rotation_and_translation` supports "
                f"shapes `(2,)` and `(3,)`. Got shape `{off.shape}`"
            )
        pose = cls.from_rotation(rotation)
        pose.offset_in_angstroms = off
        return pose

    @classmethod
    def from_translation(cls, offset_in_angstroms):
        pose = cls.from_rotation(SO3(np.array([1.0, 0.0, 0.0, 0.0])))
        pose.offset_in_angstroms = _f64(offset_in_angstroms)
        return pose

    def to_inverse_rotation(self):
        return self.from_rotation_and_translation(self.rotation.inverse(), self.offset_in_angstroms)


def _stack_offsets(x, y, z):
    parts = [x, y] if z is None else [x, y, z]
    return np.stack(np.broadcast_arrays(*[_f64(p) for p in parts]), axis=-1)


class EulerAnglePose(AbstractPose):
    """zyz extrinsic Euler angles in degrees, _pose.py:203-280."""

    def __init__(
        self,
        offset_x_in_angstroms=0.0,
        offset_y_in_angstroms=0.0,
        phi_angle=0.0,
        theta_angle=0.0,
        psi_angle=0.0,
        *,
        offset_z_in_angstroms=None,
    ):
        self.offset_in_angstroms = _stack_offsets(
            offset_x_in_angstroms, offset_y_in_angstroms, offset_z_in_angstroms
        )
        self.phi_angle = _f64(phi_angle)
        self.theta_angle = _f64(theta_angle)
        self.psi_angle = _f64(psi_angle)

    @property
    def rotation(self) -> SO3:
        """R3 @ R2 @ R1, _pose.py:257-272."""
        r1 = SO3.from_z_radians(np.deg2rad(self.phi_angle))
        r2 = SO3.from_y_radians(np.deg2rad(self.theta_angle))
        r3 = SO3.from_z_radians(np.deg2rad(self.psi_angle))
        return r3 @ r2 @ r1

    @classmethod
    def from_rotation(cls, rotation: SO3):
        ang = convert_quaternion_to_euler_angles(rotation.wxyz, "zyz", True)
        return cls(phi_angle=ang[..., 0], theta_angle=ang[..., 1], psi_angle=ang[..., 2])


class QuaternionPose(AbstractPose):
    """Unit-quaternion pose, _pose.py:283-335."""

    def __init__(
        self,
        offset_x_in_angstroms=0.0,
        offset_y_in_angstroms=0.0,
        wxyz=(1.0, 0.0, 0.0, 0.0),
        *,
        offset_z_in_angstroms=None,
    ):
        self.offset_in_angstroms = _stack_offsets(
            offset_x_in_angstroms, offset_y_in_angstroms, offset_z_in_angstroms
        )
        w = _f64(wxyz)
        if w.shape[-1] != 4:
            raise ValueError(
                "Expected `wxyz` to be a sequence of floats with "
                f"length 4, but found `len(wxyz) = {w.shape[-1]}`."
            )
        self.wxyz = w

    @property
    def rotation(self) -> SO3:
        return SO3(self.wxyz).normalize()

    @classmethod
    def from_rotation(cls, rotation: SO3):
        return cls(wxyz=rotation.wxyz)


# ---------------------------------------------------------------------------
# CTF / transfer theory, cryojax/simulator/_transfer_theory
# ---------------------------------------------------------------------------


class AstigmaticCTF:
    """transfer_function.py:84-169."""

    def __init__(
        self,
        defocus_in_angstroms=10000.0,
        astigmatism_in_angstroms=0.0,
        astigmatism_angle=0.0,
        spherical_aberration_in_mm=2.7,
    ):
        self.defocus_in_angstroms = _f64(defocus_in_angstroms)
        self.astigmatism_in_angstroms = _f64(astigmatism_in_angstroms)
        self.astigmatism_angle = _f64(astigmatism_angle)
        cs = _f64(spherical_aberration_in_mm)
        if np.any(cs < 0):  # error_if_negative, transfer_function.py:126-128
            raise ValueError("`spherical_aberration_in_mm` must be non-negative")
        self.spherical_aberration_in_mm = cs


class ContrastTransferTheory:
    """transfer_theory.py:21-120 (projection-approximation branch)."""

    def __init__(self, ctf, envelope=None, amplitude_contrast_ratio=0.1, phase_shift=0.0):
        """`envelope`: a Fourier operator (`cryojax_amd.transforms`, or any callable of the
        frequency grid in inverse Angstroms returning an array of the rfft image's shape) that
        multiplies the contrast spectrum (transfer_theory.py:117-118)."""
        ac = _f64(amplitude_contrast_ratio)
        if np.any((ac < 0) | (ac > 1)):  # error_if_not_fractional, transfer_theory.py:50
            raise ValueError("`amplitude_contrast_ratio` must be in [0, 1]")
        self.ctf = ctf
        self.envelope = envelope
        self.amplitude_contrast_ratio = ac
        self.phase_shift = _f64(phase_shift)

    def total_phase_shift(self):
        """deg2rad(phase_shift) + atan(ac / sqrt(1 - ac^2)), transfer_function.py:66-77."""
        ac = self.amplitude_contrast_ratio
        with np.errstate(divide="ignore"):
            acp = np.arctan(ac / np.sqrt(1.0 - ac**2))
        return np.deg2rad(self.phase_shift) + acp


# ---------------------------------------------------------------------------
# image config, cryojax/simulator/_image_config.py
# ---------------------------------------------------------------------------


class BasicImageConfig:
    """_image_config.py:280-338."""

    def __init__(self, shape, pixel_size, voltage_in_kilovolts, *, grid_helper=None, pad_options=None):
        ps = _f64(pixel_size)
        kv = _f64(voltage_in_kilovolts)
        if np.any(ps <= 0) or np.any(kv <= 0):  # error_if_not_positive, :329-332
            raise ValueError("`pixel_size` and `voltage_in_kilovolts` must be positive")
        pad_options = dict(pad_options or {})
        self.shape = tuple(int(s) for s in shape)
        self.pixel_size = ps
        self.voltage_in_kilovolts = kv
        self.grid_helper = grid_helper
        self.pad_options = {
            "shape": tuple(pad_options.get("shape", self.shape)),
            "mode": pad_options.get("mode", "constant"),
            "grid_helper": pad_options.get("grid_helper", None),
        }
        ps_ = self.padded_shape
        if ps_[0] < self.shape[0] or ps_[1] < self.shape[1]:
            raise AttributeError("`padded_shape` is less than `shape` in one or more dimensions.")

    @property
    def padded_shape(self):
        return self.pad_options["shape"]

    @property
    def pad_mode(self):
        return self.pad_options["mode"]

    def crop_to_shape(self, image):
        """Crop images (..., h, w) on device to `shape` around their centre (_image_config.py:225-229)."""
        return ops.resize_images(image, self.shape)

    @property
    def wavelength_in_angstroms(self):
        return wavelength_from_kilovolts(self.voltage_in_kilovolts)

    @property
    def padded_frequency_grid_in_pixels(self):
        """(padded_y, padded_x // 2 + 1, 2) frequencies in cycles per pixel: [..., 0] the rfftfreq
        of x, [..., 1] the fftfreq of y (_image_config.py:177-192, make_frequency_grid)."""
        py, px = self.padded_shape
        fx = np.fft.rfftfreq(px)
        fy = np.fft.fftfreq(py)
        return np.stack(np.broadcast_arrays(fx[None, :], fy[:, None]), axis=-1)

    @property
    def padded_frequency_grid_in_angstroms(self):
        """padded_frequency_grid_in_pixels / pixel_size with exact zeros kept
        (_image_config.py:194-201, _safe_multiply_by_constant :406-412). A batch of pixel sizes
        (a config broadcast over particles) gives one grid per image: (*batch, y, x // 2 + 1, 2)."""
        ps = _f64(self.pixel_size)
        g = self.padded_frequency_grid_in_pixels
        if ps.ndim:
            g = np.broadcast_to(g, ps.shape + g.shape)
            ps = ps[..., None, None, None]
        return np.where(g != 0.0, g / ps, 0.0)

    @property
    def padded_rfft_shape(self):
        py, px = self.padded_shape
        return (py, px // 2 + 1)

    @property
    def n_pixels(self):
        return math.prod(self.shape)

    @property
    def y_dim(self):
        return self.shape[0]

    @property
    def x_dim(self):
        return self.shape[1]


# ---------------------------------------------------------------------------
# volume, cryojax/simulator/_volume/representations/voxel_volume.py
# ---------------------------------------------------------------------------


def _frequency_slice(n):
    """make_frequency_slice((n, n), outputs_rfftfreqs=False), coordinates/_make_coordinate_grids.py:139-198."""
    f = np.fft.fftshift(np.fft.fftfreq(n))
    xv, yv = np.meshgrid(f, f, indexing="xy")
    return np.stack([xv, yv, np.zeros_like(xv)], axis=-1)[None]


class FourierVoxelGridVolume:
    """Centred complex64 Fourier grid resident in HBM, voxel_volume.py:40-119.

    `rotate_to_pose` records the pose rotation instead of materialising the
    rotated (1, N, N, 3) plane; the kernels rotate each sample on the fly.
    """

    def __init__(self, fourier_voxel_grid, frequency_slice_in_pixels=None, *, device=None, _rotation=None):
        g = fourier_voxel_grid
        if not isinstance(g, torch.Tensor):
            g = torch.as_tensor(np.asarray(g))
        if device is None:
            device = g.device if g.device.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
        self.fourier_voxel_grid = g.to(device=device, dtype=torch.complex64).contiguous()
        self._bricked = None
        self._slice = frequency_slice_in_pixels
        self._rotation = _rotation

    @property
    def shape(self):
        return tuple(self.fourier_voxel_grid.shape)

    @property
    def frequency_slice_in_pixels(self):
        """The (1, N, N, 3) plane, rotated by the recorded pose (host, float64)."""
        base = np.asarray(self._slice if self._slice is not None else _frequency_slice(self.shape[1]))
        if self._rotation is None:
            return base
        q = self._rotation.wxyz
        if q.ndim == 1:
            return self._rotation.apply(base)
        return np.stack([SO3(qi).apply(base) for qi in q.reshape(-1, 4)])

    @classmethod
    def from_real_voxel_grid(cls, real_voxel_grid, *, pad_scale=1.0, pad_mode="constant", filter=None, device=None):
        """fftshift(filter(fftn(ifftshift(pad_to_shape(x))))), voxel_volume.py:68-119.

        The 3D transform runs on device in FP64 (cryomi_fourier_voxel_grid,
        zero padding folded into the operators); the grid is kept in HBM as
        complex64 for image formation and complex128 for the pose search's
        FP64 re-scoring.
        """
        if pad_scale < 1.0:
            raise ValueError("`pad_scale` must be greater than 1.0")
        x = real_voxel_grid if isinstance(real_voxel_grid, torch.Tensor) else torch.as_tensor(
            np.asarray(real_voxel_grid))
        if x.ndim != 3:
            raise ValueError("real_voxel_grid must have shape (d, h, w)")
        padded = tuple(int(s * pad_scale) for s in x.shape)
        if padded != tuple(x.shape) and pad_mode != "constant":
            raise NotImplementedError("only pad_mode='constant' (zero padding) is implemented")
        if device is None:
            device = x.device if x.device.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
        if x.dtype not in (torch.float32, torch.float64):
            x = x.to(torch.float64)
        g128, g64 = ops.fourier_voxel_grid(x.to(device).contiguous(), padded)
        if filter is not None:
            # filter(fftn(x)) with the zero frequency in the corner, then fftshift
            # (voxel_volume.py:105-112): the centred grids times fftshift(filter)
            f = np.fft.fftshift(np.asarray(_transform_array(filter)))
            if f.shape != tuple(padded):
                raise ValueError(f"filter of shape {f.shape} does not match the padded grid {tuple(padded)}")
            ops.multiply_broadcast(g128, f)
            ops.multiply_broadcast(g64, f)
        vol = cls(g64, _frequency_slice(padded[1]), device=device)
        vol._grid_f64 = g128
        return vol

    @property
    def fourier_voxel_grid_f64(self):
        """complex128 grid in HBM (the pose search's FP64 re-scoring reads it):
        the FP64 transform for volumes from from_real_voxel_grid, else the
        complex64 grid widened."""
        if getattr(self, "_grid_f64", None) is None:
            self._grid_f64 = self.fourier_voxel_grid.to(torch.complex128)
        return self._grid_f64

    def bricked_grid(self):
        """The grid in the gather's brick layout (built once, kept in HBM)."""
        if getattr(self, "_bricked", None) is None:
            self._bricked = ops.grid_to_bricks(self.fourier_voxel_grid)
        return self._bricked

    def rotate_to_pose(self, pose, inverse=False):
        rot = pose.rotation.inverse() if inverse else pose.rotation
        if getattr(self, "_bricked", None) is None and self.shape[0] % 4 == 0:
            self.bricked_grid()
        out = FourierVoxelGridVolume.__new__(FourierVoxelGridVolume)
        out.fourier_voxel_grid = self.fourier_voxel_grid
        out._bricked = getattr(self, "_bricked", None)
        out._grid_f64 = getattr(self, "_grid_f64", None)
        out._slice = self._slice
        out._rotation = rot
        return out

    def get_representation(self, rng_key=None):
        return self


# ---------------------------------------------------------------------------
# integrator, cryojax/simulator/_volume_integrator/fourier_voxel_extract.py
# ---------------------------------------------------------------------------


class FourierSliceExtraction:
    """Linear-interpolation Fourier slice extraction, fourier_voxel_extract.py:26-222.

    `out_of_bounds_mode="fill"` reads each trilinear corner as the reference's
    `input.at[indices].get(mode="fill", fill_value=0)` (ndimage/_map_coordinates.py:117)
    does under JAX, which normalises negative indices first: a corner index in
    [-N, -1] wraps to i + N, any other index outside [0, N) reads 0.
    `wrap_negative_indices=False` (an addition of this package, named after the
    JAX `.at[].get` argument with that meaning) makes every index outside [0, N)
    read 0 instead.
    """

    is_projection_approximation = True

    def __init__(self, *, outputs_integral=True, correction_mask=None, out_of_bounds_mode="fill",
                 fill_value=0.0 + 0.0j, wrap_negative_indices=True):
        if correction_mask is not None:
            raise NotImplementedError("sinc correction masks are not implemented by the fused kernel")
        if out_of_bounds_mode != "fill" or fill_value != 0:
            raise NotImplementedError("only out_of_bounds_mode='fill' with fill_value=0 is implemented")
        self.outputs_integral = outputs_integral
        self.correction_mask = None
        self.out_of_bounds_mode = out_of_bounds_mode
        self.fill_value = fill_value
        self.wrap_negative_indices = bool(wrap_negative_indices)

    @property
    def gather_flags(self):
        """The kernel flag of this integrator's out-of-range corner rule."""
        return 0 if self.wrap_negative_indices else _lib.GATHER_FILL

    def integrate(self, volume_representation, image_config, outputs_real_space=False):
        """fourier_voxel_extract.py:77-146: rotated slice (rfft layout) or its irfftn."""
        vol = volume_representation
        if not isinstance(vol, FourierVoxelGridVolume):
            raise ValueError(
                "Supported types for `volume_representation` are "
                "`FourierVoxelGridVolume` and FourierVoxelSplineVolume`."
            )
        n = vol.shape[1]
        if vol.shape != (n, n, n):
            raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
        rot = vol._rotation if vol._rotation is not None else SO3.identity()
        flags = (0 if outputs_real_space else _lib.OUTPUT_FOURIER) | self.gather_flags
        p = _pack(rot, None, None, image_config, self, n)
        out = _run(vol, image_config, p, flags)
        return out


def _check_image_config(image_config, n):
    """The kernels form square images at padded_shape: the volume's box (the slice itself) or any
    other supported box (the integrator's resize branch); shape may be anything up to it."""
    ph, pw = image_config.padded_shape
    if ph != pw:
        raise NotImplementedError(f"only square padded shapes are implemented, got {image_config.padded_shape}")
    if image_config.pad_mode != "constant":
        raise NotImplementedError("only pad_options mode='constant' is implemented")
    if not _lib.box_supported(ph):
        raise NotImplementedError(f"padded box {ph} is not supported (even, 8..512, 2^a 3^b 5^c)")


def _pack(rot, pose, transfer_theory, image_config, integrator, n):
    """Per-image cryomi_image_params table for the batch these leaves describe."""
    mats = rot.sandwich_matrix()
    leaves = [mats[..., 0, 0], image_config.pixel_size, image_config.voltage_in_kilovolts]
    off = np.zeros(2) if pose is None else pose.offset_in_angstroms
    leaves.append(off[..., 0])
    defocus = astig = angle = cs = phase = 0.0
    if transfer_theory is not None:
        ctf = transfer_theory.ctf
        defocus = ctf.defocus_in_angstroms
        if pose is not None and pose.offset_z_in_angstroms is not None:
            defocus = defocus + pose.offset_z_in_angstroms
        astig = ctf.astigmatism_in_angstroms
        angle = ctf.astigmatism_angle
        cs = ctf.spherical_aberration_in_mm
        phase = transfer_theory.total_phase_shift()
        leaves += [defocus, astig, angle, cs, phase]
    b = _batch_size(*leaves)
    nb = 1 if b is None else b
    amp = image_config.pixel_size if integrator.outputs_integral else 1.0
    p = _lib.pack_params(
        mats.reshape(-1, 9),
        np.asarray(off, dtype=np.float64).reshape(-1, off.shape[-1])[:, :2],
        defocus,
        astig,
        angle,
        cs,
        wavelength_from_kilovolts(image_config.voltage_in_kilovolts),
        image_config.pixel_size,
        phase,
        amp,
        nb,
    )
    return p, b


def _slice_images(vol, params_dev, flags):
    n = vol.shape[1]
    if n % 4 == 0:
        return ops.simulate_images(vol.bricked_grid(), params_dev, flags | _lib.GRID_BRICKED, box=n)
    return ops.simulate_images(vol.fourier_voxel_grid, params_dev, flags)


def _run(vol, image_config, packed, flags):
    """compute_fourier_image (with OUTPUT_FOURIER) or its irfftn at padded_shape, per image.

    padded_shape equal to the volume's box: one fused launch. Otherwise the integrator's resize
    branch (fourier_voxel_extract.py:132-138): the bare slice at the box, its irfftn, a centred
    crop or zero pad to padded_shape and an rfftn, then the fix-ups, CTF and shift at padded_shape
    (cryomi_simulate_from_spectrum).
    """
    p, b = packed
    n = vol.shape[1]
    _check_image_config(image_config, n)
    dev = vol.fourier_voxel_grid.device
    pbox = image_config.padded_shape[0]
    if pbox == n:
        out = _slice_images(vol, ops.params_to_device(p, dev), flags)
    else:
        bare = np.array(p, copy=True)
        bare[:, _lib.PARAM_FIELDS.index("amplitude")] = 1.0  # the integral's pixel-size factor comes after
        slices = _slice_images(vol, ops.params_to_device(bare, dev), _lib.OUTPUT_FOURIER | (flags & _lib.GATHER_FILL))
        resized = ops.resize_images(ops.irfft2(slices, n), (pbox, pbox))
        spec = ops.rfft2(resized)
        keep = flags & (_lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.OUTPUT_FOURIER)
        out = ops.simulate_from_spectrum(spec, ops.params_to_device(p, dev), keep)
    return out if b is not None else out[0]


# ---------------------------------------------------------------------------
# image models, cryojax/simulator/_image_model/base_image_model.py
# ---------------------------------------------------------------------------


class AbstractImageModel:
    normalizes_signal: bool

    def _flags(self, outputs_real_space):
        raise NotImplementedError

    def simulate(self, rng_key=None, *, removes_padding=True, outputs_real_space=True, mask=None, filter=None):
        """base_image_model.py:68-105 and _maybe_postprocess (:194-213).

        The image is formed at padded_shape. With removes_padding, postprocess filters the rfft
        image, crops to shape, normalises when normalizes_signal (over signal_region, :146-155,
        181-189) and masks (:134-158); with no mask, crop or normalisation the fast path filters
        and inverts at shape (:119-133). Without removes_padding the mask and filter are not
        applied (:203-213). Filters and masks are `cryojax_amd.transforms` objects (or any
        object whose `get()` / `array` gives the array they multiply by).
        """
        if rng_key is not None:
            raise NotImplementedError("stochastic volume representations are not implemented")
        cfg = self.image_config
        _check_image_config(cfg, self.volume_parametrization.shape[1])
        if not removes_padding:
            return self._simulate(outputs_real_space)
        if filter is None:
            if mask is None and cfg.padded_shape == cfg.shape and not self.normalizes_signal:
                return self._simulate(outputs_real_space)
            return self._crop_normalize_mask(self._simulate(True), outputs_real_space, mask)
        return self.postprocess(self._simulate(False), outputs_real_space=outputs_real_space, mask=mask,
                                filter=filter)

    def _crop_normalize_mask(self, image, outputs_real_space, mask=None):
        """The slow branch of postprocess after its irfftn (base_image_model.py:151-158): crop to
        shape, normalise when normalizes_signal, mask, and rfftn for a Fourier output."""
        cfg = self.image_config
        out = ops.resize_images(image, cfg.shape, normalize=self.normalizes_signal,
                                region=self.get_signal_region() if self.normalizes_signal else None)
        if mask is not None:
            ops.multiply_broadcast(out, _transform_array(mask))
        return out if outputs_real_space else ops.rfft2(out)

    def compute_fourier_image(self, rng_key=None):
        """rfft half-plane at padded_shape before postprocessing (base_image_model.py:296-328)."""
        return self._simulate(False)

    def postprocess(self, fourier_image, *, outputs_real_space=True, mask=None, filter=None):
        """base_image_model.py:107-158: the filter on the rfft image, then irfftn at padded_shape,
        crop to shape, normalisation when normalizes_signal and the mask; the fast path (filter
        and irfftn at shape) when there is no mask, crop or normalisation."""
        cfg = self.image_config
        fast = mask is None and cfg.padded_shape == cfg.shape and not self.normalizes_signal
        if filter is not None:
            f = _transform_array(filter)
            if not fast and tuple(np.shape(f)) != cfg.padded_rfft_shape:
                raise ValueError(
                    "Found that the `filter` was shape "
                    f"{tuple(np.shape(f))}, but expected it to be "
                    f"shape {cfg.padded_rfft_shape}. You may have passed a "
                    f"fitler according to the "
                    "`AbstractImageModel.image_config.shape`, "
                    "when the `AbstractImageModel.image_config.padded_shape` "
                    "was expected."
                )
            fourier_image = ops.multiply_broadcast(fourier_image.contiguous().clone(), f)
        if fast:
            return ops.irfft2(fourier_image, cfg.shape[0]) if outputs_real_space else fourier_image
        image = ops.irfft2(fourier_image.contiguous(), cfg.padded_shape[0])
        return self._crop_normalize_mask(image, outputs_real_space, mask)

    def get_pose(self):
        return self.pose

    def get_image_config(self):
        return self.image_config

    def get_signal_region(self):
        return self.signal_region


class LinearImageModel(AbstractImageModel):
    """Slice -> CTF -> shift -> irfftn, base_image_model.py:216-328."""

    def __init__(self, volume_parametrization, pose, image_config, volume_integrator, transfer_theory, *,
                 applies_translation=True, normalizes_signal=False, signal_region=None,
                 translate_mode: Literal["fft", "atom"] = "fft"):
        self.volume_parametrization = volume_parametrization
        self.pose = pose
        self.image_config = image_config
        self.volume_integrator = volume_integrator
        self.transfer_theory = transfer_theory
        self.applies_translation = applies_translation
        self.normalizes_signal = normalizes_signal
        self.signal_region = None if signal_region is None else np.asarray(signal_region, dtype=bool)
        self.translate_mode = translate_mode
        if applies_translation and translate_mode == "atom":
            raise ValueError(
                "Tried to apply translation in `translate_mode = 'atom'`, but "
                "found a volume representation that was not an `AbstractAtomVolume`."
                f"Got a `{type(volume_parametrization).__name__}` class."
            )

    def _flags_and_params(self, outputs_real_space):
        """Kernel flags and the packed per-image parameter table of this model."""
        n = self.volume_parametrization.get_representation().shape[1]
        flags = _lib.APPLY_CTF if self.transfer_theory is not None else 0
        if self.applies_translation:
            flags |= _lib.APPLY_TRANSLATION
        if not outputs_real_space:
            flags |= _lib.OUTPUT_FOURIER
        flags |= self.volume_integrator.gather_flags
        packed = _pack(self.pose.rotation, self.pose, self.transfer_theory, self.image_config,
                       self.volume_integrator, n)
        return flags, packed

    def _simulate(self, outputs_real_space):
        vol = self.volume_parametrization.get_representation()
        envelope = None if self.transfer_theory is None else self.transfer_theory.envelope
        if envelope is None:
            flags, packed = self._flags_and_params(outputs_real_space)
            return _run(vol, self.image_config, packed, flags)
        # the envelope multiplies the contrast spectrum (transfer_theory.py:117-118); every factor
        # of compute_fourier_image is per frequency, so it may come after the shift
        cfg = self.image_config
        flags, packed = self._flags_and_params(False)
        spec = _run(vol, cfg, packed, flags).contiguous()
        spec = ops.multiply_broadcast(spec, _envelope_factor(envelope, cfg, tuple(spec.shape)))
        return ops.irfft2(spec, cfg.padded_shape[0]) if outputs_real_space else spec


def _envelope_factor(envelope, cfg, spec_shape):
    """The envelope's values on the padded rfft grid, shaped to multiply `spec_shape`.

    The reference evaluates the envelope once per image, vmapped over a batch of parameters
    (relion.py:1115-1118 builds one FourierGaussian / Constant per particle). Here a batch lives
    in the operator's arrays: an output that ends in the rfft shape is used as it is (one array
    for all images, or one per image), and any other output (a Constant of per-particle scale
    factors) is a value per image, broadcast over the frequencies."""
    rfft = cfg.padded_rfft_shape
    env = np.asarray(envelope(cfg.padded_frequency_grid_in_angstroms), dtype=np.float64)
    if env.shape[-2:] != rfft:
        env = env.reshape(env.shape + (1, 1))
    if env.ndim <= 2:
        return np.ascontiguousarray(np.broadcast_to(env, rfft))
    try:
        return np.ascontiguousarray(np.broadcast_to(env, spec_shape))
    except ValueError:
        raise ValueError(f"an envelope of shape {env.shape} does not match images of shape {spec_shape}") from None


class ProjectionImageModel(LinearImageModel):
    """Slice -> shift -> irfftn without a transfer theory, base_image_model.py:331-420."""

    def __init__(self, volume_parametrization, pose, image_config, volume_integrator, *,
                 applies_translation=True, normalizes_signal=False, signal_region=None,
                 translate_mode: Literal["fft", "atom"] = "fft"):
        super().__init__(volume_parametrization, pose, image_config, volume_integrator, None,
                         applies_translation=applies_translation, normalizes_signal=normalizes_signal,
                         signal_region=signal_region, translate_mode=translate_mode)


# ---------------------------------------------------------------------------
# noise model, cryojax/simulator/_noise_model/gaussian_noise_model.py
# ---------------------------------------------------------------------------


def _key_to_seed(rng_key):
    """An integer seed from an int or any array-like key (JAX keys are absent)."""
    if rng_key is None:
        raise ValueError("an rng_key (an integer seed) is required")
    a = np.asarray(rng_key).astype(np.uint64).ravel()
    seed = np.uint64(0x9E3779B97F4A7C15)
    for v in a:
        seed = np.uint64((int(seed) * 6364136223846793005 + int(v) + 1442695040888963407) % 2**64)
    return int(seed)


class GaussianWhiteNoiseModel:
    """Additive white Gaussian noise on an image model (gaussian_noise_model.py:128-251).

    sample = compute_signal + compute_noise; compute_signal =
    signal_scale_factor * simulate() + signal_offset; compute_noise draws a
    real-valued Gaussian rfft half-plane of std sqrt(N^2 variance) with the
    zero frequency set to 0 and passes it through the image model's
    postprocess (so normalizes_signal normalises the noise too, as in the
    reference). The random stream is Philox keyed by rng_key, not JAX's.
    """

    def __init__(self, image_model, variance=1.0, signal_scale_factor=1.0, signal_offset=0.0):
        if not float(variance) > 0:
            raise ValueError("variance must be positive")
        if not float(signal_scale_factor) > 0:
            raise ValueError("signal_scale_factor must be positive")
        self.image_model = image_model
        self.variance = float(variance)
        self.signal_scale_factor = float(signal_scale_factor)
        self.signal_offset = float(signal_offset)

    def compute_signal(self, *, outputs_real_space=True, mask=None, filter=None):
        if not outputs_real_space:
            raise NotImplementedError("Fourier-space signals of the noise model are not implemented")
        img = self.image_model.simulate(outputs_real_space=True, mask=mask, filter=filter)
        return ops.noisy_images(img, None, self.signal_scale_factor, self.signal_offset)

    def compute_noise(self, rng_key, *, outputs_real_space=True, mask=None, filter=None):
        if not outputs_real_space:
            raise NotImplementedError("Fourier-space noise is not implemented")
        cfg = self.image_model.get_image_config()
        n = cfg.padded_shape[0]
        vol = self.image_model.volume_parametrization.get_representation()
        lead = _batch_shape(self.image_model)
        count = int(np.prod(lead)) if lead else 1
        spec = ops.white_noise_spectrum(count, n, self.variance, _key_to_seed(rng_key),
                                        vol.fourier_voxel_grid.device)
        noise = self.image_model.postprocess(spec, outputs_real_space=True, mask=mask, filter=filter)
        return noise.reshape(lead + tuple(noise.shape[-2:])) if lead else noise[0]

    def sample(self, rng_key, *, outputs_real_space=True, mask=None, filter=None):
        if not outputs_real_space:
            raise NotImplementedError("Fourier-space samples are not implemented")
        img = self.image_model.simulate(outputs_real_space=True, mask=mask, filter=filter)
        noise = self.compute_noise(rng_key, mask=mask, filter=filter)
        return ops.noisy_images(img.contiguous(), noise.contiguous(), self.signal_scale_factor, self.signal_offset)

    def log_likelihood(self, observed, *, mask=None, filter=None):
        """sum over pixels of log N(observed | signal, variance); per image for a batch."""
        sim = self.compute_signal(mask=mask, filter=filter)
        obs = torch.as_tensor(observed, dtype=torch.float32).to(sim.device).contiguous()
        ll = ops.gaussian_log_likelihood(sim.contiguous(), obs.expand_as(sim).contiguous(), self.variance)
        return ll


def _batch_shape(image_model):
    """Leading batch shape of a (vmapped) image model's per-image leaves."""
    flags, (p, b) = image_model._flags_and_params(True)
    return () if b is None else (int(b),)


def make_image_model(volume_parametrization, image_config, pose, transfer_theory=None, volume_integrator=None,
                     detector=None, *, applies_translation=True, normalizes_signal=False, signal_region=None,
                     simulates_quantity=False, quantity_mode="contrast", translate_mode="fft"):
    """cryojax/simulator/_api_utils.py:36-203 for FourierVoxelGridVolume."""
    if volume_integrator is None:
        if isinstance(volume_parametrization, FourierVoxelGridVolume):
            volume_integrator = FourierSliceExtraction(outputs_integral=simulates_quantity)
        else:
            raise ValueError(
                "Could not select default integrator for volume of "
                f"type {type(volume_parametrization).__name__}. If using a custom potential "
                "please directly pass an integrator."
            )
    kw = dict(applies_translation=applies_translation, normalizes_signal=normalizes_signal,
              signal_region=signal_region, translate_mode=translate_mode)
    if transfer_theory is None:
        return ProjectionImageModel(volume_parametrization, pose, image_config, volume_integrator, **kw)
    if simulates_quantity:
        raise NotImplementedError(
            "simulates_quantity=True (WeakPhaseScatteringTheory models) is outside the fused path"
        )
    return LinearImageModel(volume_parametrization, pose, image_config, volume_integrator, transfer_theory, **kw)
