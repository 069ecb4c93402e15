"""FP64 NumPy restatement of cryoJAX's Fourier-slice + CTF + iFFT image formation.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

Path restated (all citations under `cryojax/` in the reference):
  make_image_model -> LinearImageModel.simulate
    -> FourierVoxelGridVolume.rotate_to_pose      (rotate the central plane)
    -> FourierSliceExtraction.integrate           (trilinear gather, rfft half)
    -> ContrastTransferTheory.propagate_object    (AstigmaticCTF multiply)
    -> AbstractImageModel._phase_shift_translate  (zero self-conjugate bins, shift)
    -> postprocess -> irfftn                      (C2R + fftshift)

Pinning (tests/test_oracle.py):
  * wavelength: closed form value at 300 kV.
  * rotations: ZYZ known-answer matrices from the reference's
    tests/test_agree_with_cistem.py:88-116 (golden json), Euler round trips
    tests/test_pose.py:65-87, identity/axis rotations.
  * trilinear gather: brute-force triple loop on tiny grids; integer
    coordinates return grid values exactly; JAX fill/wrap semantics by
    construction (parity unpinned for out-of-range samples, see below).
  * slice at identity pose == z-projection of the real volume (projection-slice
    theorem; same invariant as tests/test_volume_integrators.py:197-250).
  * 90-degree rotation about z permutes the projection exactly.
  * fast half-plane path == faithful full-plane path (extract all N^2 samples,
    ifftshift, keep [:, :N//2+1]) on random rotations.
  * CTF: closed form at selected frequencies; |CTF| <= 1; CTF(0) = -sin(ac phase).
  * translation: integer-pixel shift equals np.roll away from the zeroed bins.
  * irfftn: numpy.fft FP64 textbook DFT (irfft2 of a true rfft2 round trips).
  * pad_to_shape: the left pad aligns the centres (pad_left = P//2 - s//2)
    for all four size parities; fourier_grid_from_real of a single voxel,
    padded or not, equals the closed-form plane wave
    exp(-2 pi i sum_axes (q - P//2)(m - s//2) / P) (tests/test_oracle_fft.py).
  * filter / mask / CTF envelope in simulate and compute_fourier_image, and the
    volume filter of fourier_grid_from_real (tests/test_oracle_transforms.py):
    unit factors change nothing; a constant envelope scales the image; a filter
    keeping only F(0, 0) leaves F(0, 0) / P^2 at the padded size (so it acts
    before the inverse transform and the crop); the mask acts after the
    normalisation; a volume filter keeping only [0, 0, 0] leaves the volume's
    sum at the centred index. fourier_gaussian: its value at 0 and the product
    rule of Gaussians.

Out-of-range corners: by default (wrap_negative_indices=True) the gather
follows JAX's documented normalise-then-gather rule for
`.at[indices].get(mode="fill")` (indices in [-N, -1] wrap, others fill 0);
wrap_negative_indices=False fills every index outside [0, N) instead. Both
readings are pinned by the brute-force triple loop (which implements each rule
independently) and by closed forms on tiny grids; which one JAX applies cannot
be checked here (no reference output exercises it), so the choice between them
stays parity unpinned, and the tests report the per-image difference between
the two readings next to every GPU error.
"""

from __future__ import annotations

import itertools
import math

import numpy as np

# ----------------------------------------------------------------------------
# constants / grids
# ----------------------------------------------------------------------------


def wavelength_from_kilovolts(voltage_in_kilovolts):
    """Relativistic electron wavelength in Angstroms.

    cryojax/constants/_physical_constants.py:10-29.
    """
    v = 1000.0 * np.asarray(voltage_in_kilovolts, dtype=np.float64)
    return 12.2639 / (v + 0.97845e-6 * v**2) ** 0.5


def _freq_1d(size, grid_spacing, rfft):
    # cryojax/coordinates/_make_coordinate_grids.py:301-317
    fn = np.fft.rfftfreq if rfft else np.fft.fftfreq
    return fn(size, grid_spacing)


def make_frequency_grid(shape, grid_spacing=1.0, outputs_rfftfreqs=True):
    """2D frequency grid with zero in the corner, [..., 0] = x, [..., 1] = y.

    cryojax/coordinates/_make_coordinate_grids.py:67-98, 259-298 (ndim == 2).
    """
    ny, nx = shape
    y = _freq_1d(ny, grid_spacing, False)
    x = _freq_1d(nx, grid_spacing, outputs_rfftfreqs)
    xv, yv = np.meshgrid(x, y, indexing="xy")
    return np.stack([xv, yv], axis=-1)


def make_coordinate_grid_1d(size, grid_spacing=1.0):
    """Real-space coordinates, index size//2 is zero.

    cryojax/coordinates/_make_coordinate_grids.py:308-309.
    """
    return np.fft.fftshift(np.fft.fftfreq(size, 1.0 / grid_spacing)) * size


def make_frequency_slice(shape, outputs_rfftfreqs=False):
    """Centred (1, N, N, 3) frequency slice with a zero z component.

    cryojax/coordinates/_make_coordinate_grids.py:139-198.
    """
    fs = make_frequency_grid(shape, 1.0, outputs_rfftfreqs)
    if outputs_rfftfreqs:
        fs = np.fft.fftshift(fs, axes=(0,))
    else:
        fs = np.fft.fftshift(fs, axes=(0, 1))
    fs = np.pad(fs, ((0, 0), (0, 0), (0, 1)), mode="constant", constant_values=0.0)
    return fs[None]


def safe_multiply_by_constant(grid, constant):
    """cryojax/simulator/_image_config.py:406-412."""
    return np.where(grid != 0.0, constant * grid, 0.0)


# ----------------------------------------------------------------------------
# SO(3) (quaternions, wxyz)
# ----------------------------------------------------------------------------


def quat_compose(q0, q1):
    """Hamilton product q0 (x) q1. cryojax/rotations/_lie_group.py:102-115."""
    w0, x0, y0, z0 = np.moveaxis(np.asarray(q0, dtype=np.float64), -1, 0)
    w1, x1, y1, z1 = np.moveaxis(np.asarray(q1, dtype=np.float64), -1, 0)
    return np.stack(
        [
            -x0 * x1 - y0 * y1 - z0 * z1 + w0 * w1,
            x0 * w1 + y0 * z1 - z0 * y1 + w0 * x1,
            -x0 * z1 + y0 * w1 + z0 * x1 + w0 * y1,
            x0 * y1 - y0 * x1 + z0 * w1 + w0 * z1,
        ],
        axis=-1,
    )


def quat_inverse(q):
    """cryojax/rotations/_lie_group.py:117-119."""
    return np.asarray(q, dtype=np.float64) * np.array([1.0, -1.0, -1.0, -1.0])


def quat_normalize(q):
    """cryojax/rotations/_lie_group.py:315-317."""
    q = np.asarray(q, dtype=np.float64)
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def quat_apply(q, v):
    """Rotate vectors v (..., 3) by the sandwich q (0, v) q^-1.

    cryojax/rotations/_lie_group.py:96-100.
    """
    v = np.asarray(v, dtype=np.float64)
    q = np.broadcast_to(np.asarray(q, dtype=np.float64), v.shape[:-1] + (4,))
    padded = np.concatenate([np.zeros(v.shape[:-1] + (1,)), v], axis=-1)
    return quat_compose(quat_compose(q, padded), quat_inverse(q))[..., 1:]


def so3_epsilon(dtype):
    """cryojax/rotations/_lie_group.py:519-524."""
    return {np.dtype("float32"): 1e-5, np.dtype("float64"): 1e-10}[np.dtype(dtype)]


def so3_exp(tangent, dtype=np.float64):
    """Exponential map with the reference's Taylor branch.

    cryojax/rotations/_lie_group.py:233-271.
    """
    t = np.asarray(tangent, dtype=np.float64)
    theta_sq = np.sum(t * t, axis=-1)
    theta_p4 = theta_sq * theta_sq
    use_taylor = theta_sq < so3_epsilon(dtype)
    safe_theta = np.sqrt(np.where(use_taylor, 1.0, theta_sq))
    half = 0.5 * safe_theta
    real = np.where(use_taylor, 1.0 - theta_sq / 8.0 + theta_p4 / 384.0, np.cos(half))
    imag = np.where(
        use_taylor, 0.5 - theta_sq / 48.0 + theta_p4 / 3840.0, np.sin(half) / safe_theta
    )
    return np.concatenate([real[..., None], imag[..., None] * t], axis=-1)


def _axis_tangent(angle, axis):
    angle = np.asarray(angle, dtype=np.float64)
    t = np.zeros(angle.shape + (3,))
    t[..., axis] = angle
    return t


def euler_to_quat(phi, theta, psi):
    """EulerAnglePose.rotation: R3 @ R2 @ R1 with zyz extrinsic angles in degrees.

    cryojax/simulator/_pose.py:257-272; SO3.from_{y,z}_radians
    cryojax/rotations/_lie_group.py:126-134.
    """
    r1 = so3_exp(_axis_tangent(np.deg2rad(phi), 2))
    r2 = so3_exp(_axis_tangent(np.deg2rad(theta), 1))
    r3 = so3_exp(_axis_tangent(np.deg2rad(psi), 2))
    return quat_compose(quat_compose(r3, r2), r1)


def quat_as_matrix(q):
    """cryojax/rotations/_lie_group.py:220-231."""
    q = np.asarray(q, dtype=np.float64)
    norm = np.sum(q * q, axis=-1, keepdims=True)
    q = q * np.sqrt(2.0 / norm)
    o = q[..., :, None] * q[..., None, :]
    m = np.empty(q.shape[:-1] + (3, 3))
    m[..., 0, 0] = 1.0 - o[..., 2, 2] - o[..., 3, 3]
    m[..., 0, 1] = o[..., 1, 2] - o[..., 3, 0]
    m[..., 0, 2] = o[..., 1, 3] + o[..., 2, 0]
    m[..., 1, 0] = o[..., 1, 2] + o[..., 3, 0]
    m[..., 1, 1] = 1.0 - o[..., 1, 1] - o[..., 3, 3]
    m[..., 1, 2] = o[..., 2, 3] - o[..., 1, 0]
    m[..., 2, 0] = o[..., 1, 3] - o[..., 2, 0]
    m[..., 2, 1] = o[..., 2, 3] + o[..., 1, 0]
    m[..., 2, 2] = 1.0 - o[..., 1, 1] - o[..., 2, 2]
    return m


def sandwich_matrix(q):
    """The linear map v -> quat_apply(q, v), as a 3x3 matrix (columns = images of e_i).

    For a unit quaternion this is the rotation matrix; for any q it is exactly
    the map the reference applies (cryojax/rotations/_lie_group.py:96-100).
    """
    q = np.asarray(q, dtype=np.float64)
    cols = [quat_apply(q, np.broadcast_to(e, q.shape[:-1] + (3,))) for e in np.eye(3)]
    return np.stack(cols, axis=-1)


def quat_to_euler(wxyz, convention="zyz", extrinsic=True):
    """cryojax/rotations/_utils.py:5-54 (returns degrees)."""
    wxyz = np.asarray(wxyz, dtype=np.float64)
    axes_map = {"x": 0, "y": 1, "z": 2}
    axes = [axes_map[a] for a in convention]
    angle_first, angle_third = (0, 2) if extrinsic else (2, 0)
    axes = axes if extrinsic else axes[::-1]
    xyzw = np.roll(wxyz, shift=-1, axis=-1)
    i, j, k = axes
    eps = 1e-7
    k = 3 - i - j
    sign = float((i - j) * (j - k) * (k - i) // 2)
    a, b, c, d = xyzw[..., 3], xyzw[..., i], xyzw[..., j], xyzw[..., k] * sign
    angles = np.empty(wxyz.shape[:-1] + (3,))
    angles[..., 1] = 2 * np.arctan2(np.hypot(c, d), np.hypot(a, b))
    case = np.where(np.abs(angles[..., 1] - np.pi) <= eps, 2, 0)
    case = np.where(np.abs(angles[..., 1]) <= eps, 1, case)
    half_sum = np.arctan2(b, a)
    half_diff = np.arctan2(d, c)
    angles[..., 0] = np.where(case == 1, 2 * half_sum, 2 * half_diff * -1)
    angles[..., angle_first] = np.where(
        case == 0, half_sum - half_diff, angles[..., angle_first]
    )
    angles[..., angle_third] = np.where(
        case == 0, half_sum + half_diff, angles[..., angle_third]
    )
    angles = (angles + np.pi) % (2 * np.pi) - np.pi
    return np.rad2deg(angles)


def so3_sample_uniform(u1, u2, u3):
    """Uniform quaternion from uniforms u1~U(0,1), u2,u3~U(0,2pi).

    cryojax/rotations/_lie_group.py:319-341 (the jax.random draw is replaced by
    explicit uniforms).
    """
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    return np.stack([a * np.sin(u2), a * np.cos(u2), b * np.sin(u3), b * np.cos(u3)], -1)


# ----------------------------------------------------------------------------
# volume
# ----------------------------------------------------------------------------


def left_right_pad(pad_dim, image_dim):
    """_get_left_vs_right_pad (cryojax/ndimage/_edges.py:175-185)."""
    pad = pad_dim - image_dim
    if pad_dim % 2 == 0 and image_dim % 2 == 1:
        return pad // 2 + 1, pad // 2
    if pad_dim % 2 == 1 and image_dim % 2 == 0:
        return pad // 2, pad // 2 + 1
    return pad // 2, pad // 2


def pad_to_shape(x, shape):
    """pad_to_shape with zero padding (cryojax/ndimage/_edges.py:110-147)."""
    return np.pad(x, [left_right_pad(p, s) for s, p in zip(x.shape, shape)], mode="constant")


def fourier_grid_from_real(real_voxel_grid, pad_scale=1.0, dtype=np.complex128, padded_shape=None,
                           fourier_filter=None):
    """FourierVoxelGridVolume.from_real_voxel_grid: fftshift(filter(fftn(ifftshift(pad_to_shape(x))))).

    cryojax/simulator/_volume/representations/voxel_volume.py:68-119,
    cryojax/ndimage/_fft.py:34-54; padded shape int(size * pad_scale) per axis
    unless given. `fourier_filter` (an array of the padded shape) multiplies the
    transform with its zero frequency in the corner (voxel_volume.py:105-109).
    """
    x = np.asarray(real_voxel_grid, dtype=np.float64)
    padded = tuple(int(s * pad_scale) for s in x.shape) if padded_shape is None else tuple(padded_shape)
    if padded != x.shape:
        x = pad_to_shape(x, padded)
    ft = np.fft.fftn(np.fft.ifftshift(x))
    if fourier_filter is not None:
        ft = ft * np.asarray(fourier_filter)
    return np.fft.fftshift(ft).astype(dtype)


# ----------------------------------------------------------------------------
# trilinear gather with JAX `.at[].get(mode="fill")` semantics
# ----------------------------------------------------------------------------


def _normalise_index(idx, n, wrap=True):
    # JAX normalises negative integer indices (i < 0 -> i + n) before a
    # FILL_OR_DROP gather; anything still outside [0, n) reads the fill value.
    # Without wrapping, every index outside [0, n) reads the fill value.
    if wrap:
        idx = np.where(idx < 0, idx + n, idx)
    valid = (idx >= 0) & (idx < n)
    return np.where(valid, idx, 0), valid


def map_coordinates_linear(grid, coords, cval=0.0, wrap_negative_indices=True):
    """Order-1 map_coordinates over a 3D grid, coords = (c_axis0, c_axis1, c_axis2).

    cryojax/ndimage/_map_coordinates.py:85-123 (corner loop, product order of
    itertools.product), :168-175 (floor / weights), :117 (fill gather).
    """
    shape = grid.shape
    per_axis = []
    for c in coords:
        c = np.asarray(c, dtype=np.float64)
        lower = np.floor(c)
        upper_w = c - lower
        lower_w = 1 - upper_w
        idx = lower.astype(np.int64)
        per_axis.append([(idx, lower_w), (idx + 1, upper_w)])
    out = None
    for items in itertools.product(*per_axis):
        valid = True
        flat = []
        w = None
        for axis, (idx, wt) in enumerate(items):
            nidx, v = _normalise_index(idx, shape[axis], wrap_negative_indices)
            flat.append(nidx)
            valid = v if valid is True else (valid & v)
            w = wt if w is None else w * wt
        contrib = np.where(valid, grid[tuple(flat)], cval)
        term = w * contrib
        out = term if out is None else out + term
    return out


def map_coordinates_linear_bruteforce(grid, point, cval=0.0, wrap_negative_indices=True):
    """Single-point trilinear interpolation by explicit triple loop (pinning aid)."""
    n = grid.shape
    base = [math.floor(p) for p in point]
    frac = [p - b for p, b in zip(point, base)]
    acc = 0.0
    for a0 in (0, 1):
        for a1 in (0, 1):
            for a2 in (0, 1):
                w = 1.0
                ids = []
                ok = True
                for ax, a in enumerate((a0, a1, a2)):
                    w *= frac[ax] if a else (1.0 - frac[ax])
                    i = base[ax] + a
                    if i < 0 and wrap_negative_indices:
                        i += n[ax]
                    if not (0 <= i < n[ax]):
                        ok = False
                    ids.append(i)
                acc += w * (grid[tuple(ids)] if ok else cval)
    return acc


# ----------------------------------------------------------------------------
# Fourier helpers
# ----------------------------------------------------------------------------


def enforce_self_conjugate_rfftn_components(rfftn_array, shape, mode="zero"):
    """2D case, includes_zero_frequency=False.

    cryojax/ndimage/_fourier_utils.py:60-132.
    """
    a = np.array(rfftn_array, copy=True)
    y_dim, x_dim = shape
    if mode == "zero":
        fn = lambda v: np.zeros_like(v)  # noqa: E731
    elif mode == "real":
        fn = lambda v: np.real(v) + 0.0j  # noqa: E731
    else:
        raise NotImplementedError(mode)
    if y_dim % 2 == 0:
        a[..., y_dim // 2, 0] = fn(a[..., y_dim // 2, 0])
    if x_dim % 2 == 0:
        a[..., 0, x_dim // 2] = fn(a[..., 0, x_dim // 2])
    if y_dim % 2 == 0 and x_dim % 2 == 0:
        a[..., y_dim // 2, x_dim // 2] = fn(a[..., y_dim // 2, x_dim // 2])
    return a


def convert_fftn_to_rfftn(fftn_array, mode="real"):
    """cryojax/ndimage/_fourier_utils.py:6-57 (2D)."""
    shape = fftn_array.shape[-2:]
    r = fftn_array[..., : shape[-1] // 2 + 1]
    if mode is not None:
        r = enforce_self_conjugate_rfftn_components(r, shape, mode=mode)
    return r


def irfftn(ft, s, workers=None):
    """fftshift(irfft2(ft, s)). cryojax/ndimage/_fft.py:57-77."""
    if workers is None:
        out = np.fft.irfftn(ft, s=s, axes=(-2, -1))
    else:
        import scipy.fft as sfft

        out = sfft.irfftn(ft, s=s, axes=(-2, -1), workers=workers)
    return np.fft.fftshift(out, axes=(-2, -1))


def rfftn(x):
    """rfftn(ifftshift(x)). cryojax/ndimage/_fft.py:80-99."""
    return np.fft.rfftn(np.fft.ifftshift(x, axes=(-2, -1)), axes=(-2, -1))


def crop_to_shape(image, shape):
    """crop_to_shape for images (cryojax/ndimage/_edges.py:15-62), over the last two axes."""
    ny, nx = image.shape[-2:]
    xc, yc = nx // 2, ny // 2
    h, w = shape
    return image[..., yc - h // 2 : yc + h // 2 + h % 2, xc - w // 2 : xc + w // 2 + w % 2]


def pad_image_to_shape(image, shape):
    """pad_to_shape with zero padding for images (_edges.py:110-147), over the last two axes."""
    lead = [(0, 0)] * (image.ndim - 2)
    pads = [left_right_pad(shape[0], image.shape[-2]), left_right_pad(shape[1], image.shape[-1])]
    return np.pad(image, lead + pads, mode="constant")


def resize_with_crop_or_pad(image, shape):
    """resize_with_crop_or_pad with zero padding (cryojax/ndimage/_edges.py:150-172)."""
    n1, n2 = image.shape[-2:]
    m1, m2 = shape
    if n1 >= m1 and n2 >= m2:
        return crop_to_shape(image, shape)
    if n1 <= m1 and n2 <= m2:
        return pad_image_to_shape(image, shape)
    if n1 <= m1 and n2 >= m2:
        return pad_image_to_shape(crop_to_shape(image, (n1, m2)), (m1, m2))
    return pad_image_to_shape(crop_to_shape(image, (m1, n2)), (m1, m2))


# ----------------------------------------------------------------------------
# slice extraction
# ----------------------------------------------------------------------------


def extract_slice_faithful(grid, quat, cval=0.0, wrap_negative_indices=True):
    """Full-plane restatement of FourierSliceExtraction.extract_fourier_slice_from_grid.

    rotate_to_pose (voxel_volume.py:30-37, _pose.py:29-66) -> logical coords
    c = k * N + N // 2 (fourier_voxel_extract.py:527-528) -> map_coordinates
    over (axis0 <- k'_z, axis1 <- k'_y, axis2 <- k'_x) (:530-540) -> ifftshift
    (:542) -> convert_fftn_to_rfftn(mode="real") (:428-443). O(N^2) samples.
    """
    n = grid.shape[0]
    if grid.shape != (n, n, n):
        raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
    fs = make_frequency_slice((n, n), outputs_rfftfreqs=False)
    rotated = quat_apply(quat, fs)
    logical = rotated * n + n // 2
    k0, k1, k2 = logical[..., 0], logical[..., 1], logical[..., 2]
    surface = map_coordinates_linear(grid, (k2, k1, k0), cval, wrap_negative_indices)[0]
    surface = np.fft.ifftshift(surface)
    return convert_fftn_to_rfftn(surface, mode="real")


def half_plane_frequency_indices(n):
    """Signed integer frequency of each kept rfft sample (i, j).

    After ifftshift the corner-layout sample (i, j) sits at
    k = (fftfreq(N)[j], fftfreq(N)[i], 0); only j <= N // 2 survive
    (cryojax/ndimage/_fourier_utils.py:41). For even N, column N/2 therefore
    samples kx = -1/2 (trap 1 in the survey of the path).
    """
    f = np.fft.fftfreq(n) * n
    kx = f[: n // 2 + 1]
    ky = f
    return kx, ky


def extract_slice(grid, quat, cval=0.0, wrap_negative_indices=True):
    """Half-plane-only evaluation of `extract_slice_faithful` (identical samples).

    Only the N x (N//2+1) kept samples are interpolated; the discarded half of
    the reference's full plane never influences the output.
    """
    n = grid.shape[0]
    if grid.shape != (n, n, n):
        raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
    quat = np.asarray(quat, dtype=np.float64)
    kx, ky = half_plane_frequency_indices(n)
    k = np.zeros((n, n // 2 + 1, 3))
    k[..., 0] = (kx / n)[None, :]
    k[..., 1] = (ky / n)[:, None]
    batch = quat.shape[:-1]
    kb = np.broadcast_to(k, batch + k.shape)
    qb = quat.reshape(batch + (1, 1, 4))
    rotated = quat_apply(qb, kb)
    logical = rotated * n + n // 2
    out = map_coordinates_linear(grid, (logical[..., 2], logical[..., 1], logical[..., 0]), cval,
                                 wrap_negative_indices)
    return enforce_self_conjugate_rfftn_components(out, (n, n), mode="real")


# ----------------------------------------------------------------------------
# CTF and translation
# ----------------------------------------------------------------------------


def amplitude_contrast_phase(ac):
    """cryojax/simulator/_transfer_theory/common_functions.py:33-40."""
    ac = np.asarray(ac, dtype=np.float64)
    return np.arctan(ac / np.sqrt(1.0 - ac**2))


def astigmatic_ctf(
    freq_angstrom,
    voltage_kv,
    defocus,
    astigmatism=0.0,
    astigmatism_angle=0.0,
    cs_mm=2.7,
    amplitude_contrast_ratio=0.1,
    phase_shift=0.0,
    defocus_offset=None,
):
    """AbstractCTF.__call__ (outputs_exp=False) for AstigmaticCTF.

    cryojax/simulator/_transfer_theory/transfer_function.py:27-81, 130-169;
    common_functions.py:9-30; coordinates/_geometry.py:5-31 (atan2(qx, qy)).
    Scalar parameters may carry a leading batch shape (broadcast against the grid).
    """
    lam = wavelength_from_kilovolts(voltage_kv)

    def b(x):
        x = np.asarray(x, dtype=np.float64)
        return x.reshape(x.shape + (1, 1))

    df = b(defocus) if defocus_offset is None else b(defocus) + b(defocus_offset)
    qx, qy = freq_angstrom[..., 0], freq_angstrom[..., 1]
    theta = np.arctan2(qx, qy)
    k_sqr = qx * qx + qy * qy
    angle = np.deg2rad(b(astigmatism_angle))
    d = df + 0.5 * b(astigmatism) * np.cos(2.0 * (theta - angle))
    cs_a = b(cs_mm) * 1e7
    lam = b(lam)
    chi = (2 * np.pi) * (-0.5 * d * lam * k_sqr + 0.25 * cs_a * lam**3 * k_sqr**2)
    phase = np.deg2rad(b(phase_shift)) + b(amplitude_contrast_phase(amplitude_contrast_ratio))
    return np.sin(chi - phase)


def translation_operator(freq_angstrom, offset_xy):
    """AbstractPose.compute_translation_operator. cryojax/simulator/_pose.py:97-114."""
    off = np.asarray(offset_xy, dtype=np.float64)
    tx = off[..., 0].reshape(off.shape[:-1] + (1, 1))
    ty = off[..., 1].reshape(off.shape[:-1] + (1, 1))
    dot = freq_angstrom[..., 0] * tx + freq_angstrom[..., 1] * ty
    return np.exp(-1.0j * (2 * np.pi * dot))


# ----------------------------------------------------------------------------
# the image model
# ----------------------------------------------------------------------------


def compute_fourier_image(
    grid,
    quat,
    offset,
    pixel_size,
    voltage_kv=300.0,
    defocus=None,
    astigmatism=0.0,
    astigmatism_angle=0.0,
    cs_mm=2.7,
    amplitude_contrast_ratio=0.1,
    phase_shift=0.0,
    applies_translation=True,
    outputs_integral=False,
    applies_ctf=True,
    faithful=False,
    wrap_negative_indices=True,
    padded_shape=None,
    envelope=None,
):
    """LinearImageModel.compute_fourier_image (image shape == box, no padding).

    cryojax/simulator/_image_model/base_image_model.py:296-328: rotate (:309) ->
    integrate (:314-316; scale by pixel size if outputs_integral,
    fourier_voxel_extract.py:139-141) -> CTF with defocus_offset = offset_z
    (:318-323) -> translate (:325-326, :160-171). With applies_ctf=False this is
    ProjectionImageModel's path (no transfer theory).
    `quat`, `offset` and the CTF scalars may carry one leading batch axis.
    """
    n = grid.shape[0]
    quat = np.asarray(quat, dtype=np.float64)
    if faithful:
        qs = quat.reshape(-1, 4)
        spec = np.stack([extract_slice_faithful(grid, q, wrap_negative_indices=wrap_negative_indices) for q in qs])
        spec = spec.reshape(quat.shape[:-1] + spec.shape[1:])
    else:
        spec = extract_slice(grid, quat, wrap_negative_indices=wrap_negative_indices)
    padded = (n, n) if padded_shape is None else tuple(int(s) for s in padded_shape)
    if padded != (n, n):
        # the integrator resizes the slice to padded_shape in real space
        # (fourier_voxel_extract.py:132-138)
        spec = rfftn(resize_with_crop_or_pad(irfftn(spec, s=(n, n)), padded))
    if outputs_integral:
        spec = spec * np.asarray(pixel_size, dtype=np.float64)
    # CTF and shift on the padded grid (transfer_theory.py:87, base_image_model.py:160-171)
    freq = safe_multiply_by_constant(
        make_frequency_grid(padded, outputs_rfftfreqs=True), 1.0 / pixel_size
    )
    offset = np.asarray(offset, dtype=np.float64)
    if applies_ctf:
        offz = offset[..., 2] if offset.shape[-1] == 3 else None
        ctf = astigmatic_ctf(
            freq,
            voltage_kv,
            defocus,
            astigmatism,
            astigmatism_angle,
            cs_mm,
            amplitude_contrast_ratio,
            phase_shift,
            defocus_offset=offz,
        )
        spec = ctf * spec
        # ContrastTransferTheory's envelope: contrast_spectrum *= envelope(frequency_grid)
        # (transfer_theory.py:117-118), frequencies in inverse Angstroms
        if envelope is not None:
            spec = spec * envelope(freq)
    if applies_translation:
        spec = enforce_self_conjugate_rfftn_components(spec, padded, mode="zero")
        spec = spec * translation_operator(freq, offset[..., :2])
    return spec


def normalize_image(image):
    """AbstractImageModel._normalize_image without a signal region.

    cryojax/simulator/_image_model/base_image_model.py:184-192 (jnp.std, ddof 0).
    """
    mean = image.mean(axis=(-2, -1), keepdims=True)
    std = image.std(axis=(-2, -1), keepdims=True)
    return (image - mean) / std


def fourier_gaussian(frequency_grid, amplitude=1.0, b_factor=1.0):
    """FourierGaussian: amplitude * exp(-b_factor k^2 / 4), k^2 the squared frequency
    (cryojax/ndimage/operators/_fourier_operator.py:226-234)."""
    k_sqr = np.sum(np.asarray(frequency_grid, dtype=np.float64) ** 2, axis=-1)
    return amplitude * np.exp(-0.25 * b_factor * k_sqr)


def simulate(grid, quat, offset, pixel_size, normalizes_signal=False, shape=None, padded_shape=None,
             removes_padding=True, outputs_real_space=True, fourier_filter=None, mask=None, **kwargs):
    """AbstractImageModel.simulate.

    cryojax/simulator/_image_model/base_image_model.py:68-158 -> irfftn
    (cryojax/ndimage/_fft.py:57-77). `shape` and `padded_shape` are those of
    BasicImageConfig (defaults: the volume's box; padded_shape defaults to
    shape, _image_config.py:427-433). The image is formed at padded_shape. The
    postprocess multiplies the rfft image by `fourier_filter` (an array of the
    padded rfft shape: AbstractFilter.__call__, _filters.py:30-33), then crops
    the inverse to shape (crop_to_shape), normalises the crop and multiplies it
    by `mask` (an array of `shape`: AbstractMask.__call__, _masks.py:25-28); with
    no mask, crop or normalisation the fast path filters and inverts at shape
    (:119-133). Without removes_padding neither is applied (:203-213). kwargs:
    those of compute_fourier_image (wrap_negative_indices, envelope included).
    """
    n = grid.shape[0]
    shape = (n, n) if shape is None else tuple(int(s) for s in shape)
    padded = shape if padded_shape is None else tuple(int(s) for s in padded_shape)
    spec = compute_fourier_image(grid, quat, offset, pixel_size, padded_shape=padded, **kwargs)
    if not removes_padding:
        return irfftn(spec, s=padded) if outputs_real_space else spec
    if fourier_filter is not None:
        spec = spec * np.asarray(fourier_filter)
    if mask is None and padded == shape and not normalizes_signal:
        return irfftn(spec, s=shape) if outputs_real_space else spec
    img = irfftn(spec, s=padded)
    if padded != shape:
        img = crop_to_shape(img, shape)
    if normalizes_signal:
        img = normalize_image(img)
    if mask is not None:
        img = img * np.asarray(mask)
    return img if outputs_real_space else rfftn(img)
