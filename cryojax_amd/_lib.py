"""ctypes binding of `libcryomi.so` (the C ABI declared in include/cryomi.h).

The library is the only compute path: if it is missing, every entry point
raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libcryomi.so"

CRYOMI_OK = 0
CRYOMI_ERR_INVALID_ARGUMENT = 1
CRYOMI_ERR_UNSUPPORTED = 2
CRYOMI_ERR_WORKSPACE = 3
CRYOMI_ERR_LAUNCH = 4

APPLY_CTF = 0x1
APPLY_TRANSLATION = 0x2
OUTPUT_FOURIER = 0x4
GRID_BRICKED = 0x8
# out-of-range corners: every index outside [0, N) reads 0 (default: [-N, -1] wraps, as JAX's
# gather with wrap_negative_indices=True)
GATHER_FILL = 0x10

# cryomi_image_params: 20 float64 fields, in header order.
PARAM_FIELDS = (
    [f"rotation{i}" for i in range(9)]
    + [
        "offset_x",
        "offset_y",
        "defocus",
        "astigmatism",
        "astigmatism_angle",
        "spherical_aberration_mm",
        "wavelength",
        "pixel_size",
        "phase_shift",
        "amplitude",
        "reserved",
    ]
)
N_PARAM = len(PARAM_FIELDS)
assert N_PARAM == 20

_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol include/*.h declares.
SIGNATURES = {
    "cryomi_last_error": (ctypes.c_char_p, []),
    "cryomi_abi_version": (ctypes.c_int, []),
    "cryomi_set_tuning": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "cryomi_get_tuning": (ctypes.c_int, [ctypes.c_char_p]),
    "cryomi_simulate_workspace_bytes": (_c_size, [ctypes.c_int, ctypes.c_int, ctypes.c_uint]),
    "cryomi_simulate": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_uint, _vp, _c_size, _vp, _vp],
    ),
    "cryomi_simulate_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_uint, _vp, _vp]),
    "cryomi_bricked_grid_bytes": (_c_size, [ctypes.c_int]),
    "cryomi_grid_to_bricks": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "cryomi_irfft2_workspace_bytes": (_c_size, [ctypes.c_int, ctypes.c_int]),
    "cryomi_irfft2": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _c_size, _vp, _vp]),
    "cryomi_simulate_from_spectrum": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_uint, _vp, _c_size, _vp, _vp],
    ),
    "cryomi_simulate_band": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_uint, ctypes.c_int, _vp, _vp],
    ),
    "cryomi_simulate_vjp_workspace_bytes": (_c_size, [ctypes.c_int, ctypes.c_int]),
    "cryomi_simulate_vjp": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_uint, _vp, _vp, _c_size, _vp, _vp, _vp],
    ),
    # include/cryomi_search.h
    "cryomi_ncc_search_workspace_bytes": (_c_size, [ctypes.c_int] * 4),
    "cryomi_ncc_search": (ctypes.c_int, [_vp, _vp, _c_size, _vp]),
    "cryomi_sgemm_nt": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "cryomi_ncc_bank_row_floats": (_c_size, [ctypes.c_int]),
    "cryomi_ncc_template_bank": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_uint, _vp, _vp]),
    "cryomi_pose_candidates": (
        ctypes.c_int,
        [_vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double,
         ctypes.c_int, _vp, _vp, _vp, _vp],
    ),
    "cryomi_ncc_search_candidates_workspace_bytes": (_c_size, [ctypes.c_int] * 4),
    "cryomi_ncc_search_candidates": (ctypes.c_int, [_vp, _vp, _vp, _c_size, _vp]),
    # include/cryomi_batch.h
    "cryomi_simulate_groups_workspace_bytes": (_c_size, [_vp, ctypes.c_int]),
    "cryomi_simulate_groups": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _c_size, _vp]),
    "cryomi_simulate_groups_graph_create": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _c_size, _vp]),
    "cryomi_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "cryomi_graph_destroy": (None, [_vp]),
    # include/cryomi_cnn.h
    "cryomi_gemm_bf16": (ctypes.c_int, [_vp, _vp]),
    "cryomi_conv3x3_s2_bf16": (ctypes.c_int, [_vp, _vp]),
    "cryomi_im2col_s2": (ctypes.c_int, [_vp] + [ctypes.c_int] * 5 + [_vp, _vp]),
    "cryomi_col2im_s2": (ctypes.c_int, [_vp, _vp] + [ctypes.c_int] * 5 + [_vp, _vp]),
    "cryomi_column_sums_bf16": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _vp, _vp]),
    "cryomi_to_bf16": (ctypes.c_int, [_vp, _c_size, _vp, _vp]),
    "cryomi_fill_zero": (ctypes.c_int, [_vp, _c_size, _vp]),
    "cryomi_pose_loss": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp],
    ),
    "cryomi_adam": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, _c_size] + [ctypes.c_float] * 6 + [_vp],
    ),
    "cryomi_crop_spectrum": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_uint, _vp, _vp],
    ),
    "cryomi_noise_normalize": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong, _vp, _vp],
    ),
    # include/cryomi_noise.h
    "cryomi_normalize_images": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "cryomi_resize_images": (ctypes.c_int, [_vp] + [ctypes.c_int] * 5 + [ctypes.c_uint, _vp, _vp, _vp]),
    "cryomi_complex128_to_complex64": (ctypes.c_int, [_vp, _c_size, _vp, _vp]),
    "cryomi_white_noise_spectrum": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong,
                                                   _vp, _vp]),
    "cryomi_noisy_images": (ctypes.c_int, [_vp, _vp, _c_size, ctypes.c_float, ctypes.c_float, _vp, _vp]),
    "cryomi_gaussian_log_likelihood": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp,
                                                      _vp]),
    "cryomi_multiply_broadcast": (ctypes.c_int, [_vp, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, _vp,
                                                 ctypes.c_int, _vp]),
    # include/cryomi_ndimage.h
    "cryomi_fourier_crop_operator_floats": (_c_size, [ctypes.c_int] * 4 + [ctypes.c_uint]),
    "cryomi_fourier_crop_operators": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.c_uint, _vp, _vp]),
    "cryomi_fourier_crop_downsample_workspace_bytes": (_c_size, [ctypes.c_int] * 5 + [ctypes.c_uint]),
    "cryomi_fourier_crop_downsample": (
        ctypes.c_int,
        [_vp] + [ctypes.c_int] * 5 + [ctypes.c_uint, _vp, _vp, _c_size, _vp, _vp],
    ),
    "cryomi_bgemm_nt": (ctypes.c_int, [_vp, _vp]),
    "cryomi_spectrum_crop": (ctypes.c_int, [_vp] + [ctypes.c_int] * 6 + [ctypes.c_uint, ctypes.c_double, _vp, _vp]),
    "cryomi_complex_split": (ctypes.c_int, [_vp, ctypes.c_int, _c_size, _vp, _vp, _vp]),
    "cryomi_complex_combine": (ctypes.c_int, [_vp, _vp, _c_size, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp,
                                              _vp]),
    # include/cryomi_fft.h
    "cryomi_fourier_voxel_grid_workspace_bytes": (_c_size, [ctypes.c_int] * 6),
    "cryomi_fourier_voxel_grid": (ctypes.c_int, [_vp] + [ctypes.c_int] * 7 + [_vp, _vp, _vp, _c_size, _vp]),
    "cryomi_rfft2_f64_workspace_bytes": (_c_size, [ctypes.c_int] * 3),
    "cryomi_rfft2_f64": (ctypes.c_int, [_vp] + [ctypes.c_int] * 4 + [_vp, _vp, _c_size, _vp]),
}

REAL_F32 = 0
REAL_F64 = 1

MAX_GROUPS = 8
GEMM_RELU = 0x1
GEMM_ACCUMULATE = 0x2
CONV_FORWARD = 0
CONV_WEIGHT_GRAD = 1
CONV_DATA_GRAD = 2
CROP_PHASE_FLIP = 0x1
CROP_NORMALIZE = 0x2
CROP_HARTLEY = 0x4
CROP_BAND_INPUT = 0x8


RESIZE_NORMALIZE = 0x1

SPEC_HARTLEY = 0x1
SPEC_UNCENTRE = 0x2
SPEC_OUT_F32 = 0x4

DS_PRESERVE_MEAN = 0x1
DS_OUTPUT_FOURIER = 0x2


class BgemmStruct(ctypes.Structure):
    """cryomi_bgemm_args (include/cryomi_ndimage.h)."""

    _fields_ = [
        ("a", ctypes.c_void_p),
        ("a_batch", ctypes.c_longlong),
        ("a_seg", ctypes.c_longlong),
        ("lda", ctypes.c_int),
        ("b", ctypes.c_void_p),
        ("b_batch", ctypes.c_longlong),
        ("b_seg", ctypes.c_longlong),
        ("ldb", ctypes.c_int),
        ("c", ctypes.c_void_p),
        ("c_batch", ctypes.c_longlong),
        ("ldc", ctypes.c_int),
        ("m", ctypes.c_int),
        ("n", ctypes.c_int),
        ("k", ctypes.c_int),
        ("nseg", ctypes.c_int),
        ("batch", ctypes.c_int),
    ]


class VolumeGroupStruct(ctypes.Structure):
    """cryomi_volume_group (include/cryomi_batch.h)."""

    _fields_ = [
        ("fourier_voxel_grid", ctypes.c_void_p),
        ("params", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("box", ctypes.c_int),
        ("n_images", ctypes.c_int),
        ("flags", ctypes.c_uint),
        ("reserved", ctypes.c_int),
    ]


assert ctypes.sizeof(VolumeGroupStruct) == 40


class CryomiError(RuntimeError):
    """A nonzero status from the C ABI, with cryomi_last_error()'s message."""

    def __init__(self, status, message):
        super().__init__(f"[cryomi status {status}] {message}")
        self.status = status


_lock = threading.Lock()
_lib = None


def load():
    """Load libcryomi.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cryojax_amd._build` "
                "(or __graft_entry__.build()); there is no fallback compute path."
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status, what=""):
    if status == CRYOMI_OK:
        return
    msg = load().cryomi_last_error().decode(errors="replace")
    if status in (CRYOMI_ERR_INVALID_ARGUMENT, CRYOMI_ERR_WORKSPACE):
        raise ValueError(f"{what}: {msg}")
    if status == CRYOMI_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise CryomiError(status, f"{what}: {msg}")


def box_supported(n):
    """The image kernels' box sizes: even, 8 <= n <= 512, n = 2^a 3^b 5^c."""
    n = int(n)
    if n < 8 or n > 512 or n % 2:
        return False
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
    return n == 1


def pack_params(
    rotation,
    offset_xy,
    defocus,
    astigmatism,
    astigmatism_angle,
    spherical_aberration_mm,
    wavelength,
    pixel_size,
    phase_shift,
    amplitude,
    n,
):
    """Host float64 (n, 20) table in cryomi_image_params layout."""
    p = np.zeros((n, N_PARAM), dtype=np.float64)
    p[:, 0:9] = np.broadcast_to(np.asarray(rotation, dtype=np.float64).reshape(-1, 9), (n, 9))
    off = np.broadcast_to(np.asarray(offset_xy, dtype=np.float64).reshape(-1, 2), (n, 2))
    p[:, 9:11] = off
    cols = [defocus, astigmatism, astigmatism_angle, spherical_aberration_mm, wavelength,
            pixel_size, phase_shift, amplitude]
    for k, v in enumerate(cols):
        p[:, 11 + k] = np.broadcast_to(np.asarray(v, dtype=np.float64).reshape(-1), (n,))
    return p
