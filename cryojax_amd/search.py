"""Pose grid search: run_grid_search / MinimumSearchMethod mirror and the NCC fast path.

`run_grid_search(fn, MinimumSearchMethod(batch_size=...), tree_grid, args)`
keeps the reference's semantics for any host callable (batched loop over
C-order raveled indices, remainder pass, strict `<` with the lowest raveled
index winning ties; cryojax/jax_util/_grid_search/search_loop.py:24-158,
search_method.py:149-310). When `fn` is an `NccObjective` and the grid is
(rotations, offsets), the whole search runs on the GPU in one call to
cryomi_ncc_search (include/cryomi_search.h).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, ops
from .rotations import SO3

# ---------------------------------------------------------------------------
# C ABI binding (include/cryomi_search.h)
# ---------------------------------------------------------------------------


class _NccArgs(ctypes.Structure):
    _fields_ = [
        ("grid_c64", ctypes.c_void_p),
        ("grid_c128", ctypes.c_void_p),
        ("box", ctypes.c_int),
        ("n_images", ctypes.c_int),
        ("observed_rfft", ctypes.c_void_p),
        ("image_params", ctypes.c_void_p),
        ("n_rotations", ctypes.c_int),
        ("rotations", ctypes.c_void_p),
        ("n_offsets", ctypes.c_int),
        ("offsets", ctypes.c_void_p),
        ("tie_window", ctypes.c_double),
        ("best_index", ctypes.c_void_p),
        ("best_score", ctypes.c_void_p),
        ("overflow", ctypes.c_void_p),
        ("bank", ctypes.c_void_p),
        ("flags", ctypes.c_uint),
    ]


def _gather_flags(wrap_negative_indices):
    return 0 if wrap_negative_indices else _lib.GATHER_FILL


def _load():
    return _lib.load()


def sgemm_nt(a, b):
    """C = A . B^T on MFMA (float32 cuda tensors, K % 16 == 0)."""
    lib = _load()
    ops._require_cuda(a, "A", torch.float32)
    ops._require_cuda(b, "B", torch.float32)
    m, k = a.shape
    n, k2 = b.shape
    if k != k2:
        raise ValueError("A and B must share K")
    c = torch.empty((m, n), dtype=torch.float32, device=a.device)
    with torch.cuda.device(a.device):
        st = lib.cryomi_sgemm_nt(a.data_ptr(), b.data_ptr(), c.data_ptr(), m, n, k, ops._stream_ptr(a.device))
    _lib.check(st, "cryomi_sgemm_nt")
    return c


def observed_spectrum(images, device=None):
    """rfftn(ifftshift(y)) of real observed images (..., N, N) as a complex128
    device tensor, cryojax/ndimage/_fft.py:80-99 (FP64 DFT on device,
    cryomi_rfft2_f64). Host arrays are copied to `device` (default: the
    current ROCm device) first.
    """
    x = images if isinstance(images, torch.Tensor) else torch.as_tensor(np.asarray(images))
    if device is None:
        device = x.device if x.device.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
    if x.dtype not in (torch.float32, torch.float64):
        x = x.to(torch.float64)
    return ops.rfft2_f64(x.to(device).contiguous())


def ncc_search(grid_c64, grid_c128, observed_rfft, image_params, rotations, offsets, tie_window=0.0, bank=None,
               wrap_negative_indices=True):
    """Exhaustive NCC search on the GPU (cryomi_ncc_search).

    grid_c64 / grid_c128: (N, N, N) cuda tensors of the centred Fourier grid;
    observed_rfft: (n_img, N, N/2+1) complex128 cuda tensor; image_params:
    (n_img, 20) float64 cuda tensor (CTF fields used); rotations: (n_rot, 9)
    float64; offsets: (n_off, 2) float64 Angstrom (all cuda tensors).
    bank: optional template_bank(grid_c64, rotations) result (skips building
    templates; built with the same wrap_negative_indices). wrap_negative_indices:
    the gather's out-of-range corner rule (False = CRYOMI_GATHER_FILL).
    Returns (best raveled index int64, best -NCC float64, overflow int32).
    """
    lib = _load()
    n = int(grid_c64.shape[0])
    n_img = int(observed_rfft.shape[0])
    n_rot = int(rotations.shape[0])
    n_off = int(offsets.shape[0])
    dev = grid_c64.device
    for t, name, dt in ((grid_c64, "grid_c64", torch.complex64), (grid_c128, "grid_c128", torch.complex128),
                        (observed_rfft, "observed_rfft", torch.complex128), (image_params, "image_params",
                                                                             torch.float64),
                        (rotations, "rotations", torch.float64), (offsets, "offsets", torch.float64)):
        ops._require_cuda(t, name, dt)
    best_i = torch.empty(n_img, dtype=torch.int64, device=dev)
    best_s = torch.empty(n_img, dtype=torch.float64, device=dev)
    over = torch.empty(n_img, dtype=torch.int32, device=dev)
    nbytes = lib.cryomi_ncc_search_workspace_bytes(n, max(n_img, 1), max(n_rot, 1), max(n_off, 1))
    ws = ops._default_ws.get(dev, nbytes)
    args = _NccArgs(grid_c64.data_ptr(), grid_c128.data_ptr(), n, n_img, observed_rfft.data_ptr(),
                    image_params.data_ptr(), n_rot, rotations.data_ptr(), n_off, offsets.data_ptr(),
                    float(tie_window), best_i.data_ptr(), best_s.data_ptr(), over.data_ptr(),
                    None if bank is None else bank.data_ptr(), _gather_flags(wrap_negative_indices))
    with torch.cuda.device(dev):
        st = lib.cryomi_ncc_search(ctypes.addressof(args), ws.data_ptr(), ws.numel(), ops._stream_ptr(dev))
    _lib.check(st, "cryomi_ncc_search")
    return best_i, best_s, over


class _CandArgs(ctypes.Structure):
    _fields_ = [
        ("bank", ctypes.c_void_p),
        ("cand", ctypes.c_void_p),
        ("cand_count", ctypes.c_void_p),
        ("cap", ctypes.c_int),
        ("flags", ctypes.c_uint),
        ("rescored", ctypes.c_void_p),
    ]


REUSE_PREPARED = 0x1
MERGE = 0x2


def template_bank(grid_c64, rotations, wrap_negative_indices=True):
    """FP32 slice templates of every rotation (cryomi_ncc_template_bank): (n_rot, row) float32."""
    lib = _load()
    ops._require_cuda(grid_c64, "grid_c64", torch.complex64)
    ops._require_cuda(rotations, "rotations", torch.float64)
    n = int(grid_c64.shape[0])
    row = lib.cryomi_ncc_bank_row_floats(n)
    bank = torch.empty((int(rotations.shape[0]), row), dtype=torch.float32, device=grid_c64.device)
    with torch.cuda.device(grid_c64.device):
        st = lib.cryomi_ncc_template_bank(grid_c64.data_ptr(), n, rotations.data_ptr(), int(rotations.shape[0]),
                                          _gather_flags(wrap_negative_indices), bank.data_ptr(),
                                          ops._stream_ptr(grid_c64.device))
    _lib.check(st, "cryomi_ncc_template_bank")
    return bank


def pose_candidates(rotations, max_angle_deg, cap, *, hypotheses=None, hyp_stride=None, n_hyp=None, center_index=None,
                    n_offsets=1, n_images=None, out=None):
    """Grid rotations near each image's hypotheses (cryomi_pose_candidates).

    Either hypotheses (float32, image i's h-th 3x3 at [i * hyp_stride + 9 h])
    or center_index (int64 raveled search results). Returns (cand (n, cap)
    int32, count (n,) int32, overflow (n,) int32).
    """
    lib = _load()
    ops._require_cuda(rotations, "rotations", torch.float64)
    dev = rotations.device
    if center_index is not None:
        n = int(center_index.shape[0])
    else:
        n = int(n_images if n_images is not None else hypotheses.shape[0])
    if out is None:
        out = (torch.empty((n, cap), dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
               torch.empty(n, dtype=torch.int32, device=dev))
    cand, count, over = out
    with torch.cuda.device(dev):
        st = lib.cryomi_pose_candidates(
            None if hypotheses is None else hypotheses.data_ptr(), int(hyp_stride or 0), int(n_hyp or 0),
            None if center_index is None else center_index.data_ptr(), int(n_offsets), n, rotations.data_ptr(),
            int(rotations.shape[0]), float(max_angle_deg), int(cap), cand.data_ptr(), count.data_ptr(),
            over.data_ptr(), ops._stream_ptr(dev))
    _lib.check(st, "cryomi_pose_candidates")
    return cand, count, over


def reserve_candidates_workspace(grid_c128, n_images, n_offsets, cap):
    """Grows the shared workspace to what a candidate search with up to `cap`
    candidates needs, so that later REUSE_PREPARED calls find it in place."""
    nbytes = _load().cryomi_ncc_search_candidates_workspace_bytes(int(grid_c128.shape[0]), max(int(n_images), 1),
                                                                   max(int(n_offsets), 1), int(cap))
    if nbytes:
        ops._default_ws.get(grid_c128.device, nbytes)


def ncc_search_candidates(grid_c128, observed_rfft, image_params, rotations, offsets, bank, cand, cand_count, *,
                          best=None, flags=0, tie_window=0.0, workspace=None, rescored=None,
                          wrap_negative_indices=True):
    """NCC search restricted to per-image candidate rotations (cryomi_ncc_search_candidates).

    best: optional (best_index int64, best_score float64, overflow int32)
    tensors, updated in place (required with MERGE). Returns them.
    """
    lib = _load()
    n = int(grid_c128.shape[0])
    n_img = int(observed_rfft.shape[0])
    n_off = int(offsets.shape[0])
    cap = int(cand.shape[1])
    dev = grid_c128.device
    if best is None:
        best = (torch.empty(n_img, dtype=torch.int64, device=dev), torch.empty(n_img, dtype=torch.float64, device=dev),
                torch.empty(n_img, dtype=torch.int32, device=dev))
    bi, bs, over = best
    nbytes = lib.cryomi_ncc_search_candidates_workspace_bytes(n, max(n_img, 1), max(n_off, 1), cap)
    if nbytes == 0:
        raise ValueError("unsupported candidate search shape (box, n_offsets <= 32, 1 <= cap <= 8192)")
    if workspace is None and flags & REUSE_PREPARED and not ops._default_ws.holds(dev, nbytes):
        raise ValueError("REUSE_PREPARED: the shared workspace would be reallocated and lose the prepared data; "
                         "run the preparing call with the larger cap first or pass workspace=")
    ws = workspace if workspace is not None else ops._default_ws.get(dev, nbytes)
    args = _NccArgs(None, grid_c128.data_ptr(), n, n_img, observed_rfft.data_ptr(), image_params.data_ptr(),
                    int(rotations.shape[0]), rotations.data_ptr(), n_off, offsets.data_ptr(), float(tie_window),
                    bi.data_ptr(), bs.data_ptr(), over.data_ptr(), None, _gather_flags(wrap_negative_indices))
    cargs = _CandArgs(bank.data_ptr(), cand.data_ptr(), cand_count.data_ptr(), cap, int(flags),
                      None if rescored is None else rescored.data_ptr())
    with torch.cuda.device(dev):
        st = lib.cryomi_ncc_search_candidates(ctypes.addressof(args), ctypes.addressof(cargs), ws.data_ptr(),
                                              ws.numel(), ops._stream_ptr(dev))
    _lib.check(st, "cryomi_ncc_search_candidates")
    return best


# ---------------------------------------------------------------------------
# run_grid_search / MinimumSearchMethod mirror
# ---------------------------------------------------------------------------


class MinimumState:
    def __init__(self, current_minimum_eval, current_best_raveled_index, current_eval=None):
        self.current_minimum_eval = current_minimum_eval
        self.current_best_raveled_index = current_best_raveled_index
        self.current_eval = current_eval


class MinimumSolution:
    def __init__(self, value, stats, state):
        self.value = value
        self.stats = stats
        self.state = state


class MinimumSearchMethod:
    """search_method.py:149-310 (elementwise minimum of fn over the grid)."""

    def __init__(self, *, stores_solution_value=True, stores_current_eval=False, batch_size=None):
        self.stores_solution_value = stores_solution_value
        self.stores_current_eval = stores_current_eval
        self.batch_size = batch_size


def tree_grid_shape(tree_grid):
    leaves = list(tree_grid)
    if not leaves:
        raise ValueError("The pytree passed to `tree_grid_shape` should have at least one leaf.")
    return tuple(int(np.shape(leaf)[0]) for leaf in leaves)


def tree_grid_unravel_index(raveled, tree_grid):
    """C-order unravel with jnp.unravel_index's clipping (pytree_manipulation.py:78-118)."""
    shape = tree_grid_shape(tree_grid)
    total = math.prod(shape)
    r = np.clip(np.asarray(raveled, dtype=np.int64), -total, total - 1)
    r = np.where(r < 0, r + total, r)
    return np.unravel_index(r, shape)


def tree_grid_take(tree_grid, index):
    return tuple(np.asarray(leaf)[i] for leaf, i in zip(tree_grid, index))


class NccObjective:
    """fn(grid_point=(rotation, offset), args) = -NCC of each observed image
    against its CTF'd template at that pose; the GPU path evaluates the whole
    grid at once.

    volume: FourierVoxelGridVolume; observed_images: (n_img, N, N) real;
    image_config: BasicImageConfig; transfer_theory: ContrastTransferTheory with
    per-image CTF leaves (arrays of length n_img); wrap_negative_indices: the
    templates' out-of-range corner rule (FourierSliceExtraction's).
    """

    def __init__(self, volume, observed_images, image_config, transfer_theory, wrap_negative_indices=True):
        self.volume = volume
        self.wrap_negative_indices = bool(wrap_negative_indices)
        self.image_config = image_config
        self.transfer_theory = transfer_theory
        self.observed = observed_spectrum(observed_images, device=volume.fourier_voxel_grid.device)
        n = volume.shape[0]
        self.grid_c128 = volume.fourier_voxel_grid_f64
        self.n = n

    def image_params(self):
        from .simulator import wavelength_from_kilovolts

        ctf = self.transfer_theory.ctf
        n_img = self.observed.shape[0]
        return _lib.pack_params(np.eye(3).reshape(1, 9), np.zeros((1, 2)), ctf.defocus_in_angstroms,
                                ctf.astigmatism_in_angstroms, ctf.astigmatism_angle,
                                ctf.spherical_aberration_in_mm,
                                wavelength_from_kilovolts(self.image_config.voltage_in_kilovolts),
                                self.image_config.pixel_size, self.transfer_theory.total_phase_shift(), 1.0, n_img)


def _rotation_matrices(leaf):
    a = np.asarray(leaf, dtype=np.float64)
    if a.ndim == 2 and a.shape[1] == 4:
        return SO3(a).normalize().sandwich_matrix().reshape(-1, 9)
    if a.ndim == 3 and a.shape[1:] == (3, 3):
        return a.reshape(-1, 9)
    raise ValueError("rotation leaf must be (n, 4) quaternions wxyz or (n, 3, 3) matrices")


def run_grid_search(fn, method, tree_grid, args=None, *, is_leaf=None, progress_bar=False,
                    total_progress_bar_updates=20):
    """search_loop.py:24-158 with MinimumSearchMethod."""
    if not isinstance(method, MinimumSearchMethod):
        raise NotImplementedError("only MinimumSearchMethod is implemented")
    shape = tree_grid_shape(tree_grid)
    grid_size = math.prod(shape)
    if isinstance(fn, NccObjective):
        rots, offs = tree_grid
        dev = fn.volume.fourier_voxel_grid.device
        bi, bs, _ = ncc_search(
            fn.volume.fourier_voxel_grid,
            fn.grid_c128.to(dev),
            fn.observed.to(dev),
            ops.params_to_device(fn.image_params(), dev),
            torch.from_numpy(_rotation_matrices(rots)).to(dev),
            torch.from_numpy(np.asarray(offs, dtype=np.float64).reshape(-1, 2)).to(dev),
            wrap_negative_indices=fn.wrap_negative_indices,
        )
        best_idx = bi.cpu().numpy()
        best = bs.cpu().numpy()
    else:
        best, best_idx = _host_search(fn, tree_grid, args, method.batch_size, grid_size)
    state = MinimumState(best, best_idx)
    value = None
    if method.stores_solution_value:
        idx = tree_grid_unravel_index(best_idx.ravel() if np.ndim(best_idx) else best_idx, tree_grid)
        pts = tree_grid_take(tree_grid, idx)
        shp = np.shape(best_idx)
        value = tuple(p.reshape(shp + p.shape[1:]) if p.ndim > 1 else p.reshape(shp) for p in pts)
    return MinimumSolution(value, {"grid_size": grid_size}, state)


def _host_search(fn, tree_grid, args, batch_size, grid_size):
    p0 = tree_grid_take(tree_grid, tree_grid_unravel_index(0, tree_grid))
    out0 = np.asarray(fn(tuple(x[None] for x in p0), args) if batch_size else fn(p0, args), dtype=np.float64)
    f_shape = out0.shape[1:] if batch_size else out0.shape
    best = np.full(f_shape, np.inf)
    best_idx = np.zeros(f_shape, dtype=np.int64)
    if batch_size is None:
        for it in range(grid_size):
            v = np.asarray(fn(tree_grid_take(tree_grid, tree_grid_unravel_index(it, tree_grid)), args))
            less = v < best
            best, best_idx = np.where(less, v, best), np.where(less, it, best_idx)
        return best, best_idx
    n_iter = grid_size // batch_size
    iters = list(range(n_iter))
    if grid_size % batch_size:
        iters += list(range(n_iter, n_iter + grid_size % batch_size))
    for it in iters:
        idx = np.linspace(it * batch_size, (it + 1) * batch_size - 1, batch_size).astype(np.int32)
        vals = np.asarray(fn(tree_grid_take(tree_grid, tree_grid_unravel_index(idx, tree_grid)), args))
        am = np.argmin(vals, axis=0)
        v = np.min(vals, axis=0)
        ridx = np.take(idx, am).astype(np.int64)
        less = v < best
        best, best_idx = np.where(less, v, best), np.where(less, ridx, best_idx)
    return best, best_idx


def super_fibonacci_quaternions(n):
    """Deterministic near-uniform rotations (Alexa, "Super-Fibonacci Spirals", CVPR 2022)."""
    psi = 1.533751168755204288118041  # psi^4 = psi + 4
    phi = math.sqrt(2.0)
    s = np.arange(n, dtype=np.float64) + 0.5
    r, rr = np.sqrt(s / n), np.sqrt(1.0 - s / n)
    a, b = 2 * np.pi * s / phi, 2 * np.pi * s / psi
    return np.stack([r * np.sin(a), r * np.cos(a), rr * np.sin(b), rr * np.cos(b)], -1)


def offset_grid(n_side, step):
    """n_side x n_side in-plane offsets in Angstrom, C order over (x, y)."""
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * step
    ox, oy = np.meshgrid(c, c, indexing="ij")
    return np.stack([ox.ravel(), oy.ravel()], -1)
