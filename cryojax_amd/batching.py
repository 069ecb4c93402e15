"""Uneven batches in one persistent launch, and their HIP-graph replay.

The reference forms a batch by jax.vmap over the poses of one image model
(cryojax/simulator/_image_model/base_image_model.py:68-105), one compiled
program per (volume, box). Here several such vmapped models - each its own
FourierVoxelGridVolume, box and pose/CTF table - are formed by one kernel
launch (cryomi_simulate_groups, include/cryomi_batch.h): workgroups take
images from a shared queue, largest box first. A GroupGraph captures that
launch once and replays it; between replays the caller rewrites the
parameter tables in place (decode-style stepping: same shapes, new poses).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib, ops


@dataclass
class VolumeGroup:
    """One vmapped image model on device buffers.

    grid   : complex64 grid, either (box, box, box) standard layout or the flat
             output of ops.grid_to_bricks (then bricked=True)
    params : float64 (n, 20) device tensor of cryomi_image_params
    flags  : CRYOMI_APPLY_CTF / CRYOMI_APPLY_TRANSLATION / CRYOMI_GATHER_FILL bits
    out    : float32 (n, box, box) device tensor (allocated when None)
    """

    grid: torch.Tensor
    box: int
    params: torch.Tensor
    flags: int = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION
    bricked: bool = True
    out: torch.Tensor | None = None

    def __post_init__(self):
        ops._require_cuda(self.grid, "grid", torch.complex64)
        ops._require_cuda(self.params, "params", torch.float64)
        n = int(self.params.shape[0])
        if self.params.ndim != 2 or self.params.shape[1] != _lib.N_PARAM:
            raise ValueError(f"params must have shape (n, {_lib.N_PARAM})")
        if self.bricked:
            need = _lib.load().cryomi_bricked_grid_bytes(self.box)
            if need == 0 or self.grid.numel() * 8 < need:
                raise ValueError("bricked grid is smaller than cryomi_bricked_grid_bytes(box)")
        elif tuple(self.grid.shape) != (self.box,) * 3:
            raise AttributeError("Only cubic boxes are supported for fourier slice extraction.")
        if self.out is None:
            self.out = torch.empty((n, self.box, self.box), dtype=torch.float32, device=self.grid.device)
        else:
            ops._require_cuda(self.out, "out", torch.float32)
            if tuple(self.out.shape) != (n, self.box, self.box):
                raise ValueError("out must have shape (n, box, box)")

    def _struct(self):
        flags = int(self.flags) | (_lib.GRID_BRICKED if self.bricked else 0)
        return _lib.VolumeGroupStruct(self.grid.data_ptr(), self.params.data_ptr(), self.out.data_ptr(), self.box,
                                      int(self.params.shape[0]), flags, 0)


def _table(groups):
    if not 1 <= len(groups) <= _lib.MAX_GROUPS:
        raise ValueError(f"between 1 and {_lib.MAX_GROUPS} groups per launch")
    dev = groups[0].grid.device
    if any(g.grid.device != dev for g in groups):
        raise ValueError("all groups must live on one device")
    arr = (_lib.VolumeGroupStruct * len(groups))(*[g._struct() for g in groups])
    return arr, dev


def _workspace(lib, arr, n, dev, workspace):
    nbytes = lib.cryomi_simulate_groups_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), n)
    if nbytes == 0:
        # the description is invalid: let the call report why
        return None
    if workspace is None:
        return torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
    if workspace.numel() < nbytes:
        raise ValueError("workspace smaller than cryomi_simulate_groups_workspace_bytes()")
    return workspace


def simulate_groups(groups, workspace=None):
    """Form every group's images in one launch; returns the list of outputs."""
    lib = _lib.load()
    arr, dev = _table(groups)
    ws = _workspace(lib, arr, len(groups), dev, workspace)
    with torch.cuda.device(dev):
        st = lib.cryomi_simulate_groups(ctypes.cast(arr, ctypes.c_void_p), len(groups),
                                        None if ws is None else ws.data_ptr(), 0 if ws is None else ws.numel(),
                                        ops._stream_ptr(dev))
    _lib.check(st, "cryomi_simulate_groups")
    return [g.out for g in groups]


class GroupGraph:
    """cryomi_simulate_groups captured once into a HIP graph.

    The graph holds the groups' device pointers: rewrite `groups[i].params`
    in place (for example with `params.copy_(new)`) and call `replay()` to form
    the new images into the same output tensors.
    """

    def __init__(self, groups, workspace=None):
        self._lib = _lib.load()
        self.groups = list(groups)
        arr, self.device = _table(self.groups)
        self.workspace = _workspace(self._lib, arr, len(self.groups), self.device, workspace)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            st = self._lib.cryomi_simulate_groups_graph_create(
                ctypes.cast(arr, ctypes.c_void_p), len(self.groups),
                None if self.workspace is None else self.workspace.data_ptr(),
                0 if self.workspace is None else self.workspace.numel(), ctypes.byref(handle))
        _lib.check(st, "cryomi_simulate_groups_graph_create")
        self._handle = handle

    def replay(self):
        if self._handle is None:
            raise RuntimeError("GroupGraph was closed")
        with torch.cuda.device(self.device):
            st = self._lib.cryomi_graph_launch(self._handle, ops._stream_ptr(self.device))
        _lib.check(st, "cryomi_graph_launch")
        return [g.out for g in self.groups]

    def close(self):
        if self._handle is not None:
            self._lib.cryomi_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_image_models(image_models, workspace=None):
    """simulate() of several (vmapped) image models in one launch.

    Each model is a LinearImageModel / ProjectionImageModel over a
    FourierVoxelGridVolume with a batched pose (the shape a jax.vmap over
    make_image_model(...).simulate() would see). Returns one (n, box, box)
    tensor per model, equal to model.simulate().
    """
    from . import simulator as S

    groups = []
    for m in image_models:
        if not isinstance(m, S.LinearImageModel):
            raise TypeError("simulate_image_models takes LinearImageModel / ProjectionImageModel instances")
        if m.normalizes_signal:
            raise NotImplementedError("normalizes_signal=True is not implemented by the fused kernel")
        vol = m.volume_parametrization.get_representation()
        n = vol.shape[1]
        S._check_image_config(m.image_config, n)
        flags, packed = m._flags_and_params(True)
        p, _ = packed
        dev = vol.fourier_voxel_grid.device
        groups.append(VolumeGroup(vol.bricked_grid(), n, ops.params_to_device(p, dev), flags, True))
    outs = simulate_groups(groups, workspace)
    return outs
