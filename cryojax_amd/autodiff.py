"""Differentiable batched image formation: a torch.autograd.Function over the
fused forward and its hand-written VJP (cryomi_simulate / cryomi_simulate_vjp).

The counterpart of wrapping the FFI call in jax.custom_vjp: the forward runs
the fused HIP kernel, the backward runs the adjoint kernels (the rfft2 of the
cotangent, per-sample g_P with per-image pose / offset / defocus sums, then the
transpose of the trilinear gather: samples binned by brick of voxels and each
brick summed in LDS and added to the grid once, without global float atomics)
and chains dL/dR to the quaternion on the host. Complex gradients follow PyTorch:
grid.grad = dL/dRe + i dL/dIm (JAX reports the conjugate).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, ops
from .rotations import rotation_and_jacobian


def _pack(quats, offsets, defocus, consts):
    r, _ = rotation_and_jacobian(quats)
    n = quats.shape[0]
    return _lib.pack_params(r.reshape(n, 9), offsets, defocus, consts["astigmatism"], consts["astigmatism_angle"],
                            consts["cs_mm"], consts["wavelength"], consts["pixel_size"], consts["phase_shift"],
                            consts.get("amplitude", 1.0), n)


def simulate_vjp(grid, params_dev, flags, cotangent, bricks=None):
    """(grid_grad complex64, param_grads float64 (n, 9)) for a float32 cotangent.

    bricks: optional ops.grid_to_bricks(grid) of the same grid; the gather then reads the bricked
    layout (CRYOMI_GRID_BRICKED). The gradient keeps the grid's standard layout.
    """
    lib = _lib.load()
    n = int(grid.shape[0])
    n_img = int(params_dev.shape[0])
    ops._require_cuda(cotangent, "cotangent", torch.float32)
    gg = torch.zeros_like(grid)
    gp = torch.zeros((n_img, 9), dtype=torch.float64, device=grid.device)
    nbytes = lib.cryomi_simulate_vjp_workspace_bytes(n, max(n_img, 1))
    ws = ops._default_ws.get(grid.device, nbytes)
    src = grid
    if bricks is not None:
        if bricks.numel() * 8 < lib.cryomi_bricked_grid_bytes(n):
            raise ValueError("bricks is smaller than cryomi_bricked_grid_bytes(box)")
        src = bricks
        flags = flags | _lib.GRID_BRICKED
    with torch.cuda.device(grid.device):
        st = lib.cryomi_simulate_vjp(src.data_ptr(), n, n_img, params_dev.data_ptr(), flags,
                                     cotangent.data_ptr(), ws.data_ptr(), ws.numel(), gg.data_ptr(), gp.data_ptr(),
                                     ops._stream_ptr(grid.device))
    _lib.check(st, "cryomi_simulate_vjp")
    return gg, gp


class SimulateImages(torch.autograd.Function):
    """images = simulate(grid, quaternions, offsets, defocus) with a HIP VJP.

    grid: complex64 (N, N, N) cuda tensor; quats (B, 4) wxyz, offsets (B, 2) A,
    defocus (B,) A: float64 tensors; consts: astigmatism, astigmatism_angle,
    cs_mm, wavelength, pixel_size, phase_shift (radians), amplitude, and
    optionally wrap_negative_indices (default True; False = CRYOMI_GATHER_FILL).
    """

    @staticmethod
    def forward(ctx, grid, quats, offsets, defocus, consts):
        q = quats.detach().cpu().numpy()
        o = offsets.detach().cpu().numpy()
        d = defocus.detach().cpu().numpy()
        params = ops.params_to_device(_pack(q, o, d, consts), grid.device)
        flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION
        if not consts.get("wrap_negative_indices", True):
            flags |= _lib.GATHER_FILL
        imgs = ops.simulate_images(grid.detach(), params, flags)
        ctx.save_for_backward(grid, params)
        ctx.q = q
        ctx.dev = (quats.device, offsets.device, defocus.device)
        ctx.flags = flags
        return imgs

    @staticmethod
    def backward(ctx, ct):
        grid, params = ctx.saved_tensors
        gg, gp = simulate_vjp(grid.detach(), params, ctx.flags, ct.contiguous().float())
        gp = gp.cpu().numpy()
        _, jac = rotation_and_jacobian(ctx.q)  # (B, 3, 3, 4)
        dl_dr = np.zeros((gp.shape[0], 3, 3))
        dl_dr[:, :, 0:2] = gp[:, 0:6].reshape(-1, 3, 2)  # only columns 0, 1 of R act on (kx, ky, 0)
        gq = np.einsum("bij,bijq->bq", dl_dr, jac)
        qd, od, dd = ctx.dev
        return (gg, torch.from_numpy(gq).to(qd), torch.from_numpy(gp[:, 6:8].copy()).to(od),
                torch.from_numpy(gp[:, 8].copy()).to(dd), None)


def simulate(grid, quats, offsets, defocus, **consts):
    return SimulateImages.apply(grid, quats, offsets, defocus, consts)
