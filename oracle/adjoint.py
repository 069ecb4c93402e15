"""FP64 analytic adjoint of the forward image model (reverse-mode restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Loss L = sum over images and pixels of (image - target)^2, image from
oracle.imaging.simulate with a QuaternionPose (wxyz normalised,
cryojax/simulator/_pose.py:322-330), 2D offsets and per-image defocus. The
reference differentiates this path with jax.grad (README.md:99-123): the
transpose of the trilinear gather is a scatter-add, floor has zero derivative
(cryojax/ndimage/_map_coordinates.py:171-175), and the quaternion gradient
flows through normalize (cryojax/rotations/_lie_group.py:316-317).

Complex gradients use g = dL/dRe + i dL/dIm (PyTorch's convention); JAX
reports the conjugate. jax.grad itself cannot run here (JAX is absent), so the
adjoint is pinned by central finite differences of the FP64 forward
(tests/test_oracle_adjoint.py) and is "parity unpinned" against jax.grad.
"""

from __future__ import annotations

import itertools

import numpy as np

from . import imaging as O


def _sandwich_and_jacobian(q):
    """R(q / |q|) (the rotate_coordinates map) and dR/dq, shapes (3, 3), (3, 3, 4)."""
    q = np.asarray(q, dtype=np.float64)
    nq = np.linalg.norm(q)
    u = q / nq
    w, x, y, z = u
    r = np.array(
        [
            [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
        ]
    )
    # dR/du
    du = np.zeros((3, 3, 4))
    du[:, :, 0] = [[2 * w, -2 * z, 2 * y], [2 * z, 2 * w, -2 * x], [-2 * y, 2 * x, 2 * w]]
    du[:, :, 1] = [[2 * x, 2 * y, 2 * z], [2 * y, -2 * x, -2 * w], [2 * z, 2 * w, -2 * x]]
    du[:, :, 2] = [[-2 * y, 2 * x, 2 * w], [2 * x, 2 * y, 2 * z], [-2 * w, 2 * z, -2 * y]]
    du[:, :, 3] = [[-2 * z, -2 * w, 2 * x], [2 * w, -2 * z, 2 * y], [2 * x, 2 * y, 2 * z]]
    # du/dq = (I - u u^T) / |q|
    dudq = (np.eye(4) - np.outer(u, u)) / nq
    return r, du @ dudq


def loss_and_grads(grid, quats, offsets, defocus, targets, pixel_size, voltage_kv=300.0, astigmatism=0.0,
                   astigmatism_angle=0.0, cs_mm=2.7, ac=0.1, wrap_negative_indices=True):
    """Sum-squared loss and its gradients w.r.t. (grid, quats, offsets, defocus).

    grid: complex128 (N, N, N); quats (B, 4) wxyz (normalised inside);
    offsets (B, 2); defocus (B,); targets (B, N, N); wrap_negative_indices: the
    gather's out-of-range corner rule (oracle.imaging.map_coordinates_linear).
    Returns (L, g_grid complex128, g_quats (B, 4), g_offsets (B, 2), g_defocus (B,)).
    """
    n = grid.shape[0]
    m = n // 2
    quats = np.asarray(quats, dtype=np.float64)
    b = quats.shape[0]
    astig = np.broadcast_to(np.asarray(astigmatism, dtype=np.float64), (b,))
    angle = np.broadcast_to(np.asarray(astigmatism_angle, dtype=np.float64), (b,))
    qn = O.quat_normalize(quats)
    imgs = O.simulate(grid, qn, offsets, pixel_size, voltage_kv=voltage_kv, defocus=defocus, astigmatism=astig,
                      astigmatism_angle=angle, cs_mm=cs_mm, amplitude_contrast_ratio=ac,
                      wrap_negative_indices=wrap_negative_indices)
    resid = imgs - targets
    loss = float(np.sum(resid**2))
    dimg = 2.0 * resid
    # image = fftshift(irfft2(F)): dL/dF = (c_kx / N^2) rfft2(ifftshift(dL/dimage)),
    # c = 1 on columns 0 and N/2, 2 elsewhere
    c = np.full(m + 1, 2.0)
    c[0] = 1.0
    c[m] = 1.0
    g_f = np.fft.rfftn(np.fft.ifftshift(dimg, axes=(-2, -1)), axes=(-2, -1)) * (c / (n * n))
    freq = O.safe_multiply_by_constant(O.make_frequency_grid((n, n), outputs_rfftfreqs=True), 1.0 / pixel_size)
    qx, qy = freq[..., 0], freq[..., 1]
    lam = O.wavelength_from_kilovolts(voltage_kv)
    kx, ky = O.half_plane_frequency_indices(n)
    g_grid = np.zeros_like(grid, dtype=np.complex128)
    g_q = np.zeros((b, 4))
    g_off = np.zeros((b, 2))
    g_df = np.zeros(b)
    self_conj = np.zeros((n, m + 1), dtype=bool)
    self_conj[m, 0] = self_conj[0, m] = self_conj[m, m] = True
    for i in range(b):
        r, dr = _sandwich_and_jacobian(quats[i])
        # forward pieces of this image
        k3 = np.zeros((n, m + 1, 3))
        k3[..., 0] = kx[None, :]
        k3[..., 1] = ky[:, None]
        cc = np.einsum("ab,klb->kla", r, k3) + m  # logical coordinates (x, y, z components)
        coords = (cc[..., 2], cc[..., 1], cc[..., 0])  # grid axes (z, y, x)
        per_axis = []
        for cval in coords:
            lo = np.floor(cval)
            f = cval - lo
            per_axis.append((lo.astype(np.int64), f))
        p = np.zeros((n, m + 1), dtype=np.complex128)
        dp = np.zeros((3, n, m + 1), dtype=np.complex128)  # dP / d(coords along z, y, x)
        corners = []
        for bits in itertools.product((0, 1), repeat=3):
            idx = []
            valid = np.ones((n, m + 1), dtype=bool)
            ws = []
            dws = []
            for ax, bit in enumerate(bits):
                lo, f = per_axis[ax]
                ii = lo + bit
                if wrap_negative_indices:
                    ii = np.where(ii < 0, ii + n, ii)
                valid &= (ii >= 0) & (ii < n)
                idx.append(np.where((ii >= 0) & (ii < n), ii, 0))
                ws.append(f if bit else 1.0 - f)
                dws.append(1.0 if bit else -1.0)
            val = np.where(valid, grid[idx[0], idx[1], idx[2]], 0.0)
            w = ws[0] * ws[1] * ws[2]
            p += w * val
            dp[0] += dws[0] * ws[1] * ws[2] * val
            dp[1] += ws[0] * dws[1] * ws[2] * val
            dp[2] += ws[0] * ws[1] * dws[2] * val
            corners.append((idx, valid, w))
        p = np.where(self_conj, 0.0, p)  # real part at extraction, then zeroed by the shift
        ctf = O.astigmatic_ctf(freq, voltage_kv, defocus[i], astig[i], angle[i], cs_mm, ac)
        shift = O.translation_operator(freq, offsets[i])
        gf = g_f[i]
        # F = shift * ctf * P  ->  dL/dP = ctf conj(shift) g_F (zero at the self-conjugate bins)
        gp = np.where(self_conj, 0.0, ctf * np.conj(shift) * gf)
        for idx, valid, w in corners:
            np.add.at(g_grid, (idx[0][valid], idx[1][valid], idx[2][valid]), (w * gp)[valid])
        # coordinates: dL/dc_a = Re(conj(g_P) dP/dc_a); c_(x,y,z) = R k N + m
        dl_dc = np.real(np.conj(gp)[None] * dp)  # axes (z, y, x)
        dl_dcx = np.stack([dl_dc[2], dl_dc[1], dl_dc[0]])  # component order (x, y, z)
        dl_dr = np.einsum("akl,klb->ab", dl_dcx, k3)  # k3 already carries the factor N
        g_q[i] = np.einsum("ab,abq->q", dl_dr, dr)
        fval = shift * ctf * p
        # offsets: F ~ exp(-2 pi i (qx tx + qy ty))
        g_off[i, 0] = np.sum(np.real(np.conj(gf) * (-2j * np.pi * qx) * fval))
        g_off[i, 1] = np.sum(np.real(np.conj(gf) * (-2j * np.pi * qy) * fval))
        # defocus: ctf = sin(chi - phi), d chi / d defocus = -pi lambda k^2
        theta = np.arctan2(qx, qy)
        k2 = qx * qx + qy * qy
        d_eff = defocus[i] + 0.5 * astig[i] * np.cos(2.0 * (theta - np.deg2rad(angle[i])))
        chi = 2 * np.pi * (-0.5 * d_eff * lam * k2 + 0.25 * cs_mm * 1e7 * lam**3 * k2**2)
        dctf = np.cos(chi - O.amplitude_contrast_phase(ac)) * (-np.pi * lam * k2)
        g_df[i] = np.sum(np.real(np.conj(gf) * shift * p) * dctf)
    return loss, g_grid, g_q, g_off, g_df
