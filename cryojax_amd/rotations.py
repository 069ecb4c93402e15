"""Host-side SO(3) in float64 (quaternions wxyz), mirroring cryojax.rotations.SO3.

Pose algebra is tiny per-image host work; the kernels receive the resulting
rotation matrices. Leading batch axes are supported throughout.
"""

from __future__ import annotations

import numpy as np


def _eps(dtype):
    # cryojax/rotations/_lie_group.py:519-524
    return {np.dtype("float32"): 1e-5, np.dtype("float64"): 1e-10}[np.dtype(dtype)]


class SO3:
    """Rotation as a (possibly batched) quaternion, cryojax/rotations/_lie_group.py:71-341."""

    def __init__(self, wxyz):
        self.wxyz = np.asarray(wxyz, dtype=np.float64)
        if self.wxyz.shape[-1] != 4:
            raise ValueError("SO3 expects wxyz with a trailing dimension of 4")

    def __repr__(self):
        return f"SO3(wxyz={self.wxyz!r})"

    def compose(self, other):
        """Hamilton product, _lie_group.py:102-115."""
        w0, x0, y0, z0 = np.moveaxis(self.wxyz, -1, 0)
        w1, x1, y1, z1 = np.moveaxis(other.wxyz, -1, 0)
        return SO3(
            np.stack(
                [
                    -x0 * x1 - y0 * y1 - z0 * z1 + w0 * w1,
                    x0 * w1 + y0 * z1 - z0 * y1 + w0 * x1,
                    -x0 * z1 + y0 * w1 + z0 * x1 + w0 * y1,
                    x0 * y1 - y0 * x1 + z0 * w1 + w0 * z1,
                ],
                axis=-1,
            )
        )

    def __matmul__(self, other):
        if isinstance(other, SO3):
            return self.compose(other)
        return self.apply(other)

    def inverse(self):
        return SO3(self.wxyz * np.array([1.0, -1.0, -1.0, -1.0]))

    def apply(self, target):
        """q (0, v) q^-1, _lie_group.py:96-100."""
        v = np.asarray(target, dtype=np.float64)
        q = SO3(np.broadcast_to(self.wxyz, np.broadcast_shapes(self.wxyz.shape[:-1], v.shape[:-1]) + (4,)))
        p = SO3(np.concatenate([np.zeros(v.shape[:-1] + (1,)), np.broadcast_to(v, q.wxyz.shape[:-1] + (3,))], -1))
        return (q @ p @ q.inverse()).wxyz[..., 1:]

    def normalize(self):
        return SO3(self.wxyz / np.linalg.norm(self.wxyz, axis=-1, keepdims=True))

    def sandwich_matrix(self):
        """Matrix of v -> apply(v) (the rotation matrix for a unit quaternion)."""
        cols = [self.apply(np.broadcast_to(e, self.wxyz.shape[:-1] + (3,))) for e in np.eye(3)]
        return np.stack(cols, axis=-1)

    def as_matrix(self):
        """_lie_group.py:220-231."""
        q = self.wxyz
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

    @classmethod
    def exp(cls, tangent):
        """_lie_group.py:233-271 (Taylor branch below the float64 epsilon)."""
        t = np.asarray(tangent, dtype=np.float64)
        th2 = np.sum(t * t, axis=-1)
        th4 = th2 * th2
        taylor = th2 < _eps(np.float64)
        safe = np.sqrt(np.where(taylor, 1.0, th2))
        real = np.where(taylor, 1.0 - th2 / 8.0 + th4 / 384.0, np.cos(0.5 * safe))
        imag = np.where(taylor, 0.5 - th2 / 48.0 + th4 / 3840.0, np.sin(0.5 * safe) / safe)
        return cls(np.concatenate([real[..., None], imag[..., None] * t], axis=-1))

    @classmethod
    def _axis(cls, angle, axis):
        a = np.asarray(angle, dtype=np.float64)
        t = np.zeros(a.shape + (3,))
        t[..., axis] = a
        return cls.exp(t)

    @classmethod
    def from_x_radians(cls, angle):
        return cls._axis(angle, 0)

    @classmethod
    def from_y_radians(cls, angle):
        return cls._axis(angle, 1)

    @classmethod
    def from_z_radians(cls, angle):
        return cls._axis(angle, 2)

    @classmethod
    def identity(cls):
        return cls(np.array([1.0, 0.0, 0.0, 0.0]))

    @classmethod
    def sample_uniform(cls, rng, shape=()):
        """Uniform on SO(3) from a numpy Generator (map of _lie_group.py:319-341)."""
        u1 = rng.uniform(0.0, 1.0, shape)
        u2 = rng.uniform(0.0, 2 * np.pi, shape)
        u3 = rng.uniform(0.0, 2 * np.pi, shape)
        a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
        return cls(np.stack([a * np.sin(u2), a * np.cos(u2), b * np.sin(u3), b * np.cos(u3)], -1))


def convert_quaternion_to_euler_angles(wxyz, convention="zyz", extrinsic=True):
    """cryojax/rotations/_utils.py:5-54 (degrees).

    The method is the direct quaternion-to-Euler conversion of Bernardes and Viollet,
    "Quaternion to Euler angles conversion: a direct, general and computationally efficient
    method", PLoS ONE 17(11): e0276302 (2022), as the reference adapted it (through
    jax-scipy-spatial): the case analysis on the half-sum and half-difference angles and the
    1e-7 degeneracy threshold are that method's, kept so that angles match bit for bit.
    """
    if len(convention) != 3 or not all(a in "xyz" for a in convention):
        raise ValueError(f"`convention` should be three of 'x', 'y', 'z'. Got '{convention}'")
    if convention[0] == convention[1] or convention[1] == convention[2]:
        raise ValueError(f"`convention` cannot repeat axes in a row. Got '{convention}'.")
    if convention[0] != convention[2]:
        raise ValueError(f"`convention` must repeat its first and last axis. Got '{convention}'.")
    wxyz = np.asarray(wxyz, dtype=np.float64)
    amap = {"x": 0, "y": 1, "z": 2}
    axes = [amap[a] for a in convention]
    first, third = (0, 2) if extrinsic else (2, 0)
    axes = axes if extrinsic else axes[::-1]
    xyzw = np.roll(wxyz, -1, axis=-1)
    i, j, _ = axes
    k = 3 - i - j
    sign = float((i - j) * (j - k) * (k - i) // 2)
    a, b, c, d = xyzw[..., 3], xyzw[..., i], xyzw[..., j], xyzw[..., k] * sign
    ang = np.empty(wxyz.shape[:-1] + (3,))
    ang[..., 1] = 2 * np.arctan2(np.hypot(c, d), np.hypot(a, b))
    case = np.where(np.abs(ang[..., 1] - np.pi) <= 1e-7, 2, 0)
    case = np.where(np.abs(ang[..., 1]) <= 1e-7, 1, case)
    hs, hd = np.arctan2(b, a), np.arctan2(d, c)
    ang[..., 0] = np.where(case == 1, 2 * hs, -2 * hd)
    ang[..., first] = np.where(case == 0, hs - hd, ang[..., first])
    ang[..., third] = np.where(case == 0, hs + hd, ang[..., third])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    return np.rad2deg(ang)


def rotation_and_jacobian(wxyz):
    """R(q / |q|) and dR/dq (shapes (..., 3, 3) and (..., 3, 3, 4)) for the
    sandwich map of SO3.apply (cryojax/rotations/_lie_group.py:96-100) after
    QuaternionPose's normalisation (cryojax/simulator/_pose.py:322-330)."""
    q = np.asarray(wxyz, dtype=np.float64)
    nq = np.linalg.norm(q, axis=-1, keepdims=True)
    u = q / nq
    w, x, y, z = np.moveaxis(u, -1, 0)
    r = np.stack(
        [
            np.stack([w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
            np.stack([2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)], -1),
            np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z], -1),
        ],
        -2,
    )
    du = np.stack(
        [
            np.stack([np.stack([w, -z, y], -1), np.stack([z, w, -x], -1), np.stack([-y, x, w], -1)], -2),
            np.stack([np.stack([x, y, z], -1), np.stack([y, -x, -w], -1), np.stack([z, w, -x], -1)], -2),
            np.stack([np.stack([-y, x, w], -1), np.stack([x, y, z], -1), np.stack([-w, z, -y], -1)], -2),
            np.stack([np.stack([-z, -w, x], -1), np.stack([w, -z, y], -1), np.stack([x, y, z], -1)], -2),
        ],
        -1,
    ) * 2.0
    eye = np.eye(4)
    dudq = (eye - u[..., :, None] * u[..., None, :]) / nq[..., None]
    return r, np.einsum("...abk,...kq->...abq", du, dudq)
