"""CPU restatement of the pose-proposal CNN (forward, backward, loss, Adam).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The network has no counterpart in the reference (cryoJAX has no learned
pose proposals), so this module restates the build's own definition:
NHWC activations, 3x3 stride-2 padding-1 convolutions as im2col GEMMs
(column (ty*3 + tx)*C + c, padded to Kp), ReLU, a 2048-512-96 MLP, and the
multi-hypothesis pose loss of include/cryomi_cnn.h. `bf16=True` rounds every
tensor the GPU stores in bfloat16 (weights, activations, output gradients)
with round-to-nearest-even, so the GPU path can be compared tightly.

Pinning (tests/test_oracle_cnn.py): backward and loss gradients against
central finite differences in float64; im2col against a direct convolution
(scipy.signal.correlate); Adam against its closed form after one step.
"""

from __future__ import annotations

import numpy as np

CONV = [(1, 32, 64), (32, 64, 32), (64, 128, 16), (128, 128, 8)]  # (c_in, c_out, input side)
FC = [(2048, 512), (512, 96)]
N_HYP = 8
N_OUT = 96


def kpad(c_in):
    return ((9 * c_in + 7) // 8) * 8


def layer_shapes():
    """(name, weight shape, bias length) in parameter order."""
    out = []
    for i, (ci, co, _) in enumerate(CONV):
        out.append((f"conv{i + 1}", (co, kpad(ci)), co))
    for i, (fi, fo) in enumerate(FC):
        out.append((f"fc{i + 1}", (fo, fi), fo))
    return out


def bf16_round(x):
    """float32 -> nearest-even bfloat16 -> float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, (co, k), nb in layer_shapes():
        fan_in = 9 * CONV[int(name[-1]) - 1][0] if name.startswith("conv") else k
        w = np.zeros((co, k), dtype=np.float32)
        scale = np.sqrt(2.0 / fan_in) if name != "fc2" else 0.01
        w[:, :fan_in] = rng.standard_normal((co, fan_in)) * scale
        p[name + ".w"] = w
        p[name + ".b"] = np.zeros(nb, dtype=np.float32)
    return p


def im2col(x, kp):
    """x (n, h, w, c) -> (n * h/2 * w/2, kp)."""
    n, h, w, c = x.shape
    ho, wo = h // 2, w // 2
    xp = np.zeros((n, h + 2, w + 2, c), dtype=x.dtype)
    xp[:, 1:h + 1, 1:w + 1] = x
    cols = np.zeros((n, ho, wo, kp), dtype=x.dtype)
    for ty in range(3):
        for tx in range(3):
            tap = ty * 3 + tx
            cols[..., tap * c:(tap + 1) * c] = xp[:, ty:ty + h:2, tx:tx + w:2, :]
    return cols.reshape(n * ho * wo, kp)


def col2im(dcol, shape):
    """Adjoint of im2col: dcol (n * h/2 * w/2, kp) -> (n, h, w, c)."""
    n, h, w, c = shape
    ho, wo = h // 2, w // 2
    d = dcol.reshape(n, ho, wo, -1)
    dxp = np.zeros((n, h + 2, w + 2, c), dtype=dcol.dtype)
    for ty in range(3):
        for tx in range(3):
            tap = ty * 3 + tx
            dxp[:, ty:ty + h:2, tx:tx + w:2, :] += d[..., tap * c:(tap + 1) * c]
    return dxp[:, 1:h + 1, 1:w + 1]


def forward(p, x, bf16=False, dtype=np.float32):
    """x (n, 64, 64) -> out (n, 96) and the cache for backward."""
    r = bf16_round if bf16 else (lambda a: a)
    a = r(np.asarray(x, dtype=dtype).reshape(x.shape[0], 64, 64, 1))
    cache = {"a0": a}
    for i, (ci, co, s) in enumerate(CONV):
        w = r(p[f"conv{i + 1}.w"].astype(dtype))
        col = im2col(a, w.shape[1])
        y = col @ w.T + p[f"conv{i + 1}.b"].astype(dtype)
        a = r(np.maximum(y, 0)).reshape(x.shape[0], s // 2, s // 2, co)
        cache[f"col{i + 1}"] = col
        cache[f"a{i + 1}"] = a
    f = a.reshape(x.shape[0], -1)
    w1 = r(p["fc1.w"].astype(dtype))
    h = r(np.maximum(f @ w1.T + p["fc1.b"].astype(dtype), 0))
    w2 = r(p["fc2.w"].astype(dtype))
    out = h @ w2.T + p["fc2.b"].astype(dtype)
    cache.update(f=f, h=h)
    return out, cache


def pose_loss(out, rot, off, eps=0.05, off_weight=0.01):
    """Mean over images of the multi-hypothesis loss; returns (loss, dL/dout, winners)."""
    out = np.asarray(out, dtype=np.float64)
    n = out.shape[0]
    R = out[:, :72].reshape(n, N_HYP, 9)
    T = out[:, 72:88].reshape(n, N_HYP, 2)
    logits = out[:, 88:96]
    d = np.sum((R - rot.reshape(n, 1, 9)) ** 2, axis=-1)
    e = np.sum((T - off.reshape(n, 1, 2)) ** 2, axis=-1)
    w = np.argmin(d, axis=1)
    c = np.full((n, N_HYP), eps / (N_HYP - 1))
    c[np.arange(n), w] = 1 - eps
    z = logits - logits.max(axis=1, keepdims=True)
    prob = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
    loss_i = np.sum(c * (d + off_weight * e), axis=1) - np.log(prob[np.arange(n), w])
    g = np.zeros_like(out)
    g[:, :72] = (2 * c[..., None] * (R - rot.reshape(n, 1, 9))).reshape(n, 72) / n
    g[:, 72:88] = (2 * off_weight * c[..., None] * (T - off.reshape(n, 1, 2))).reshape(n, 16) / n
    onehot = np.zeros((n, N_HYP))
    onehot[np.arange(n), w] = 1
    g[:, 88:96] = (prob - onehot) / n
    return float(loss_i.mean()), g, w


def backward(p, cache, dout, bf16=False, dtype=np.float32):
    """Gradients of sum(out * dout) w.r.t. every parameter."""
    r = bf16_round if bf16 else (lambda a: a)
    g = {}
    dout = r(np.asarray(dout, dtype=dtype))
    h, f = cache["h"], cache["f"]
    g["fc2.w"] = dout.T @ h
    g["fc2.b"] = dout.sum(0)
    dh = r((dout @ r(p["fc2.w"].astype(dtype))) * (h > 0))
    g["fc1.w"] = dh.T @ f
    g["fc1.b"] = dh.sum(0)
    dy = r((dh @ r(p["fc1.w"].astype(dtype))) * (f > 0))
    n = dout.shape[0]
    for i in range(len(CONV), 0, -1):
        ci, co, s = CONV[i - 1]
        dy2 = dy.reshape(-1, co)
        col = cache[f"col{i}"]
        g[f"conv{i}.w"] = dy2.T @ col
        g[f"conv{i}.b"] = dy2.sum(0)
        if i > 1:
            dcol = dy2 @ r(p[f"conv{i}.w"].astype(dtype))
            a_in = cache[f"a{i - 1}"]
            dy = r(col2im(dcol, (n, s, s, ci)) * (a_in > 0))
    return g


def adam(p, m, v, g, t, lr=3e-4, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1**t)) / (np.sqrt(v / (1 - b2**t)) + eps)
    return p, m, v


def orthonormalize(r9):
    """Rotation from an unconstrained 3x3 (rows) by Gram-Schmidt on its first two columns."""
    m = np.asarray(r9, dtype=np.float64).reshape(-1, 3, 3)
    a1, a2 = m[..., :, 0], m[..., :, 1]
    b1 = a1 / np.linalg.norm(a1, axis=-1, keepdims=True)
    u2 = a2 - np.sum(b1 * a2, axis=-1, keepdims=True) * b1
    b2 = u2 / np.linalg.norm(u2, axis=-1, keepdims=True)
    b3 = np.cross(b1, b2)
    return np.stack([b1, b2, b3], axis=-1)
