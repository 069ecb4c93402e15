"""Pin the CPU CNN restatement: float64 central differences for the backward
pass and the pose loss, a direct convolution for im2col, Adam's closed form."""

import numpy as np
import pytest
from scipy.signal import correlate

from oracle import cnn as C


def test_im2col_gemm_equals_direct_convolution(rng):
    n, h, c, co = 2, 8, 3, 4
    x = rng.standard_normal((n, h, h, c))
    w = rng.standard_normal((co, 3, 3, c))
    wk = np.zeros((co, C.kpad(c)))
    wk[:, :9 * c] = w.reshape(co, 9 * c)
    y = (C.im2col(x, C.kpad(c)) @ wk.T).reshape(n, h // 2, h // 2, co)
    for b in range(n):
        for o in range(co):
            full = sum(correlate(np.pad(x[b, :, :, k], 1), w[o, :, :, k], mode="valid") for k in range(c))
            np.testing.assert_allclose(y[b, :, :, o], full[::2, ::2], rtol=1e-12, atol=1e-12)


def test_col2im_is_the_adjoint(rng):
    shape = (2, 8, 8, 3)
    x = rng.standard_normal(shape)
    d = rng.standard_normal((2 * 16, C.kpad(3)))
    lhs = np.sum(C.im2col(x, C.kpad(3)) * d)
    rhs = np.sum(x * C.col2im(d, shape))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_backward_matches_central_differences(rng):
    p = {k: v.astype(np.float64) for k, v in C.init_params(3).items()}
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    x = rng.standard_normal((2, 64, 64))
    dout = rng.standard_normal((2, C.N_OUT))
    out, cache = C.forward(p, x, dtype=np.float64)
    g = C.backward(p, cache, dout, dtype=np.float64)

    def f(pp):
        return float(np.sum(C.forward(pp, x, dtype=np.float64)[0] * dout))

    h = 1e-6
    for name in ["conv1.w", "conv2.w", "conv3.b", "conv4.w", "fc1.w", "fc1.b", "fc2.w", "fc2.b"]:
        for _ in range(3):
            idx = tuple(int(rng.integers(0, s)) for s in p[name].shape)
            if name.startswith("conv") and name.endswith(".w"):
                ci = C.CONV[int(name[4]) - 1][0]
                idx = (idx[0], int(rng.integers(0, 9 * ci)))
            pp, pm = dict(p), dict(p)
            pp[name] = p[name].copy()
            pm[name] = p[name].copy()
            pp[name][idx] += h
            pm[name][idx] -= h
            fd = (f(pp) - f(pm)) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, g[name][idx])


def test_pose_loss_gradient_and_winner(rng):
    n = 5
    out = rng.standard_normal((n, C.N_OUT))
    rot = C.orthonormalize(rng.standard_normal((n, 9))).reshape(n, 9)
    off = rng.standard_normal((n, 2))
    loss, g, w = C.pose_loss(out, rot, off, eps=0.05, off_weight=0.3)
    d = np.sum((out[:, :72].reshape(n, 8, 9) - rot[:, None]) ** 2, -1)
    assert np.array_equal(w, np.argmin(d, 1))
    h = 1e-6
    for _ in range(25):
        i, q = int(rng.integers(0, n)), int(rng.integers(0, C.N_OUT))
        op, om = out.copy(), out.copy()
        op[i, q] += h
        om[i, q] -= h
        # the winner must not switch inside the step
        if not (np.array_equal(C.pose_loss(op, rot, off)[2], w) and np.array_equal(C.pose_loss(om, rot, off)[2], w)):
            continue
        fd = (C.pose_loss(op, rot, off, 0.05, 0.3)[0] - C.pose_loss(om, rot, off, 0.05, 0.3)[0]) / (2 * h)
        assert abs(fd - g[i, q]) <= 1e-6, (i, q, fd, g[i, q])


def test_adam_first_step_closed_form(rng):
    p = rng.standard_normal(10)
    g = rng.standard_normal(10)
    p1, m1, v1 = C.adam(p, np.zeros(10), np.zeros(10), g, 1, lr=1e-3, eps=1e-12)
    # after one step m_hat = g and v_hat = g^2, so the update is lr * sign(g)
    np.testing.assert_allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=1e-12)


def test_bf16_round_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5, 3.0e-39], dtype=np.float32)
    r = C.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even (mantissa ...0)
    assert r[2] == np.float32(1.0 + 4 * 2**-8)  # tie -> even (...10)
    assert r[3] == 1.0
    assert r[4] == -2.5


def test_orthonormalize_gives_rotations(rng):
    r = C.orthonormalize(rng.standard_normal((6, 9)))
    for m in r:
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(m) > 0
