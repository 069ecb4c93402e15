"""Pin the FP64 analytic adjoint with central finite differences of the FP64 forward (CPU)."""

import numpy as np
import pytest

from oracle import adjoint as A
from oracle import imaging as O
from cryojax_amd.synthetic import gaussian_mixture_volume


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(5)
    n = 16
    vol = gaussian_mixture_volume(n, 12, 3, seed=6)
    grid = O.fourier_grid_from_real(vol)
    # add edge-heavy noise so wrap / fill corners carry weight
    grid = grid + 0.05 * np.abs(grid).max() * (rng.standard_normal(grid.shape) + 1j * rng.standard_normal(grid.shape))
    b = 2
    quats = rng.standard_normal((b, 4))
    offs = rng.uniform(-2, 2, (b, 2))
    df = rng.uniform(8000, 12000, b)
    targets = rng.standard_normal((b, n, n)) * 10.0
    return grid, quats, offs, df, targets


def loss(grid, quats, offs, df, targets):
    return A.loss_and_grads(grid, quats, offs, df, targets, 1.3, astigmatism=150.0, astigmatism_angle=20.0)[0]


def test_adjoint_matches_central_differences(problem):
    grid, quats, offs, df, targets = problem
    L, gg, gq, go, gd = A.loss_and_grads(grid, quats, offs, df, targets, 1.3, astigmatism=150.0,
                                         astigmatism_angle=20.0)
    rng = np.random.default_rng(7)
    # grid entries (real and imaginary parts), including the wrapped edge planes
    for _ in range(12):
        idx = tuple(rng.integers(0, 16, 3))
        for part, unit in ((0, 1.0), (1, 1j)):
            h = 1e-3
            gp, gm = grid.copy(), grid.copy()
            gp[idx] += h * unit
            gm[idx] -= h * unit
            fd = (loss(gp, quats, offs, df, targets) - loss(gm, quats, offs, df, targets)) / (2 * h)
            an = gg[idx].real if part == 0 else gg[idx].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (idx, part, fd, an)
    # quaternion components (the loss sees q / |q|)
    for i in range(2):
        for c in range(4):
            h = 1e-6
            qp, qm = quats.copy(), quats.copy()
            qp[i, c] += h
            qm[i, c] -= h
            fd = (loss(grid, qp, offs, df, targets) - loss(grid, qm, offs, df, targets)) / (2 * h)
            assert abs(fd - gq[i, c]) <= 1e-4 * max(1.0, abs(gq[i, c])), (i, c, fd, gq[i, c])
        # radial direction carries no gradient
        assert abs(np.dot(gq[i], quats[i])) < 1e-6 * np.abs(gq[i]).max()
    for i in range(2):
        for c in range(2):
            h = 1e-5
            op, om = offs.copy(), offs.copy()
            op[i, c] += h
            om[i, c] -= h
            fd = (loss(grid, quats, op, df, targets) - loss(grid, quats, om, df, targets)) / (2 * h)
            assert abs(fd - go[i, c]) <= 1e-5 * max(1.0, abs(go[i, c])), (i, c, fd, go[i, c])
        h = 1e-3
        dp, dm = df.copy(), df.copy()
        dp[i] += h
        dm[i] -= h
        fd = (loss(grid, quats, offs, dp, targets) - loss(grid, quats, offs, dm, targets)) / (2 * h)
        assert abs(fd - gd[i]) <= 1e-5 * max(1.0, abs(gd[i])), (i, fd, gd[i])
