"""Hand-written VJP (cryomi_simulate_vjp, torch.autograd over the fused forward)
against the FP64 analytic adjoint and FP64 central differences.

Bar: <= 1e-4 relative error on gradients (grid, quaternion, offsets, defocus).
"""

import numpy as np
import pytest
import torch

from cryojax_amd import autodiff
from cryojax_amd.simulator import wavelength_from_kilovolts
from oracle import adjoint as A
from tests.helpers import fourier_grid, make_volume

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
CONSTS = dict(astigmatism=150.0, astigmatism_angle=20.0, cs_mm=2.7, wavelength=float(wavelength_from_kilovolts(300.0)),
              pixel_size=1.3, phase_shift=float(np.arctan(0.1 / np.sqrt(1 - 0.01))), amplitude=1.0)


def setup(n, b, seed, edge_noise=0.02):
    rng = np.random.default_rng(seed)
    vol = make_volume(n, max(20, n), n // 4, seed)
    g = fourier_grid(vol)
    g = g + edge_noise * np.abs(g).max() * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
    g = g.astype(np.complex64).astype(np.complex128)  # the device grid, exactly
    quats = rng.standard_normal((b, 4))
    offs = rng.uniform(-3, 3, (b, 2))
    df = rng.uniform(8000, 20000, b)
    targets = rng.standard_normal((b, n, n))
    return g, quats, offs, df, targets


def gpu_grads(g, quats, offs, df, targets, wrap=True):
    grid = torch.from_numpy(g.astype(np.complex64)).to(DEV).requires_grad_(True)
    q = torch.from_numpy(quats).requires_grad_(True)
    o = torch.from_numpy(offs).requires_grad_(True)
    d = torch.from_numpy(df).requires_grad_(True)
    imgs = autodiff.simulate(grid, q, o, d, wrap_negative_indices=wrap, **CONSTS)
    loss = ((imgs - torch.from_numpy(targets).float().to(DEV)) ** 2).sum()
    loss.backward()
    return float(loss), grid.grad.cpu().numpy(), q.grad.numpy(), o.grad.numpy(), d.grad.numpy()


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


# boxes up to 256 bin samples by brick (8: one brick spans x and y, so wrapped corner pairs
# share a brick; 12 and 18: ragged last bricks; 256: 16 x 16 x 8 bricks, the largest); 288 and
# 512 (the largest box, narrowest column strips) take the sample-window path
@pytest.mark.parametrize("n,b,seed", [(32, 4, 1), (64, 3, 2), (8, 6, 8), (12, 5, 9), (18, 3, 3), (256, 2, 4),
                                      (288, 1, 5), (512, 1, 7), (160, 2, 10), (10, 4, 11), (320, 1, 12)])
def test_vjp_matches_fp64_adjoint(n, b, seed):
    """Both out-of-range corner rules, each against the FP64 adjoint of the same rule."""
    g, quats, offs, df, targets = setup(n, b, seed)
    ref = {}
    for wrap in (True, False):
        print(f"[vjp] box={n} images={b} wrap={wrap}", flush=True)
        L, gg, gq, go, gd = gpu_grads(g, quats, offs, df, targets, wrap)
        Lr, rgg, rgq, rgo, rgd = A.loss_and_grads(g, quats, offs, df, targets, CONSTS["pixel_size"],
                                                  astigmatism=CONSTS["astigmatism"],
                                                  astigmatism_angle=CONSTS["astigmatism_angle"],
                                                  wrap_negative_indices=wrap)
        ref[wrap] = rgg
        print(f"box={n} wrap={wrap} rel errors: grid {rel(gg, rgg):.2e} quat {rel(gq, rgq):.2e} "
              f"offset {rel(go, rgo):.2e} defocus {rel(gd, rgd):.2e}")
        assert abs(L - Lr) <= 1e-5 * Lr
        assert rel(gg, rgg) <= 1e-4
        for a_, b_ in ((gq, rgq), (go, rgo), (gd, rgd)):
            assert np.all(np.abs(a_ - b_) <= 1e-4 * np.abs(b_).max(axis=-1, keepdims=True) if b_.ndim > 1 else
                          np.abs(a_ - b_) <= 1e-4 * np.abs(b_).max())
    print(f"box={n} grid-gradient wrap-vs-fill sensitivity {rel(ref[False], ref[True]):.2e}")


def test_central_differences_on_32_coordinates():
    """Central differences of the FP64 forward loss at 32 random coordinates
    (grid entries, quaternion, offset, defocus) vs the GPU VJP. The loss is
    linear in the grid (h = 1e-3 is exact up to rounding) but only piecewise
    smooth in the pose: trilinear weights kink where a sample crosses a voxel
    face, and h = 1e-3 on a quaternion moves samples by ~0.03 voxels, across
    many faces. Pose steps are therefore 1e-6 (quaternion), 1e-5 (offsets)
    and 1e-3 (defocus, smooth)."""
    n, b = 32, 2
    g, quats, offs, df, targets = setup(n, b, 9)
    _, gg, gq, go, gd = gpu_grads(g, quats, offs, df, targets)

    def loss(g_, q_, o_, d_):
        return A.loss_and_grads(g_, q_, o_, d_, targets, CONSTS["pixel_size"], astigmatism=CONSTS["astigmatism"],
                                astigmatism_angle=CONSTS["astigmatism_angle"])[0]

    rng = np.random.default_rng(10)
    h = 1e-3
    checked = 0
    scale_g = np.abs(gg).max()
    # 20 grid coordinates near the sampled region (real and imaginary parts)
    m = n // 2
    while checked < 20:
        idx = tuple(int(v) for v in rng.integers(m - 10, m + 10, 3))
        part = checked % 2
        unit = 1.0 if part == 0 else 1j
        gp, gm = g.copy(), g.copy()
        gp[idx] += h * unit
        gm[idx] -= h * unit
        fd = (loss(gp, quats, offs, df) - loss(gm, quats, offs, df)) / (2 * h)
        an = gg[idx].real if part == 0 else gg[idx].imag
        assert abs(fd - an) <= 1e-4 * scale_g, (idx, part, fd, an)
        checked += 1
    # 4 quaternion, 4 offset, 2 defocus, 2 more quaternion coordinates
    for i, c in [(0, 0), (0, 1), (1, 2), (1, 3), (0, 2), (1, 0)]:
        qp, qm = quats.copy(), quats.copy()
        qp[i, c] += 1e-6
        qm[i, c] -= 1e-6
        fd = (loss(g, qp, offs, df) - loss(g, qm, offs, df)) / 2e-6
        assert abs(fd - gq[i, c]) <= 1e-4 * np.abs(gq).max(), (i, c, fd, gq[i, c])
    for i, c in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        op, om = offs.copy(), offs.copy()
        op[i, c] += 1e-5
        om[i, c] -= 1e-5
        fd = (loss(g, quats, op, df) - loss(g, quats, om, df)) / 2e-5
        assert abs(fd - go[i, c]) <= 1e-4 * np.abs(go).max(), (i, c, fd, go[i, c])
    for i in range(2):
        dp, dm = df.copy(), df.copy()
        dp[i] += h
        dm[i] -= h
        fd = (loss(g, quats, offs, dp) - loss(g, quats, offs, dm)) / (2 * h)
        assert abs(fd - gd[i]) <= 1e-4 * np.abs(gd).max(), (i, fd, gd[i])


def test_vjp_1024_images_128_box():
    """1024 images of 128^2 from a 128^3 grid: per-image pose/offset/defocus
    gradients of a 16-image subset vs the FP64 adjoint; the grid gradient of the
    whole batch equals the sum of the grid gradients of four 256-image chunks
    (linearity of the scatter-add)."""
    n, b = 128, 1024
    g, quats, offs, df, targets = setup(n, b, 21, edge_noise=0.0)
    _, gg, gq, go, gd = gpu_grads(g, quats, offs, df, targets)
    sub = np.random.default_rng(3).choice(b, 16, replace=False)
    _, _, rgq, rgo, rgd = A.loss_and_grads(g, quats[sub], offs[sub], df[sub], targets[sub], CONSTS["pixel_size"],
                                           astigmatism=CONSTS["astigmatism"],
                                           astigmatism_angle=CONSTS["astigmatism_angle"])
    for a_, b_ in ((gq[sub], rgq), (go[sub], rgo)):
        assert np.all(np.abs(a_ - b_) <= 1e-4 * np.abs(b_).max(axis=-1, keepdims=True))
    assert np.all(np.abs(gd[sub] - rgd) <= 1e-4 * np.abs(rgd))
    acc = np.zeros_like(gg)
    for c in range(4):
        s = slice(256 * c, 256 * (c + 1))
        acc += gpu_grads(g, quats[s], offs[s], df[s], targets[s])[1]
    assert rel(gg, acc) <= 1e-5


def test_vjp_bins_images_in_chunks():
    """4500 images of a 16^3 grid: the brick lists are built for 2048 images at a time
    (three passes accumulating into one grid gradient); all of it against the FP64 adjoint."""
    n, b = 16, 4500
    g, quats, offs, df, targets = setup(n, b, 6)
    _, gg, gq, go, gd = gpu_grads(g, quats, offs, df, targets)
    _, rgg, rgq, rgo, rgd = A.loss_and_grads(g, quats, offs, df, targets, CONSTS["pixel_size"],
                                             astigmatism=CONSTS["astigmatism"],
                                             astigmatism_angle=CONSTS["astigmatism_angle"])
    print(f"4500 x 16^2 rel errors: grid {rel(gg, rgg):.2e} quat {rel(gq, rgq):.2e}")
    assert rel(gg, rgg) <= 1e-4
    assert np.all(np.abs(gq - rgq) <= 1e-4 * np.abs(rgq).max(axis=-1, keepdims=True))


@pytest.mark.parametrize("wrap", [True, False])
def test_vjp_config_1024_images_snr_half(wrap):
    """The differentiation workload: 1024 images of 128^2 from a 128^3 Gaussian-mixture volume,
    random SO(3) poses and defocus, sum-squared loss against noisy targets at SNR 0.5 (targets
    simulated at nearby true poses, white noise of variance signal variance / 0.5). Per-image
    quaternion / offset / defocus gradients of 64 images, and the grid gradient of those 64
    images, against the FP64 adjoint of the same out-of-range rule."""
    from oracle import imaging as O
    from tests.helpers import draw

    n, b = 128, 1024
    rng = np.random.default_rng(404)
    vol = make_volume(n, 500, 40, 404)
    g = fourier_grid(vol).astype(np.complex64).astype(np.complex128)
    d = draw(b, 405, offset_max=5.0)
    q_true = d["wxyz"]
    quats = q_true + 0.01 * rng.standard_normal(q_true.shape)
    offs = d["offset"] + rng.uniform(-0.5, 0.5, (b, 2))
    df = d["defocus"]
    sub = np.random.default_rng(7).choice(b, 64, replace=False)
    # targets: clean images at the true poses (the GPU forward, checked against the oracle on the
    # subset) plus noise at SNR 0.5
    grid_t = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    with torch.no_grad():
        clean = autodiff.simulate(grid_t, torch.from_numpy(q_true), torch.from_numpy(d["offset"]),
                                  torch.from_numpy(df), wrap_negative_indices=wrap, **CONSTS).cpu().numpy()
    sig = clean.std(axis=(-2, -1), keepdims=True)
    targets = clean + rng.standard_normal(clean.shape) * sig / np.sqrt(0.5)
    print(f"[vjp] config 1024 x 128^2 SNR 0.5 wrap={wrap}", flush=True)
    L, gg_all, gq, go, gd = gpu_grads(g, quats, offs, df, targets, wrap)
    _, gg, _, _, _ = gpu_grads(g, quats[sub], offs[sub], df[sub], targets[sub], wrap)
    Lr, rgg, rgq, rgo, rgd = A.loss_and_grads(g, quats[sub], offs[sub], df[sub], targets[sub], CONSTS["pixel_size"],
                                              astigmatism=CONSTS["astigmatism"],
                                              astigmatism_angle=CONSTS["astigmatism_angle"],
                                              wrap_negative_indices=wrap)
    print(f"1024 x 128^2 wrap={wrap}: 64-image grid gradient rel {rel(gg, rgg):.2e}, quat "
          f"{rel(gq[sub], rgq):.2e}, offset {rel(go[sub], rgo):.2e}, defocus {rel(gd[sub], rgd):.2e}", flush=True)
    assert rel(gg, rgg) <= 1e-4
    for a_, b_ in ((gq[sub], rgq), (go[sub], rgo)):
        assert np.all(np.abs(a_ - b_) <= 1e-4 * np.abs(b_).max(axis=-1, keepdims=True))
    assert np.all(np.abs(gd[sub] - rgd) <= 1e-4 * np.abs(rgd).max())
    assert np.isfinite(gg_all).all()


@pytest.mark.parametrize("n,b,wrap", [(128, 64, True), (128, 64, False), (64, 40, True), (256, 6, True)])
def test_vjp_bricked_grid_equals_standard(n, b, wrap):
    """CRYOMI_GRID_BRICKED: the gather reads the bricked copy (one x pair per (z, y) row); the
    gradients equal those of the standard layout (checked against the FP64 adjoint by the tests
    above) to FP32 rounding: the corner sums run in another order."""
    from cryojax_amd import _lib, ops
    from tests.helpers import draw, packed_params

    g = fourier_grid(make_volume(n, 200, n // 4, 61))
    grid = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    d = draw(b, 62, offset_max=4.0)
    prm = ops.params_to_device(packed_params(d, 1.0), DEV)
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | (0 if wrap else _lib.GATHER_FILL)
    ct = torch.randn((b, n, n), dtype=torch.float32, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3))
    print(f"[vjp] bricked vs standard box={n} images={b} wrap={wrap}", flush=True)
    gg0, gp0 = autodiff.simulate_vjp(grid, prm, flags, ct)
    gg1, gp1 = autodiff.simulate_vjp(grid, prm, flags, ct, bricks=ops.grid_to_bricks(grid))
    assert rel(gg1.cpu().numpy(), gg0.cpu().numpy()) <= 1e-5
    a, c = gp1.cpu().numpy(), gp0.cpu().numpy()
    assert np.all(np.abs(a - c) <= 1e-5 * np.abs(c).max(axis=0, keepdims=True) + 1e-12)


def test_vjp_graph_replay_equals_eager():
    """cryomi_simulate_vjp captured in a HIP graph (its brick counts run on a second stream, forked
    from and joined to the captured one): every replay, with eager work in between, gives the
    eager gradients to FP32 rounding (the list order, and so the summation order, varies)."""
    from cryojax_amd import _lib, ops
    from tests.helpers import draw, packed_params

    n, b = 64, 40
    lib = _lib.load()
    grid = torch.from_numpy(fourier_grid(make_volume(n, 120, n // 4, 71)).astype(np.complex64)).to(DEV)
    prm = ops.params_to_device(packed_params(draw(b, 72, offset_max=4.0), 1.0), DEV)
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION
    ct = torch.randn((b, n, n), dtype=torch.float32, device=DEV, generator=torch.Generator(device=DEV).manual_seed(5))
    ws = torch.empty(lib.cryomi_simulate_vjp_workspace_bytes(n, b), dtype=torch.uint8, device=DEV)
    gg = torch.zeros_like(grid)
    gp = torch.zeros((b, 9), dtype=torch.float64, device=DEV)

    def call():
        gg.zero_()
        gp.zero_()
        st = lib.cryomi_simulate_vjp(grid.data_ptr(), n, b, prm.data_ptr(), flags, ct.data_ptr(), ws.data_ptr(),
                                     ws.numel(), gg.data_ptr(), gp.data_ptr(), ops._stream_ptr(grid.device))
        _lib.check(st, "cryomi_simulate_vjp")

    print(f"[vjp] graph replay box={n} images={b}", flush=True)
    call()
    torch.cuda.synchronize()
    ref_g, ref_p = gg.clone(), gp.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        call()
    for r in range(3):
        gg.fill_(1.0)
        junk = torch.randn(1 << 20, device=DEV).sum()  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        assert rel(gg.cpu().numpy(), ref_g.cpu().numpy()) <= 1e-5, r
        a, c = gp.cpu().numpy(), ref_p.cpu().numpy()
        assert np.all(np.abs(a - c) <= 1e-5 * np.abs(c).max(axis=0, keepdims=True) + 1e-12), r
