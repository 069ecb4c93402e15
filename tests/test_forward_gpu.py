"""GPU parity of the fused forward (cryomi_simulate) against the FP64 oracle.

Bar: per-image relative L2 <= 1e-5 against the float64 oracle on the same
seeded inputs (the reported error per image is printed).
"""

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, ops
from tests.helpers import convention_report, draw, fourier_grid, make_volume, oracle_images, packed_params, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
DEV = "cuda:0"


def run_gpu(grid128, params, flags):
    g = torch.from_numpy(grid128.astype(np.complex64)).to(DEV)
    p = ops.params_to_device(params, DEV)
    out = ops.simulate_images(g, p, flags)
    torch.cuda.synchronize()
    return out.cpu().numpy()


FULL = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION


@pytest.mark.parametrize("n,n_gauss,radius,ps,seed", [(64, 200, 20, 2.0, 1), (48, 60, 12, 1.5, 2),
                                                      (96, 300, 30, 1.0, 3), (128, 500, 40, 1.0, 4),
                                                      (192, 800, 60, 1.0, 5), (54, 100, 13, 1.5, 6),
                                                      (384, 400, 96, 1.0, 7), (512, 400, 128, 1.0, 8),
                                                      (160, 500, 50, 1.0, 9), (80, 200, 25, 2.0, 10),
                                                      (100, 200, 30, 1.5, 11), (120, 300, 36, 1.0, 12),
                                                      (10, 8, 2, 2.0, 13)])
def test_simulate_parity(n, n_gauss, radius, ps, seed):
    """Both out-of-range corner rules, each against the oracle of the same rule; the rule's
    per-image sensitivity is printed beside the error (box 64 is the CPU-runnable case:
    64^3, 200 Gaussians, 20-voxel ball, 2.0 A, offsets within 5 A)."""
    print(f"[parity] box={n} seed={seed}", flush=True)
    vol = make_volume(n, n_gauss, radius, seed)
    g = fourier_grid(vol)
    d = draw(8, seed + 100, offset_max=5.0 if n == 64 else 10.0)
    params = packed_params(d, ps)
    refs = {w: oracle_images(g, d, ps, wrap_negative_indices=w) for w in (True, False)}
    for wrap in (True, False):
        gpu = run_gpu(g, params, FULL | (0 if wrap else _lib.GATHER_FILL))
        err, _ = convention_report(f"box={n} wrap={wrap}", gpu, refs[wrap], refs[not wrap])
        assert np.all(err <= TOL), (wrap, err)


@pytest.mark.parametrize("wrap", [True, False])
def test_simulate_parity_256_subset(wrap):
    """The batched 256^3 -> 256^2 workload: 4096 poses (offsets within 10 A, defocus
    8000-20000 A, astigmatism 0-800 A) on the bricked grid, FP64 oracle on a 64-image subset."""
    n = 256
    print(f"[parity] box=256 images=4096 wrap={wrap}", flush=True)
    vol = make_volume(n, 2000, 80, 7)
    g = fourier_grid(vol)
    d = draw(4096, 11)
    params = packed_params(d, 1.0)
    gt = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    flags = FULL | _lib.GRID_BRICKED | (0 if wrap else _lib.GATHER_FILL)
    gpu = ops.simulate_images(ops.grid_to_bricks(gt), ops.params_to_device(params, DEV), flags, box=n)
    idx = np.random.default_rng(5).choice(4096, 64, replace=False)
    sub = {k: v[idx] for k, v in d.items()}
    gsub = gpu[torch.from_numpy(idx).to(DEV)].cpu().numpy()
    ref = oracle_images(g, sub, 1.0, wrap_negative_indices=wrap)
    other = oracle_images(g, sub, 1.0, wrap_negative_indices=not wrap)
    err, _ = convention_report(f"box=256 wrap={wrap}", gsub, ref, other)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("flags", [0, _lib.APPLY_CTF, _lib.APPLY_TRANSLATION, FULL])
def test_flag_combinations(flags):
    n = 64
    vol = make_volume(n, 200, 20, 21)
    g = fourier_grid(vol)
    d = draw(6, 22, offset_max=5.0)
    gpu = run_gpu(g, packed_params(d, 2.0), flags)
    ref = oracle_images(g, d, 2.0, applies_ctf=bool(flags & _lib.APPLY_CTF),
                        applies_translation=bool(flags & _lib.APPLY_TRANSLATION))
    err = rel_l2(gpu, ref)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("flags", [FULL, _lib.APPLY_CTF])
def test_fourier_output_parity(flags):
    n = 64
    vol = make_volume(n, 200, 20, 31)
    g = fourier_grid(vol)
    d = draw(5, 32, offset_max=5.0)
    gpu = run_gpu(g, packed_params(d, 2.0), flags | _lib.OUTPUT_FOURIER)
    ref = oracle_images(g, d, 2.0, fourier=True, applies_translation=bool(flags & _lib.APPLY_TRANSLATION))
    assert gpu.shape == (5, n, n // 2 + 1)
    err = rel_l2(gpu, ref)
    assert np.all(err <= TOL), err


def test_identity_pose_is_z_projection():
    n = 64
    vol = make_volume(n, 200, 20, 41)
    g = fourier_grid(vol)
    d = {"wxyz": np.array([[1.0, 0, 0, 0]]), "offset": np.zeros((1, 2)), "defocus": np.zeros(1),
         "astigmatism": np.zeros(1), "astigmatism_angle": np.zeros(1)}
    gpu = run_gpu(g, packed_params(d, 1.0), 0)[0]
    proj = vol.astype(np.float64).sum(0)
    assert rel_l2(gpu, proj) < TOL


def test_irfft2_parity(rng):
    n = 128
    x = rng.standard_normal((5, n, n // 2 + 1)) + 1j * rng.standard_normal((5, n, n // 2 + 1))
    t = torch.from_numpy(x.astype(np.complex64)).to(DEV)
    out = ops.irfft2(t, n).cpu().numpy()
    ref = np.fft.fftshift(np.fft.irfftn(x, s=(n, n), axes=(-2, -1)), axes=(-2, -1))
    assert np.all(rel_l2(out, ref) <= TOL)


def test_zero_images_and_errors():
    n = 32
    g = torch.zeros((n, n, n), dtype=torch.complex64, device=DEV)
    p = torch.zeros((0, 20), dtype=torch.float64, device=DEV)
    out = ops.simulate_images(g, p, FULL)
    assert out.shape == (0, n, n)
    g2 = torch.zeros((42, 42, 42), dtype=torch.complex64, device=DEV)  # 42 = 2*3*7: unsupported
    with pytest.raises(NotImplementedError):
        ops.simulate_images(g2, torch.zeros((1, 20), dtype=torch.float64, device=DEV), FULL)


def bricks_reference(g):
    """Host statement of the bricked layout: each (z, y) row padded to entries
    x = -1 .. 3 (n // 3 + 1) - 2 (x = -1 the wrapped x = n - 1, x >= n zero),
    brick b = the four entries x = 3b - 1 .. 3b + 2; bricks of 2x2x4 (z, y, x)
    row-major over (z // 2, y // 2, b), elements (z & 1, y & 1, x) inside."""
    n = g.shape[0]
    nb = n // 3 + 1
    row = np.zeros((n, n, 3 * nb + 1), dtype=g.dtype)  # row[..., xs] holds x = xs - 1
    row[:, :, 1 : n + 1] = g
    row[:, :, 0] = g[:, :, n - 1]
    bricks = np.stack([row[:, :, 3 * b : 3 * b + 4] for b in range(nb)], axis=2)  # (z, y, b, 4)
    b = bricks.reshape(n // 2, 2, n // 2, 2, nb, 4).transpose(0, 2, 4, 1, 3, 5)
    return np.ascontiguousarray(b).reshape(-1)


@pytest.mark.parametrize("n", [8, 48, 64])
def test_grid_to_bricks_layout(n, rng):
    g = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))).astype(np.complex64)
    out = ops.grid_to_bricks(torch.from_numpy(g).to(DEV)).cpu().numpy()
    np.testing.assert_array_equal(out, bricks_reference(g))


@pytest.mark.parametrize("n,seed", [(64, 51), (96, 52), (160, 54), (192, 55), (256, 53)])
def test_bricked_grid_matches_standard(n, seed):
    """Same samples, same accumulation order: the bricked layout is bit-exact (160 and 192 take
    the register former's row offsets for boxes that are not powers of two; 96 the generic
    kernel)."""
    print(f"[bricked] box={n}", flush=True)
    vol = make_volume(n, 200 if n < 256 else 1000, n // 4, seed)
    g = fourier_grid(vol).astype(np.complex64)
    d = draw(8, seed)
    p = ops.params_to_device(packed_params(d, 1.0), DEV)
    gt = torch.from_numpy(g).to(DEV)
    for flags in (FULL, FULL | _lib.OUTPUT_FOURIER, FULL | _lib.GATHER_FILL, FULL | _lib.OUTPUT_FOURIER | _lib.GATHER_FILL):
        std = ops.simulate_images(gt, p, flags)
        brk = ops.simulate_images(ops.grid_to_bricks(gt), p, flags | _lib.GRID_BRICKED, box=n)
        torch.cuda.synchronize()
        assert torch.equal(std, brk)


@pytest.mark.parametrize("n,bricked", [(32, False), (32, True), (48, False), (64, True), (64, False)])
def test_edge_heavy_random_grid(n, bricked, rng):
    """A random complex grid puts full-size values at the box edges, where the
    rotated plane's corners wrap or fill (indices in [-N, -1]) or fill (>= N):
    every fill / wrap / seam case of the gather is exercised at full weight, under
    both out-of-range rules."""
    g = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    g = g.astype(np.complex64).astype(np.complex128)
    d = draw(12, 1000 + n)
    p = ops.params_to_device(packed_params(d, 1.0), DEV)
    gt = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    grid_arg, flags, box = (ops.grid_to_bricks(gt), _lib.GRID_BRICKED, n) if bricked else (gt, 0, None)
    for wrap in (True, False):
        rule = 0 if wrap else _lib.GATHER_FILL
        for extra in (_lib.OUTPUT_FOURIER, 0):
            out = ops.simulate_images(grid_arg, p, FULL | flags | extra | rule, box=box).cpu().numpy()
            ref = oracle_images(g, d, 1.0, fourier=bool(extra), wrap_negative_indices=wrap)
            other = oracle_images(g, d, 1.0, fourier=bool(extra), wrap_negative_indices=not wrap)
            err, sens = convention_report(f"edge box={n} bricked={bricked} wrap={wrap} fourier={bool(extra)}", out,
                                          ref, other)
            assert np.all(err <= TOL), (wrap, extra, err)
            assert np.mean(sens > 1e-3) > 0.5  # the rules really differ on these grids (most poses)


@pytest.mark.parametrize("n,n_images,bricked", [(256, 4096, True), (128, 9000, True), (64, 20000, False)])
def test_plane_ordered_schedule_forms_identical_images(n, n_images, bricked):
    """Large batches run on the plane-ordered schedule (images reordered by slice normal across
    XCDs); every image must equal, bit for bit, the same image formed in a small batch (natural
    order), and land at its own index."""
    vol = make_volume(n, 100, n // 4, 17)
    g = torch.from_numpy(fourier_grid(vol).astype(np.complex64)).to(DEV)
    grid = ops.grid_to_bricks(g) if bricked else g
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | (_lib.GRID_BRICKED if bricked else 0)
    d = draw(n_images, 99, offset_max=4.0)
    prm = ops.params_to_device(packed_params(d, 1.0), DEV)
    print(f"[forward] ordered schedule box={n} images={n_images} bricked={bricked}", flush=True)
    full = ops.simulate_images(grid, prm, flags, box=n)
    idx = np.random.default_rng(1).choice(n_images, 512, replace=False)
    sub = ops.simulate_images(grid, prm[torch.from_numpy(idx).to(DEV)].contiguous(), flags, box=n)
    torch.testing.assert_close(full[torch.from_numpy(idx).to(DEV)], sub, rtol=0, atol=0)
