"""Uneven (volume, box, pose-count) groups in one persistent launch, its HIP
graph replay, and the tuning knobs: every group equals its own cryomi_simulate
call bit for bit, and a subset of images meets the FP64 oracle at 1e-5."""

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, batching, ops
from tests.helpers import convention_report, draw, fourier_grid, make_volume, oracle_images, packed_params, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def volume(n, seed):
    vol = make_volume(n, max(20, n), n // 4, seed)
    g = fourier_grid(vol).astype(np.complex64)
    return g.astype(np.complex128), torch.from_numpy(g).to(DEV)


def group(n, count, seed, pixel_size=1.5, bricked=True, flags=_lib.APPLY_CTF | _lib.APPLY_TRANSLATION):
    g128, gdev = volume(n, seed)
    d = draw(count, seed + 100)
    p = ops.params_to_device(packed_params(d, pixel_size), DEV)
    grid = ops.grid_to_bricks(gdev) if bricked else gdev
    return batching.VolumeGroup(grid, n, p, flags, bricked), (g128, gdev, d)


def single(gr, gdev):
    flags = gr.flags | (_lib.GRID_BRICKED if gr.bricked else 0)
    return ops.simulate_images(gr.grid, gr.params, flags, box=gr.box)


@pytest.fixture(autouse=True)
def reset_tuning():
    lib = _lib.load()
    yield
    lib.cryomi_set_tuning(b"groups.threads", 256)
    lib.cryomi_set_tuning(b"groups.wg_per_cu", 0)


@pytest.mark.parametrize("bricked", [True, False])
def test_groups_equal_single_launches_and_oracle(bricked):
    specs = [(64, 7, 1), (192, 5, 2), (128, 11, 3), (256, 3, 4), (128, 0, 5)]
    groups, hosts = zip(*[group(n, c, s, bricked=bricked) for n, c, s in specs])
    print("case: groups", specs, "bricked", bricked, flush=True)
    outs = batching.simulate_groups(list(groups))
    torch.cuda.synchronize()
    for gr, (g128, gdev, d), out in zip(groups, hosts, outs):
        if gr.params.shape[0] == 0:
            continue
        ref = single(gr, gdev)
        assert torch.equal(out, ref), f"box {gr.box}: grouped != single launch"
        k = min(2, gr.params.shape[0])
        sub = {key: v[:k] for key, v in d.items()}
        err = rel_l2(out[:k].cpu().numpy(), oracle_images(g128, sub, 1.5))
        print(f"box {gr.box}: rel L2 vs FP64 oracle {err.max():.2e}")
        assert np.all(err <= 1e-5)


@pytest.mark.parametrize("threads,wg", [(512, 0), (256, 1), (512, 1), (256, 2)])
def test_tuning_knobs_change_nothing(threads, wg):
    lib = _lib.load()
    groups, _ = zip(*[group(n, c, s) for n, c, s in [(256, 9, 1), (128, 17, 2), (192, 6, 3)]])
    print("case: threads", threads, "wg_per_cu", wg, flush=True)
    base = [o.clone() for o in batching.simulate_groups(list(groups))]
    _lib.check(lib.cryomi_set_tuning(b"groups.threads", threads), "set")
    _lib.check(lib.cryomi_set_tuning(b"groups.wg_per_cu", wg), "set")
    outs = batching.simulate_groups(list(groups))
    for a, b in zip(base, outs):
        assert torch.equal(a, b)


def test_graph_replays_follow_rewritten_parameters():
    specs = [(128, 40, 11), (64, 25, 12), (192, 8, 13)]
    groups, hosts = zip(*[group(n, c, s) for n, c, s in specs])
    graph = batching.GroupGraph(list(groups))
    try:
        for step in range(4):
            print("case: graph replay", step, flush=True)
            # decode-style step: new poses / CTF written in place into the captured tables
            for i, gr in enumerate(groups):
                d = draw(gr.params.shape[0], 1000 * step + i)
                gr.params.copy_(torch.from_numpy(packed_params(d, 1.5)))
            graph.replay()
            torch.cuda.synchronize()
            got = [gr.out.clone() for gr in groups]
            # eager work between replays
            x = torch.randn(1 << 20, device=DEV)
            (x * 2).sum().item()
            for gr, (_, gdev, _), o in zip(groups, hosts, got):
                assert torch.equal(o, single(gr, gdev)), f"replay {step} box {gr.box}"
    finally:
        graph.close()


@pytest.mark.parametrize("wrap", [True, False])
def test_uneven_batch_12000_poses(wrap):
    """128^3 / 192^3 / 256^3 with 2000 / 4000 / 6000 poses, one launch, under both
    out-of-range corner rules; 8 images per group against the FP64 oracle of the same rule."""
    rule = 0 if wrap else _lib.GATHER_FILL
    specs = [(128, 2000, 21), (192, 4000, 22), (256, 6000, 23)]
    print(f"case: uneven 12000 poses wrap={wrap}", flush=True)
    groups, hosts = zip(*[group(n, c, s, pixel_size=1.0, flags=_lib.APPLY_CTF | _lib.APPLY_TRANSLATION | rule)
                          for n, c, s in specs])
    outs = batching.simulate_groups(list(groups))
    torch.cuda.synchronize()
    for gr, (g128, gdev, d), out in zip(groups, hosts, outs):
        ref = single(gr, gdev)
        assert torch.equal(out, ref)
        idx = np.random.default_rng(gr.box).choice(gr.params.shape[0], 8, replace=False)
        sub = {key: v[idx] for key, v in d.items()}
        err, _ = convention_report(f"uneven box {gr.box} n={gr.params.shape[0]} wrap={wrap}",
                                   out[torch.from_numpy(idx).to(DEV)].cpu().numpy(),
                                   oracle_images(g128, sub, 1.0, wrap_negative_indices=wrap),
                                   oracle_images(g128, sub, 1.0, wrap_negative_indices=not wrap))
        assert np.all(err <= 1e-5)


def test_group_errors():
    g64, _ = group(64, 2, 1)
    g96 = batching.VolumeGroup(torch.zeros((96, 96, 96), dtype=torch.complex64, device=DEV), 96,
                               torch.zeros((1, 20), dtype=torch.float64, device=DEV), bricked=False)
    with pytest.raises(NotImplementedError, match="box"):
        batching.simulate_groups([g64, g96])
    gstd, _ = group(64, 2, 2, bricked=False)
    with pytest.raises(NotImplementedError, match="layout"):
        batching.simulate_groups([g64, gstd])
    with pytest.raises(ValueError):
        batching.simulate_groups([g64] * 9)
    with pytest.raises(ValueError):
        batching.simulate_groups([g64], workspace=torch.empty(16, dtype=torch.uint8, device=DEV))


def test_simulate_image_models_matches_simulate():
    import cryojax_amd.simulator as cxs

    models = []
    for n, b, seed in [(128, 5, 31), (64, 3, 32)]:
        vol = make_volume(n, 40, n // 4, seed)
        volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(vol)
        config = cxs.BasicImageConfig((n, n), 1.5, 300.0)
        d = draw(b, seed)
        pose = cxs.QuaternionPose(d["offset"][:, 0], d["offset"][:, 1], d["wxyz"])
        tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(d["defocus"], d["astigmatism"], d["astigmatism_angle"]))
        models.append(cxs.make_image_model(volume, config, pose, tt))
    outs = batching.simulate_image_models(models)
    for m, o in zip(models, outs):
        assert torch.equal(o, m.simulate())
