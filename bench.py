"""Benchmark: batched fused slice + CTF + shift + irfft2 forward on one MI355X.

Workload (the batched-forward configuration): 256^3 float32 Gaussian-mixture
volume (2000 Gaussians, 80-voxel ball, pixel 1.0 A) stored as a complex64
Fourier grid in HBM; 4096 poses uniform on SO(3), offsets +-10 A, defocus
8000-20000 A, astigmatism 0-800 A; output 4096 real 256x256 float32 images.
One step = one cryomi_simulate launch over all 4096 images (inputs resident
in HBM before timing). Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Beside the timed step, the CPU leg times the FP32 restatement (cpu_baseline)
and the checker reports each image's error against the FP64 oracle on a
64-image subset (parity), both outside the timed region.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 256
N_IMAGES = 4096
PIXEL = 1.0
VOLTAGE = 300.0
AC = 0.1
CS = 2.7
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK = 157.3e12  # dense FP32 matrix FLOP/s
BF16_MFMA_PEAK = 2.5e15  # dense BF16 matrix FLOP/s


def bytes_per_image(n):
    """Algorithmic bytes: 8 complex64 corners per rfft sample + float32 output + params."""
    return 64 * n * (n // 2 + 1) + 4 * n * n + 64


def build_inputs(seed=2026):
    from cryojax_amd.rotations import SO3
    from cryojax_amd.simulator import wavelength_from_kilovolts
    from cryojax_amd import _lib
    from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume

    vol = gaussian_mixture_volume(N, 2000, 80, seed)
    grid128 = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(vol.astype(np.float64))))
    grid = grid128.astype(np.complex64)
    d = draw_image_parameters(N_IMAGES, seed + 1, offset_max=10.0)
    mats = SO3(d["wxyz"]).normalize().sandwich_matrix()
    params = _lib.pack_params(mats.reshape(-1, 9), d["offset"], d["defocus"], d["astigmatism"],
                              d["astigmatism_angle"], CS, wavelength_from_kilovolts(VOLTAGE), PIXEL,
                              np.arctan(AC / np.sqrt(1 - AC * AC)), 1.0, N_IMAGES)
    return grid, params, mats, d, grid128


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, else None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def parity_check(grid_h128, params_h, d, dev, n_check=64, seed=5):
    """The checker: per-image relative L2 of the timed workload's images against the FP64 oracle,
    on a 64-image subset, formed outside the timed region by the same bricked launch; and the
    per-image difference between the two out-of-range corner rules of the gather (the oracle's
    wrap-vs-fill sensitivity), which no reference output can settle."""
    import torch

    from cryojax_amd import _lib, ops
    from oracle import imaging as O

    idx = np.sort(np.random.default_rng(seed).choice(N_IMAGES, n_check, replace=False))
    g64 = torch.from_numpy(grid_h128.astype(np.complex64)).to(dev)
    prm = ops.params_to_device(params_h, dev)
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.GRID_BRICKED
    out = ops.simulate_images(ops.grid_to_bricks(g64), prm, flags, box=N)
    gpu = out[torch.from_numpy(idx).to(dev)].double().cpu().numpy()
    q = O.quat_normalize(d["wxyz"][idx])
    kw = dict(voltage_kv=VOLTAGE, defocus=d["defocus"][idx], astigmatism=d["astigmatism"][idx],
              astigmatism_angle=d["astigmatism_angle"][idx], cs_mm=CS, amplitude_contrast_ratio=AC)
    ref = O.simulate(grid_h128, q, d["offset"][idx], PIXEL, **kw)
    fill = O.simulate(grid_h128, q, d["offset"][idx], PIXEL, wrap_negative_indices=False, **kw)

    def rel(a, b):
        return np.sqrt(np.sum((a - b) ** 2, axis=(1, 2)) / np.sum(b ** 2, axis=(1, 2)))

    err, sens = rel(gpu, ref), rel(fill, ref)
    return {
        "images": n_check,
        "oracle": "FP64 NumPy restatement (complex128 grid before the complex64 rounding)",
        "rel_l2_max": float(err.max()),
        "rel_l2_median": float(np.median(err)),
        "rel_l2_per_image": [float(f"{e:.3e}") for e in err],
        "tolerance": 1e-5,
        "corner_rule": "JAX wrap of indices in [-N, -1] (CRYOMI_GATHER_FILL selects fill-only)",
        "wrap_vs_fill_sensitivity_max": float(sens.max()),
        "wrap_vs_fill_sensitivity_median": float(np.median(sens)),
    }


def cpu_baseline(grid, mats, d, seconds):
    from oracle.cpu_baseline import simulate_fp32

    # bounded sample: as many 64-image batches as fit in ~`seconds`
    done, t0 = 0, time.perf_counter()
    threads = None
    while True:
        s = done % N_IMAGES
        e = s + 64
        _, threads = simulate_fp32(grid, mats[s:e], d["offset"][s:e], d["defocus"][s:e], d["astigmatism"][s:e],
                                   d["astigmatism_angle"][s:e], PIXEL, VOLTAGE, AC, CS)
        done += e - s
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {
        "value": done / dt,
        "unit": "images/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} images of 256x256 from the same 256^3 grid and parameter draws, FP32 NumPy/SciPy "
                  f"restatement of the path, {dt:.1f} s on the host",
    }


SEARCH_BOX = 128
SEARCH_IMAGES = 1000
SEARCH_ROTATIONS = 30000
SEARCH_OFFSET_SIDE = 5
SEARCH_SNR = 0.1


def _observed_set(g64, n_images, seed, dev):
    """Noisy observed images (SNR 0.1) at random poses, as complex128 spectra."""
    import torch

    from cryojax_amd import _lib, ops, search
    from cryojax_amd.synthetic import draw_image_parameters, pack_drawn_parameters

    d = draw_image_parameters(n_images, seed, offset_max=2.0)
    prm_dev = ops.params_to_device(pack_drawn_parameters(d, PIXEL, VOLTAGE, AC, CS), dev)
    clean = ops.simulate_images(g64, prm_dev, _lib.APPLY_CTF | _lib.APPLY_TRANSLATION)
    clean = (clean - clean.mean(dim=(1, 2), keepdim=True)) / clean.std(dim=(1, 2), unbiased=False, keepdim=True)
    gen = torch.Generator(device=dev).manual_seed(seed + 1)
    noisy = np.sqrt(SEARCH_SNR) * clean + torch.randn(clean.shape, generator=gen, device=dev)
    y = search.observed_spectrum(noisy)
    return d, prm_dev, y


def pose_search_bench(dev, steps):
    """Plain exhaustive GPU grid search.

    1000 observed 128x128 images simulated (fused forward) at random poses
    from a 128^3 Gaussian-mixture volume, normalised, plus white Gaussian
    noise at SNR 0.1; scored by NCC against 30000 super-Fibonacci rotations x
    5 x 5 offsets (1 A steps). Inputs (grids, observed spectra, grids of
    poses) are resident before timing; one step = one cryomi_ncc_search call.
    """
    import torch

    from cryojax_amd import search
    from cryojax_amd.rotations import SO3
    from cryojax_amd.synthetic import gaussian_mixture_volume

    n = SEARCH_BOX
    vol = gaussian_mixture_volume(n, 1000, 40, 77)
    g128 = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(vol.astype(np.float64))))
    g64 = torch.from_numpy(g128.astype(np.complex64)).to(dev)
    g128_dev = torch.from_numpy(g128).to(dev)
    d, prm_dev, y = _observed_set(g64, SEARCH_IMAGES, 78, dev)
    quats = search.super_fibonacci_quaternions(SEARCH_ROTATIONS)
    rmats = SO3(quats).sandwich_matrix().reshape(-1, 9)
    rots = torch.from_numpy(rmats).to(dev)
    offs = torch.from_numpy(search.offset_grid(SEARCH_OFFSET_SIDE, 1.0)).to(dev)
    search.ncc_search(g64, g128_dev, y, prm_dev, rots, offs)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bi, bs, over = search.ncc_search(g64, g128_dev, y, prm_dev, rots, offs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    best = bi.cpu().numpy()
    n_off = SEARCH_OFFSET_SIDE**2
    q_best = quats[best // n_off]
    dots = np.abs(np.sum(q_best * d["wxyz"] / np.linalg.norm(d["wxyz"], axis=1, keepdims=True), axis=1))
    err_deg = np.rad2deg(2 * np.arccos(np.clip(dots, -1, 1)))
    kdim = 2 * n * (n // 2 + 1)
    flops = 2.0 * SEARCH_IMAGES * n_off * SEARCH_ROTATIONS * kdim
    plain = {
        "metric": "pose-search images/s (plain exhaustive GPU grid search)",
        "value": round(SEARCH_IMAGES / dt, 1),
        "unit": "images/s",
        "seconds_per_step": round(dt, 4),
        "workload": f"{SEARCH_IMAGES} x {n}x{n} images at SNR {SEARCH_SNR}, {SEARCH_ROTATIONS} rotations x "
                    f"{SEARCH_OFFSET_SIDE}x{SEARCH_OFFSET_SIDE} offsets, NCC, FP32 MFMA + FP64 re-scoring",
        "contraction_tflops": round(flops / dt / 1e12, 1),
        "contraction_frac_of_fp32_mfma_peak": round(flops / dt / FP32_MFMA_PEAK, 3),
        "overflow_images": int(over.sum().item()),
        "median_error_to_truth_deg": round(float(np.median(err_deg)), 2),
        "fraction_within_7p5deg_of_truth": round(float(np.mean(err_deg < 7.5)), 3),
    }

    return plain


# per-image FLOPs of the pose network (algorithmic: K = 9 c_in per 3x3 layer): forward, and
# forward + weight gradients + data gradients (no data gradient into the input image)
def _cnn_flops_per_image():
    from cryojax_amd import cnn

    fwd = 0.0
    conv1 = 0.0
    for i, (ci, co, s) in enumerate(cnn.CONV):
        f = 2.0 * (s // 2) ** 2 * co * 9 * ci
        fwd += f
        conv1 = f if i == 0 else conv1
    for fi, fo in cnn.FC:
        fwd += 2.0 * fi * fo
    return fwd, 3 * fwd - conv1


def learned_search_bench(dev, grid_h, grid_h128, steps, train_images=2_000_000):
    """Learned-proposal pose search against the plain exhaustive search on the 256^3 volume.

    The pose network trains on 2 M images the fused simulator forms on the fly from the
    256^3 Gaussian-mixture volume (SNR uniform 0.05-0.5; only the 64^2 band the
    fourier_crop_downsample input reads is simulated). Observed set: 1000 noisy 256x256 images
    (SNR 0.1) at random poses. Both searches score NCC against 30000 super-Fibonacci rotations x
    5 x 5 offsets (1 A steps); inputs are resident before timing. One plain step = one
    cryomi_ncc_search call; one learned step = network input (fourier_crop_downsample + phase
    flip) + proposals + exact scoring of the grid rotations within 15 deg of the eight proposals
    + grid hill climb (12 deg balls) + exhaustive search for the images whose best -NCC lies
    above the threshold calibrated on a separate set's misses. Accuracy is judged against the
    exhaustive argmax.
    """
    import torch

    from cryojax_amd import learned, search
    from cryojax_amd.rotations import SO3

    n = int(grid_h.shape[0])
    g64 = torch.from_numpy(grid_h).to(dev)
    g128_dev = torch.from_numpy(grid_h128).to(dev)
    d, prm_dev, y = _observed_set(g64, SEARCH_IMAGES, 78, dev)
    quats = search.super_fibonacci_quaternions(SEARCH_ROTATIONS)
    rmats = SO3(quats).sandwich_matrix().reshape(-1, 9)
    rots = torch.from_numpy(rmats).to(dev)
    offs = torch.from_numpy(search.offset_grid(SEARCH_OFFSET_SIDE, 1.0)).to(dev)
    n_off = SEARCH_OFFSET_SIDE**2
    bank = search.template_bank(g64, rots, True)
    search.ncc_search(g64, g128_dev, y, prm_dev, rots, offs, bank=bank)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bi, bs, over = search.ncc_search(g64, g128_dev, y, prm_dev, rots, offs, bank=bank)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    best = bi.cpu().numpy()
    _progress(f"plain search at {n}: {dt:.3f} s per {SEARCH_IMAGES} images")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trainer = learned.ProposalTrainer(g64, PIXEL, batch=512, lr=3e-4, seed=1)
    trainer.train(train_images, lr_schedule=learned.cosine_schedule(3e-4, warmup=200))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    trained = trainer.step * trainer.batch
    _progress(f"trained on {trained} images in {t_train:.2f} s")
    fwd_flops, train_flops = _cnn_flops_per_image()
    # the network alone (forward + loss + backward + Adam on one resident batch): its GEMM rate
    net = trainer.net
    xb = trainer._x
    rot_t = torch.eye(3, device=dev).reshape(1, 9).repeat(trainer.batch, 1)
    off_t = torch.zeros((trainer.batch, 2), device=dev)
    for _ in range(3):
        out = net.forward(xb, trainer.batch)
        dout, _ = net.loss(out, rot_t, off_t)
        net.backward(dout)
    torch.cuda.synchronize()
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        out = net.forward(xb, trainer.batch)
        dout, _ = net.loss(out, rot_t, off_t)
        net.backward(dout)
    torch.cuda.synchronize()
    t_net = (time.perf_counter() - t0) / reps
    net_tflops = train_flops * trainer.batch / t_net / 1e12

    ls = learned.LearnedPoseSearch(net, g64, g128_dev, rots, offs, first_radius=15.0, refine_radius=12.0,
                                   refine_steps=6, cap=1024, refine_cap=256, bank=bank)
    _, cprm, cy = _observed_set(g64, SEARCH_IMAGES, 178, dev)
    ci, _, _ = search.ncc_search(g64, g128_dev, cy, cprm, rots, offs, bank=bank)
    calib = ls.calibrate(cy, cprm, ci)
    # misses before the fallback, on the measured set (untimed)
    ni, nsc, _ = ls.search(y, prm_dev, fallback=False)
    ni = ni.cpu().numpy()
    nsc = nsc.cpu().numpy()
    raw_miss = ni != best
    ls.search(y, prm_dev)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        li, lsc, info = ls.search(y, prm_dev)
    torch.cuda.synchronize()
    dtl = (time.perf_counter() - t0) / steps
    li = li.cpu().numpy()

    def angle(a, b):
        r1 = rmats[a // n_off].reshape(-1, 3, 3)
        r2 = rmats[b // n_off].reshape(-1, 3, 3)
        return np.degrees(np.arccos(np.clip((np.einsum("nij,nij->n", r1, r2) - 1) / 2, -1, 1)))

    ang = angle(li, best)
    fb = info["fallback_mask"].cpu().numpy()
    return {
        "metric": "pose-search images/s (learned proposals + exact local scoring + exhaustive fallback)",
        "value": round(SEARCH_IMAGES / dtl, 1),
        "unit": "images/s",
        "seconds_per_step": round(dtl, 5),
        "plain_images_per_s": round(SEARCH_IMAGES / dt, 1),
        "plain_seconds_per_step": round(dt, 4),
        "speedup_vs_plain": round(dt / dtl, 2),
        "workload": f"{SEARCH_IMAGES} x {n}x{n} images at SNR {SEARCH_SNR} from the {n}^3 volume, "
                    f"{SEARCH_ROTATIONS} rotations x {SEARCH_OFFSET_SIDE}x{SEARCH_OFFSET_SIDE} offsets",
        "fraction_within_5deg_of_exhaustive_argmax": round(float(np.mean(ang <= 5.0)), 4),
        "fraction_identical_to_exhaustive_argmax": round(float(np.mean(li == best)), 4),
        "before_fallback": {
            "fraction_within_5deg": round(float(np.mean(angle(ni, best) <= 5.0)), 4),
            "misses": int(raw_miss.sum()),
            "misses_caught": int((raw_miss & fb).sum()),
        },
        "fallback_images": int(info["fallback_images"]),
        "refine_rounds": int(info["refine_rounds"]),
        "mean_first_stage_candidates": round(float(info["candidates_first"].float().mean().item()), 1),
        "network": "4x conv3x3/s2 (32/64/128/128) + MLP 2048-512-96 (8 hypotheses), BF16 MFMA, FP32 accumulate; "
                   "input fourier_crop_downsample to 64x64",
        "training": {"images": trained, "seconds": round(t_train, 2), "images_per_s": round(trained / t_train, 0),
                     "batch": trainer.batch, "optimizer": "Adam 3e-4", "snr": [0.05, 0.5],
                     "flops_per_image": train_flops,
                     "tflops_whole_step": round(train_flops * trained / t_train / 1e12, 1)},
        "network_step": {"ms": round(t_net * 1e3, 3), "tflops": round(net_tflops, 1),
                         "frac_of_bf16_mfma_peak": round(net_tflops * 1e12 / BF16_MFMA_PEAK, 4),
                         "what": "forward + loss + backward (all BF16 MFMA GEMMs) on one resident batch of 512"},
        "calibration": calib,
    }


UNEVEN = [(128, 2000), (192, 4000), (256, 6000)]


def uneven_batch_bench(dev, grid256_h, steps):
    """Uneven batch in one persistent launch, autotuned, and its graph replay.

    Gaussian-mixture volumes at 128^3 / 192^3 / 256^3 (2000 Gaussians, ball
    radius 80 N / 256 voxels; the 256^3 one is the batched-forward volume)
    with 2000 / 4000 / 6000 poses and matching image sizes, one
    cryomi_simulate_groups launch per step. The autotuner picks the grouped
    kernel's workgroup size and workgroups per CU first (timed on this
    workload); the same step is then timed as three cryomi_simulate launches
    and as a replay of the captured HIP graph.
    """
    import torch

    from cryojax_amd import _lib, autotune, batching, ops
    from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume, pack_drawn_parameters

    groups = []
    for i, (n, count) in enumerate(UNEVEN):
        if n == N:
            gh = grid256_h
        else:
            vol = gaussian_mixture_volume(n, 2000, 80 * n // 256, 500 + i)
            gh = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(vol.astype(np.float64)))).astype(np.complex64)
        d = draw_image_parameters(count, 600 + i, offset_max=10.0)
        p = ops.params_to_device(pack_drawn_parameters(d, PIXEL), dev)
        groups.append(batching.VolumeGroup(ops.grid_to_bricks(torch.from_numpy(gh).to(dev)), n, p))
    cache_path = os.environ.get("CRYOMI_TUNING_CACHE",
                                os.path.join(tempfile.gettempdir(), "cryomi_tuning.json"))
    tuned = autotune.tune_groups(groups, reps=3, warmup=1, cache=autotune.TuningCache(cache_path), use_cache=False)
    lib = _lib.load()
    import ctypes

    arr, _ = batching._table(groups)
    ws = torch.empty(lib.cryomi_simulate_groups_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(groups)),
                     dtype=torch.uint8, device=dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.GRID_BRICKED
    ms_group = timed(lambda: batching.simulate_groups(groups, ws))
    ms_seq = timed(lambda: [ops.simulate_images(g.grid, g.params, flags, out=g.out, box=g.box) for g in groups])
    graph = batching.GroupGraph(groups)
    ms_graph = timed(graph.replay)
    graph.close()
    n_img = sum(c for _, c in UNEVEN)
    alg = sum(c * bytes_per_image(n) for n, c in UNEVEN)
    return {
        "metric": "uneven-batch images/s (one persistent launch)",
        "value": round(n_img / ms_group * 1e3, 1),
        "unit": "images/s",
        "workload": "128^3/192^3/256^3 complex64 grids (bricked), 2000/4000/6000 poses, images at the volume box",
        "ms_per_launch": round(ms_group, 4),
        "algorithmic_gbs": round(alg / ms_group / 1e6, 1),
        "frac_of_hbm_peak": round(alg / ms_group / 1e6 / HBM_PEAK_GBS, 4),
        "autotuned": tuned.best,
        "autotune_table_ms": [[c, None if m is None else round(m, 4)] for c, m in tuned.table],
        "ms_three_separate_launches": round(ms_seq, 4),
        "ms_graph_replay": round(ms_graph, 4),
    }


PUB_VOLUME = 80
PUB_PAD_SCALE = 2
PUB_BATCH = 4096
PUB_REFERENCE_MS = 1.11  # eqx.filter_jit(simulate), one image (docs/examples/jax-transformations.ipynb)


def published_setup_bench(dev, steps=10):
    """The reference's one published timing, shape for shape: an 80^3 volume turned into a 160^3
    Fourier grid by from_real_voxel_grid(pad_scale=2), 80x80 images padded to the grid's 160x160,
    AstigmaticCTF with amplitude contrast 0.1, normalizes_signal=True (the notebook's own GroEL
    map is not in the reference's checkout, so the volume is a synthetic 80^3 Gaussian mixture).
    Timed three ways: the public API for one image (make_image_model(...).simulate(), wall clock,
    the counterpart of the notebook's per-image timing); the public API on a 4096-pose batch; and
    the device launches alone on resident parameters (fused 160^2 forward, then the crop to 80^2
    and its normalisation), HIP events on the launch stream."""
    import torch

    import cryojax_amd.simulator as cxs
    from cryojax_amd import _lib, ops
    from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume

    real = gaussian_mixture_volume(PUB_VOLUME, 500, 25, 901)
    volume = cxs.FourierVoxelGridVolume.from_real_voxel_grid(real, pad_scale=PUB_PAD_SCALE, device=dev)
    pbox = volume.shape[0]
    config = cxs.BasicImageConfig((PUB_VOLUME, PUB_VOLUME), 1.0, VOLTAGE, pad_options=dict(shape=volume.shape[0:2]))
    tt = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(10000.0, -100.0, 10.0), amplitude_contrast_ratio=AC)
    one = cxs.make_image_model(volume, config, cxs.EulerAnglePose(5.0, -3.0, 20.0, 80.0, -5.0), tt,
                               normalizes_signal=True)
    for _ in range(5):
        one.simulate()
    torch.cuda.synchronize()
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        one.simulate()
        torch.cuda.synchronize()
    single_ms = (time.perf_counter() - t0) / reps * 1e3
    d = draw_image_parameters(PUB_BATCH, 902, offset_max=10.0)
    pose = cxs.QuaternionPose(d["offset"][:, 0], d["offset"][:, 1], d["wxyz"])
    ttb = cxs.ContrastTransferTheory(cxs.AstigmaticCTF(d["defocus"], d["astigmatism"], d["astigmatism_angle"]),
                                     amplitude_contrast_ratio=AC)
    batch = cxs.make_image_model(volume, config, pose, ttb, normalizes_signal=True)
    batch.simulate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.simulate()
    torch.cuda.synchronize()
    api_ms = (time.perf_counter() - t0) / steps * 1e3
    flags, (p, _) = batch._flags_and_params(True)
    prm = ops.params_to_device(p, dev)
    grid = volume.bricked_grid()
    stream = torch.cuda.current_stream(dev)
    fwd_ms = _event_ms(lambda: ops.simulate_images(grid, prm, flags | _lib.GRID_BRICKED, box=pbox), steps, stream)
    padded = ops.simulate_images(grid, prm, flags | _lib.GRID_BRICKED, box=pbox)
    crop_ms = _event_ms(lambda: ops.resize_images(padded, (PUB_VOLUME, PUB_VOLUME), normalize=True), steps, stream)
    b = bytes_per_image(pbox)
    dev_ms = fwd_ms + crop_ms
    return {
        "workload": f"{PUB_VOLUME}^3 synthetic volume, pad_scale={PUB_PAD_SCALE} -> {pbox}^3 complex64 grid; "
                    f"{PUB_VOLUME}x{PUB_VOLUME} images padded to {pbox}x{pbox}, AstigmaticCTF, shift, "
                    "normalizes_signal=True",
        "reference_published_ms_per_image": PUB_REFERENCE_MS,
        "reference_note": "eqx.filter_jit(simulate) on one image, hardware not stated",
        "api_single_image_ms": round(single_ms, 4),
        "api_batch_images": PUB_BATCH,
        "api_batch_us_per_image": round(api_ms * 1e3 / PUB_BATCH, 4),
        "device_us_per_image": round(dev_ms * 1e3 / PUB_BATCH, 5),
        "device_images_per_s": round(PUB_BATCH / dev_ms * 1e3, 1),
        "forward_ms": round(fwd_ms, 4),
        "crop_normalize_ms": round(crop_ms, 4),
        "forward_frac_of_hbm_peak": round(PUB_BATCH * b / fwd_ms / 1e6 / HBM_PEAK_GBS, 4),
        "speedup_vs_published_per_image": round(PUB_REFERENCE_MS / (dev_ms / PUB_BATCH), 1),
    }


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)




def _event_ms(fn, steps, stream):
    """Average HIP-event duration of fn() on `stream` over `steps` calls (after one warm call)."""
    import torch

    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def ndimage_bench(dev, images, steps=5):
    """Device rows beside the forward, each timed with HIP events on the launch stream.

    - fourier_crop_downsample of the step's 4096 x 256^2 images to 64^2 (the CNN input size);
    - the FP64 centred 3D DFT of a 256^3 real volume (from_real_voxel_grid);
    - the FP64 rfft2 of 4096 observed 256^2 images (the search's input spectra).
    All three run Stockham FFT kernels and are priced by their algorithmic bytes (real input read,
    outputs written) against the HBM peak; the two FP64 transforms run in several in-place passes
    over their output, so their moved bytes (every pass's reads and writes) are reported beside.
    """
    import torch

    from cryojax_amd import ndimage, ops

    stream = torch.cuda.current_stream(dev)
    n, h = int(images.shape[0]), int(images.shape[-1])
    o = 64
    ds_ms = _event_ms(lambda: ndimage.fourier_crop_downsample_images(images, (o, o)), steps, stream)
    ds_bytes = n * 4 * (h * h + o * o)  # float32 images in, float32 downsampled images out
    vol = torch.randn((h, h, h), device=dev, generator=torch.Generator(device=dev).manual_seed(5))
    ft_ms = _event_ms(lambda: ops.fourier_voxel_grid(vol), steps, stream)
    ft_bytes = h ** 3 * (4 + 16 + 8)  # float32 volume in, complex128 and complex64 grids out
    # three in-place passes: x reads float32 and writes complex128, y reads and writes complex128,
    # z reads complex128 and writes complex128 + complex64
    ft_moved = h ** 3 * (4 + 16 + 32 + 16 + 16 + 8)
    kw = h // 2 + 1
    rf_ms = _event_ms(lambda: ops.rfft2_f64(images), steps, stream)
    rf_bytes = n * (4 * h * h + 16 * h * kw)  # float32 images in, complex128 half-planes out
    rf_moved = rf_bytes + n * 2 * 16 * h * kw  # + the column pass's in-place read and write
    return {
        "fourier_crop_downsample": {
            "workload": f"{n} x {h}x{h} float32 -> {o}x{o}", "ms": round(ds_ms, 4),
            "images_per_s": round(n / ds_ms * 1e3, 1), "bound": "hbm (fused FFT kernel; LDS/latency in practice)",
            "algorithmic_gbs": round(ds_bytes / ds_ms / 1e6, 1),
            "frac_of_hbm_peak": round(ds_bytes / ds_ms / 1e6 / HBM_PEAK_GBS, 3)},
        "fourier_voxel_grid_fp64": {
            "workload": f"{h}^3 float32 -> complex128 + complex64", "ms": round(ft_ms, 3), "bound": "hbm",
            "algorithmic_gbs": round(ft_bytes / ft_ms / 1e6, 1),
            "frac_of_hbm_peak": round(ft_bytes / ft_ms / 1e6 / HBM_PEAK_GBS, 3),
            "moved_gbs": round(ft_moved / ft_ms / 1e6, 1),
            "moved_frac_of_hbm_peak": round(ft_moved / ft_ms / 1e6 / HBM_PEAK_GBS, 3)},
        "observed_rfft2_fp64": {
            "workload": f"{n} x {h}x{h} float32 -> complex128 half-planes", "ms": round(rf_ms, 3),
            "images_per_s": round(n / rf_ms * 1e3, 1), "bound": "hbm",
            "algorithmic_gbs": round(rf_bytes / rf_ms / 1e6, 1),
            "frac_of_hbm_peak": round(rf_bytes / rf_ms / 1e6 / HBM_PEAK_GBS, 3),
            "moved_gbs": round(rf_moved / rf_ms / 1e6, 1),
            "moved_frac_of_hbm_peak": round(rf_moved / rf_ms / 1e6 / HBM_PEAK_GBS, 3)},
    }


def vjp_bytes_per_image(n):
    """Algorithmic bytes of one image's VJP: cotangent read, g_F write + read and g_P write + read
    (complex64 half-planes), the 8 corner operands re-read for the pose / defocus derivatives, and
    the parameters in / gradients out. The grid gradient (N^3 complex64, read + written once per
    call) is not per image and is left out."""
    nh = n * (n // 2 + 1)
    return 4 * n * n + 4 * 8 * nh + 64 * nh + 160 + 72


def adjoint_bench(dev, steps=5, n=128, n_images=1024, snr=0.5, seed=11):
    """cryomi_simulate_vjp on the adjoint workload: 1024 images of 128^2 from a 128^3 grid, cotangent
    = simulated - observed (the Gaussian log-likelihood's gradient at unit variance, SNR 0.5)."""
    import torch

    from cryojax_amd import _lib, autodiff, ops
    from cryojax_amd.rotations import SO3
    from cryojax_amd.simulator import wavelength_from_kilovolts
    from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume

    vol = torch.from_numpy(gaussian_mixture_volume(n, 1000, 40, seed).astype(np.float32)).to(dev)
    _, g64 = ops.fourier_voxel_grid(vol, complex128=False)
    d = draw_image_parameters(n_images, seed + 1, offset_max=5.0)
    mats = SO3(d["wxyz"]).normalize().sandwich_matrix()
    params = ops.params_to_device(_lib.pack_params(
        mats.reshape(-1, 9), d["offset"], d["defocus"], d["astigmatism"], d["astigmatism_angle"], CS,
        wavelength_from_kilovolts(VOLTAGE), 2.0, np.arctan(AC / np.sqrt(1 - AC * AC)), 1.0, n_images), dev)
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION
    sim = ops.simulate_images(g64, params, flags)
    unit = ops.normalize_images(sim.clone())
    noise = torch.randn(sim.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(seed))
    observed = ops.noisy_images(unit, noise, scale=float(np.sqrt(snr)))
    cot = (float(np.sqrt(snr)) * unit - observed).contiguous()  # input preparation, outside the timed region
    bricks = ops.grid_to_bricks(g64)  # the volume's bricked copy, as the forward keeps it
    stream = torch.cuda.current_stream(dev)
    ms = _event_ms(lambda: autodiff.simulate_vjp(g64, params, flags, cot, bricks=bricks), steps, stream)
    b = vjp_bytes_per_image(n)
    achieved = n_images * b / ms / 1e6
    return {
        "workload": f"{n_images} x {n}x{n} images from a {n}^3 complex64 grid (bricked copy for the gather), "
                    "cotangent = simulated - observed "
                    f"(SNR {snr}); grid gradient (complex64, accumulated per voxel brick on chip) + per-image "
                    "pose/offset/defocus gradients",
        "ms": round(ms, 3), "images_per_s": round(n_images / ms * 1e3, 1),
        "bound": "valu / lds (brick-local accumulation of the transposed gather); HBM traffic is minor",
        "bytes_per_image": b, "achieved_gbs": round(achieved, 1), "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 3),
    }


def copy_calibration():
    """A plain copy of known size: the traffic counters read for it give the
    bytes-per-count scale of FETCH_SIZE and WRITE_SIZE on this part."""
    import torch

    nbytes = 256 * 1024 * 1024
    a = torch.full((nbytes // 4,), 1.0, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-search", action="store_true", help="skip the pose-search measurement")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--no-ndimage", action="store_true", help="skip the downsample / FFT measurements")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device-to-device copies (the reference for the FETCH_SIZE / "
                         "WRITE_SIZE counter scale) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (images shard trivially; "
              "see the sharding design). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from cryojax_amd import _lib, ops

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    grid_h, params_h, mats, d, grid_h128 = build_inputs()
    grid = ops.grid_to_bricks(torch.from_numpy(grid_h).to(dev))
    params = ops.params_to_device(params_h, dev)
    flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.GRID_BRICKED
    out = torch.empty((N_IMAGES, N, N), dtype=torch.float32, device=dev)
    # launch shape (persistent workgroups per CU) from the autotuner, on this workload, before the
    # warm-up; every candidate must form bit-identical images
    from cryojax_amd import autotune

    cache_path = os.environ.get("CRYOMI_TUNING_CACHE", os.path.join(tempfile.gettempdir(), "cryomi_tuning.json"))
    tuned = autotune.tune_forward(grid, N, params, flags, wg_per_cu=(0, 2, 3, 4), reps=10, warmup=3,
                                  cache=autotune.TuningCache(cache_path), use_cache=False)
    _progress(f"forward launch shape: {tuned.best}")
    ws = torch.empty(_lib.load().cryomi_simulate_workspace_bytes(N, N_IMAGES, flags), dtype=torch.uint8,
                     device=dev)

    for _ in range(args.warmup):
        ops.simulate_images(grid, params, flags, out=out, workspace=ws, box=N)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        ops.simulate_images(grid, params, flags, out=out, workspace=ws, box=N)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3
    value = N_IMAGES * args.steps / wall

    b = bytes_per_image(N)
    achieved = N_IMAGES * b / avg_launch_s / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": "reg_image_kernel<256, false, true>",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": read_traffic("reg_image_kernel_256_traffic.txt"),
        "bytes_per_image": b,
        "avg_launch_ms": round(avg_launch_s * 1e3, 4),
        "autotuned": tuned.best,
        "autotune_table_ms": [[c, None if m is None else round(m, 4)] for c, m in tuned.table],
    }
    _progress(f"forward: {avg_launch_s * 1e3:.3f} ms per launch")
    side = None
    if not args.no_ndimage:
        side = ndimage_bench(dev, out)
        _progress("downsample / FFT rows timed")
        side["adjoint"] = adjoint_bench(dev)
        _progress("adjoint timed")
    published = None
    if not args.no_ndimage:
        published = published_setup_bench(dev)
        _progress("published setup timed")
    uneven = None
    if not args.no_uneven:
        uneven = uneven_batch_bench(dev, grid_h, steps=10)
        _progress("uneven batch timed")
    pose = learned_pose = None
    if not args.no_search:
        pose = pose_search_bench(dev, steps=2)
        _progress("plain pose search timed")
        learned_pose = learned_search_bench(dev, grid_h, grid_h128, steps=2)
        _progress("learned pose search timed")
    cpu = parity = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(grid_h, mats, d, args.cpu_seconds)
        _progress("cpu baseline timed")
        parity = parity_check(grid_h128, params_h, d, dev)
        _progress(f"parity: max rel L2 {parity['rel_l2_max']:.2e} over {parity['images']} images")
    del grid_h128
    line = {
        "metric": "slice+CTF images/s at 256\u00b3/256\u00b2; pose-search imgs/s vs plain grid at =accuracy",
        "value": round(value, 1),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 coordinates and CTF phase)",
        "data": "synthetic",
        "config": {
            "workload": "batched fused forward: 256^3 complex64 Fourier grid (2000-Gaussian mixture, 1.0 A, "
                        "bricked layout), "
                        "4096 poses -> 4096 x 256x256 float32 images",
            "box": N,
            "images_per_step": N_IMAGES,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "uneven_batch": uneven,
        "pose_search": pose,
        "learned_pose_search": learned_pose,
        "ndimage": side,
        "published_setup": published,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
