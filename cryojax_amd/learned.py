"""Learned pose proposals: training data on the fly, training, inference inputs.

The network (cryojax_amd.cnn.PoseNet) sees 64x64 images: the observed image
downsampled by fourier_crop_downsample (cryojax/ndimage/_downsample.py:91-222,
the real branch's centred Hartley crop; cryomi_crop_spectrum with
CRYOMI_CROP_HARTLEY forms rfftn of its output straight from the observed
rfft half-plane, bin for bin), with each image's CTF sign flipped out ("phase
flipping") and normalised. Training images come from the fused simulator
itself: random uniform poses, offsets and CTFs; cryomi_simulate_band forms
only the low-frequency band of each clean Fourier image that the crop reads
(a (64 / box)^2 fraction of the slice), which is normalised to unit variance
(as the band-limited box image), scaled by sqrt(SNR) and cropped, plus white
noise whose variance matches the cropped noise of an observed image (a
factor-f crop of unit white noise has variance f^2 under numpy's irfft
normalisation: the box^2-point forward and 64^2-point inverse leave a factor
box^2 / 64^2). Every step runs on the GPU through libcryomi.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _lib, cnn, ops
from .rotations import SO3
from .simulator import wavelength_from_kilovolts
from .synthetic import uniform_quaternions

INPUT = cnn.IMAGE


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def network_input(spectra, params_dev, box, *, complex128=True, snr=None, noise_std=0.0, seed=0, out=None,
                  work=None, band=False):
    """bf16 (n, 64, 64) network inputs from rfft half-planes (n, box, box/2+1).

    Observed images: complex128 spectra rfftn(ifftshift(image)), snr=None,
    noise_std=0. Training images: complex64 clean spectra from the simulator,
    snr (n,) float32 device tensor, noise_std = box / 64; with band=True the
    spectra are cryomi_simulate_band's (n, 65, 33) low-frequency bands.
    """
    lib = _lib.load()
    n = int(spectra.shape[0])
    dev = spectra.device
    st = _stream(dev)
    crop = work if work is not None else torch.empty((n, INPUT, INPUT // 2 + 1), dtype=torch.complex64, device=dev)
    flags = (_lib.CROP_PHASE_FLIP | _lib.CROP_HARTLEY | (_lib.CROP_NORMALIZE if snr is not None else 0)
             | (_lib.CROP_BAND_INPUT if band else 0))
    _lib.check(lib.cryomi_crop_spectrum(spectra.data_ptr(), 1 if complex128 else 0, n, box, INPUT, params_dev.data_ptr(),
                                        None if snr is None else snr.data_ptr(), flags, crop.data_ptr(), st),
               "cryomi_crop_spectrum")
    img = ops.irfft2(crop, INPUT)
    if out is None:
        out = torch.empty((n, INPUT, INPUT), dtype=torch.int16, device=dev)
    _lib.check(lib.cryomi_noise_normalize(img.data_ptr(), n, INPUT, float(noise_std), int(seed), out.data_ptr(), st),
               "cryomi_noise_normalize")
    return out


def draw_training_batch(rng, n, pixel_size, voltage_kv=300.0, ac=0.1, cs=2.7, offset_max=2.0,
                        defocus=(8000.0, 20000.0), astigmatism=(0.0, 800.0), snr=(0.05, 0.5)):
    """Host draws for n training images: packed params, rotation targets, offset targets, SNRs."""
    wxyz = uniform_quaternions(n, rng)
    mats = SO3(wxyz).normalize().sandwich_matrix().reshape(n, 9)
    off = rng.uniform(-offset_max, offset_max, (n, 2))
    df = rng.uniform(*defocus, n)
    ast = rng.uniform(*astigmatism, n)
    ang = rng.uniform(0.0, 180.0, n)
    prm = _lib.pack_params(mats, off, df, ast, ang, cs, wavelength_from_kilovolts(voltage_kv), pixel_size,
                           math.atan(ac / math.sqrt(1 - ac * ac)), 1.0, n)
    s = rng.uniform(*snr, n)
    return prm, mats.astype(np.float32), (off / max(offset_max, 1e-9)).astype(np.float32), s.astype(np.float32)


class ProposalTrainer:
    """Trains a PoseNet on images the fused simulator forms on the fly.

    grid: complex64 (box, box, box) device tensor (the volume to search);
    batch: images per Adam step (a multiple of 8).
    """

    def __init__(self, grid, pixel_size, *, net=None, batch=512, offset_max=2.0, snr=(0.05, 0.5), lr=3e-4,
                 seed=0, off_weight=0.05, eps=0.05, wrap_negative_indices=True, **ctf):
        self.grid = grid
        self.gather_flags = 0 if wrap_negative_indices else _lib.GATHER_FILL
        self.box = int(grid.shape[0])
        self.dev = grid.device
        self.pixel_size = float(pixel_size)
        self.net = net if net is not None else cnn.PoseNet(self.dev, seed=seed)
        if batch % 8:
            raise ValueError("batch must be a multiple of 8")
        self.batch = batch
        self.offset_max = offset_max
        self.snr = snr
        self.lr = lr
        self.off_weight = off_weight
        self.eps = eps
        self.ctf = ctf
        self.rng = np.random.default_rng(seed)
        self.step = 0
        b, m = batch, self.box // 2 + 1
        self.band = self.box > INPUT  # simulate only the band the crop reads
        if self.band:
            self._spec = torch.empty((b, INPUT + 1, INPUT // 2 + 1), dtype=torch.complex64, device=self.dev)
        else:
            self._spec = torch.empty((b, self.box, m), dtype=torch.complex64, device=self.dev)
        self._crop = torch.empty((b, INPUT, INPUT // 2 + 1), dtype=torch.complex64, device=self.dev)
        self._x = torch.empty((b, INPUT, INPUT), dtype=torch.int16, device=self.dev)
        self._ws = ops.Workspace()

    def _chunk(self, steps):
        """Device copies of the draws for `steps` batches (one host->device transfer per chunk)."""
        n = steps * self.batch
        prm, rot, off, s = draw_training_batch(self.rng, n, self.pixel_size, offset_max=self.offset_max,
                                               snr=self.snr, **self.ctf)
        to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)  # noqa: E731
        return to(prm), to(rot), to(off), to(s)

    def train(self, n_images, log_every=0, chunk_steps=64, lr_schedule=None):
        """Run ceil(n_images / batch) Adam steps; returns a list of (step, mean loss, mean min-distance)."""
        steps = -(-int(n_images) // self.batch)
        hist = []
        done = 0
        t0 = time.perf_counter()
        flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.OUTPUT_FOURIER | self.gather_flags
        while done < steps:
            cs = min(chunk_steps, steps - done)
            prm, rot, off, snr = self._chunk(cs)
            for j in range(cs):
                sl = slice(j * self.batch, (j + 1) * self.batch)
                if self.band:
                    ops.simulate_band(self.grid, prm[sl], flags, INPUT, out=self._spec)
                else:
                    ops.simulate_images(self.grid, prm[sl], flags, out=self._spec)
                network_input(self._spec, prm[sl], self.box, complex128=False, snr=snr[sl],
                              noise_std=self.box / INPUT, seed=(self.step << 20) ^ 0x9E37, out=self._x,
                              work=self._crop, band=self.band)
                out = self.net.forward(self._x, self.batch)
                dout, terms = self.net.loss(out, rot[sl], off[sl], eps=self.eps, off_weight=self.off_weight)
                self.net.backward(dout)
                lr = self.lr if lr_schedule is None else lr_schedule(self.step, steps)
                self.net.adam_step(lr=lr)
                self.step += 1
                if log_every and (self.step % log_every == 0 or done + j + 1 == steps):
                    t = terms.mean(dim=0).cpu().numpy()
                    hist.append((self.step, float(t[0]), float(t[1]), time.perf_counter() - t0))
            done += cs
        torch.cuda.synchronize()
        return hist


def cosine_schedule(base_lr, warmup=200, floor=0.05):
    """Linear warm-up, then cosine decay to floor * base_lr."""

    def lr(step, total):
        if step < warmup:
            return base_lr * (step + 1) / warmup
        t = (step - warmup) / max(total - warmup, 1)
        return base_lr * (floor + (1 - floor) * 0.5 * (1 + math.cos(math.pi * min(t, 1.0))))

    return lr


class LearnedPoseSearch:
    """Exact NCC scoring around learned proposals, with exhaustive fallback.

    Per batch of observed images (rfft half-planes, complex128, as for
    cryomi_ncc_search): the network proposes eight orientations per image;
    every grid rotation within `first_radius` degrees of one of them is scored
    exactly (cryomi_ncc_search_candidates: the exhaustive search's FP32
    contraction, tie window and FP64 re-scoring, on the same grid and raveled
    index); then up to `refine_steps` rounds score the grid ball of
    `refine_radius` around the current best (a hill climb on the grid, merged
    with the same tie rule; after the first round only around bests that
    moved), stopping once no image's best moves. Images whose candidate list overflowed, or whose best -NCC is
    above `fallback_score` (poor agreement: the proposals likely missed), are
    re-searched exhaustively with cryomi_ncc_search, whose answer stays the reference.
    """

    def __init__(self, net, grid_c64, grid_c128, rotations, offsets, *, first_radius=15.0, refine_radius=12.0,
                 refine_steps=6, cap=1024, refine_cap=256, fallback_score=None, n_hyp=cnn.N_HYP,
                 wrap_negative_indices=True, bank=None):
        from . import search

        self.wrap_negative_indices = bool(wrap_negative_indices)
        self.net = net
        self.grid_c64 = grid_c64
        self.grid_c128 = grid_c128
        self.box = int(grid_c64.shape[0])
        self.rotations = rotations
        self.offsets = offsets
        self.first_radius = float(first_radius)
        self.refine_radius = float(refine_radius)
        self.refine_steps = int(refine_steps)
        self.cap = int(cap)
        self.refine_cap = int(refine_cap)
        self.fallback_score = fallback_score
        self.n_hyp = int(n_hyp)
        # the FP32 template bank of the grid (search.template_bank), built here unless given
        self.bank = bank if bank is not None else search.template_bank(grid_c64, rotations, self.wrap_negative_indices)

    def rebuild_bank(self):
        from . import search

        self.bank = search.template_bank(self.grid_c64, self.rotations, self.wrap_negative_indices)

    def propose(self, observed_rfft, image_params):
        """Network output (n, 96) for the observed images."""
        x = network_input(observed_rfft, image_params, self.box)
        return self.net.forward(x, int(observed_rfft.shape[0]))

    def calibrate(self, observed_rfft, image_params, exhaustive_index, catch=0.7, max_fallback=0.05,
                  min_fallback=0.005):
        """Set `fallback_score` from a calibration set with known exhaustive results.

        The miss signal is the learned best score itself: an image whose proposals missed the
        exhaustive argmax ends on a worse (higher -NCC) local optimum. The threshold is the
        score above which a fraction `catch` of the calibration misses lie, kept between the
        `max_fallback` and `min_fallback` quantiles of all calibration scores (so at most /
        at least that fraction of images goes to the exhaustive search). Returns the
        calibration statistics."""
        bi, bs, _ = self.search(observed_rfft, image_params, fallback=False)
        bi = bi.cpu().numpy()
        bs = bs.cpu().numpy()
        ex = np.asarray(exhaustive_index.cpu().numpy() if torch.is_tensor(exhaustive_index) else exhaustive_index)
        miss = bi != ex
        hi = float(np.quantile(bs, 1.0 - min_fallback))  # cut at most here: >= min_fallback fall back
        lo = float(np.quantile(bs, 1.0 - max_fallback))  # cut at least here: <= max_fallback fall back
        cut = float(np.quantile(bs[miss], 1.0 - catch)) - 1e-12 if miss.any() else hi
        cut = min(max(cut, lo), hi)
        self.fallback_score = cut
        caught = miss & (bs > cut)
        return {"threshold": cut, "miss_rate": float(miss.mean()), "fallback_rate": float(np.mean(bs > cut)),
                "misses_caught": int(caught.sum()), "misses": int(miss.sum()), "signal": "learned best -NCC"}

    def search(self, observed_rfft, image_params, *, fallback=True):
        """Returns (best_index int64, best_score float64, info dict)."""
        from . import search

        n = int(observed_rfft.shape[0])
        n_off = int(self.offsets.shape[0])
        out = self.propose(observed_rfft, image_params)
        search.reserve_candidates_workspace(self.grid_c128, n, n_off, max(self.cap, self.refine_cap))
        cand, count, cover = search.pose_candidates(self.rotations, self.first_radius, self.cap, hypotheses=out,
                                                    hyp_stride=cnn.N_OUT, n_hyp=self.n_hyp, n_images=n)
        wrap = self.wrap_negative_indices
        best = search.ncc_search_candidates(self.grid_c128, observed_rfft, image_params, self.rotations, self.offsets,
                                            self.bank, cand, count, wrap_negative_indices=wrap)
        rounds = 0
        moved = None  # images whose best moved in the last round (all before the first round)
        for _ in range(self.refine_steps):  # hill climb on the grid until no image's best moves
            before = best[0].clone()
            # a ball is scored only around a best that moved: the others' balls were scored already
            centre = best[0] if moved is None else torch.where(moved, best[0], torch.full_like(best[0], -1))
            c2, n2, o2 = search.pose_candidates(self.rotations, self.refine_radius, self.refine_cap,
                                                center_index=centre, n_offsets=n_off)
            search.ncc_search_candidates(self.grid_c128, observed_rfft, image_params, self.rotations, self.offsets,
                                         self.bank, c2, n2, best=best,
                                         flags=search.REUSE_PREPARED | search.MERGE, wrap_negative_indices=wrap)
            cover = cover | o2
            rounds += 1
            moved = best[0] != before
            if not bool(moved.any()):
                break
        bi, bs, _ = best
        info = {"candidates_first": count, "fallback_images": 0, "refine_rounds": rounds}
        if fallback:
            need = (cover != 0) | (bi < 0) | (bi >= int(self.rotations.shape[0]) * n_off) | ~torch.isfinite(bs)
            if self.fallback_score is not None:
                need = need | (bs > self.fallback_score)
            idx = torch.nonzero(need).flatten()
            info["fallback_images"] = int(idx.numel())
            if idx.numel():
                fi, fs, _ = search.ncc_search(self.grid_c64, self.grid_c128, observed_rfft[idx].contiguous(),
                                              image_params[idx].contiguous(), self.rotations, self.offsets,
                                              bank=self.bank, wrap_negative_indices=wrap)
                bi[idx] = fi
                bs[idx] = fs
            info["fallback_mask"] = need
        return bi, bs, info
