"""The trained pose-proposal network against the same pipeline without
training: a short training run on simulator output must turn proposals into
an accurate search (the exhaustive argmax found for nearly every image), while
the untrained network's proposals find few of them."""

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, learned, ops, search
from cryojax_amd.rotations import SO3
from cryojax_amd.synthetic import draw_image_parameters, gaussian_mixture_volume, pack_drawn_parameters

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_training_gains_over_untrained_network():
    n, n_img, snr = 64, 200, 0.5
    vol = gaussian_mixture_volume(n, 300, 20, 5)
    g = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(vol.astype(np.float64))))
    g64 = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    g128 = torch.from_numpy(g).to(DEV)
    d = draw_image_parameters(n_img, 6, offset_max=2.0)
    p = ops.params_to_device(pack_drawn_parameters(d, 2.0), DEV)
    clean = ops.simulate_images(g64, p, _lib.APPLY_CTF | _lib.APPLY_TRANSLATION)
    clean = (clean - clean.mean(dim=(1, 2), keepdim=True)) / clean.std(dim=(1, 2), unbiased=False, keepdim=True)
    gen = torch.Generator(device=DEV).manual_seed(7)
    y = search.observed_spectrum(np.sqrt(snr) * clean + torch.randn(clean.shape, generator=gen, device=DEV))
    rmats = SO3(search.super_fibonacci_quaternions(5000)).sandwich_matrix().reshape(-1, 9)
    rots = torch.from_numpy(rmats).to(DEV)
    offs = torch.from_numpy(search.offset_grid(5, 1.0)).to(DEV)
    ei, es, _ = search.ncc_search(g64, g128, y, p, rots, offs)
    ei = ei.cpu().numpy()

    def accuracy(net):
        ls = learned.LearnedPoseSearch(net, g64, g128, rots, offs, first_radius=20.0, refine_radius=15.0,
                                       refine_steps=2, cap=512, refine_cap=128)
        li, ls_, info = ls.search(y, p, fallback=False)
        return float(np.mean(li.cpu().numpy() == ei))

    trainer = learned.ProposalTrainer(g64, 2.0, batch=256, lr=5e-4, seed=3, snr=(0.3, 1.0))
    before = accuracy(trainer.net)
    hist = trainer.train(300_000, log_every=200, lr_schedule=learned.cosine_schedule(5e-4, warmup=100))
    after = accuracy(trainer.net)
    print(f"exhaustive argmax found: untrained {before:.3f}, trained {after:.3f}; loss {hist[0][1]:.3f} -> "
          f"{hist[-1][1]:.3f}")
    assert hist[-1][1] < 0.7 * hist[0][1]
    assert after >= 0.9
    assert after >= before + 0.5
