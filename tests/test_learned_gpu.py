"""Learned-proposal search: candidate generation against brute force, the
candidate-restricted NCC search against the exhaustive search on the same
candidates (index for index), merging, and a trained network's gain."""

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, learned, ops, search
from cryojax_amd.rotations import SO3
from oracle import cnn as C
from tests.helpers import draw, fourier_grid, make_volume, packed_params

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rotations(n):
    return SO3(search.super_fibonacci_quaternions(n)).sandwich_matrix().reshape(-1, 9)


def brute_ball(hyp9, rmats, angle):
    """Indices of grid rotations within `angle` degrees of any hypothesis, and those within 1e-4 of the edge."""
    R = C.orthonormalize(hyp9.reshape(-1, 9)).reshape(-1, 9)
    tr = rmats @ R.T  # (n_rot, n_hyp)
    thr = 1 + 2 * np.cos(np.radians(angle))
    inside = np.nonzero((tr >= thr).any(axis=1))[0]
    edge = np.nonzero((np.abs(tr - thr) < 1e-4).any(axis=1))[0]
    return inside, set(edge.tolist())


def test_pose_candidates_match_bruteforce(rng):
    rmats = rotations(3000)
    rd = torch.from_numpy(rmats).to(DEV)
    n, stride = 6, 96
    out = rng.standard_normal((n, stride)).astype(np.float32)
    cand, cnt, over = search.pose_candidates(rd, 20.0, 512, hypotheses=torch.from_numpy(out).to(DEV), hyp_stride=stride,
                                             n_hyp=8, n_images=n)
    cand, cnt, over = cand.cpu().numpy(), cnt.cpu().numpy(), over.cpu().numpy()
    for i in range(n):
        want, edge = brute_ball(out[i, :72].astype(np.float64), rmats, 20.0)
        got = cand[i, :cnt[i]]
        assert np.all(np.diff(got) > 0), "ascending order"
        assert set(got.tolist()) ^ set(want.tolist()) <= edge
        assert over[i] == 0
    # a cap below the count flags overflow and keeps the lowest indices
    c2, n2, o2 = search.pose_candidates(rd, 20.0, 8, hypotheses=torch.from_numpy(out).to(DEV), hyp_stride=stride,
                                        n_hyp=8, n_images=n)
    assert np.all(n2.cpu().numpy() == 8) and np.all(o2.cpu().numpy() == 1)
    np.testing.assert_array_equal(c2.cpu().numpy(), cand[:, :8])


def test_center_candidates(rng):
    rmats = rotations(2000)
    rd = torch.from_numpy(rmats).to(DEV)
    n_off = 25
    centers = np.array([0, 37 * n_off + 3, 1999 * n_off + 24, -1, 2000 * n_off], dtype=np.int64)
    cand, cnt, over = search.pose_candidates(rd, 15.0, 256, center_index=torch.from_numpy(centers).to(DEV),
                                             n_offsets=n_off)
    cand, cnt = cand.cpu().numpy(), cnt.cpu().numpy()
    for i, ci in enumerate(centers[:3]):
        want, edge = brute_ball(rmats[ci // n_off], rmats, 15.0)
        got = cand[i, :cnt[i]]
        assert ci // n_off in got
        assert set(got.tolist()) ^ set(want.tolist()) <= edge
    assert cnt[3] == 0 and cnt[4] == 0  # missing results match nothing


@pytest.fixture(scope="module")
def problem():
    n, n_img, n_rot = 64, 10, 1500
    vol = make_volume(n, 60, 16, 3)
    g = fourier_grid(vol)
    g64 = torch.from_numpy(g.astype(np.complex64)).to(DEV)
    g128 = torch.from_numpy(g).to(DEV)
    d = draw(n_img, 5, offset_max=1.5)
    p = ops.params_to_device(packed_params(d, 2.0), DEV)
    clean = ops.simulate_images(g64, p, _lib.APPLY_CTF | _lib.APPLY_TRANSLATION)
    clean = (clean - clean.mean(dim=(1, 2), keepdim=True)) / clean.std(dim=(1, 2), unbiased=False, keepdim=True)
    gen = torch.Generator(device=DEV).manual_seed(6)
    noisy = np.sqrt(0.3) * clean + torch.randn(clean.shape, generator=gen, device=DEV)
    y = search.observed_spectrum(noisy)
    rmats = rotations(n_rot)
    rd = torch.from_numpy(rmats).to(DEV)
    offs = torch.from_numpy(search.offset_grid(3, 1.0)).to(DEV)
    bank = search.template_bank(g64, rd)
    return dict(g64=g64, g128=g128, p=p, y=y, rmats=rmats, rd=rd, offs=offs, bank=bank, n_img=n_img, n_rot=n_rot)


def cand_tensors(lists, cap):
    n = len(lists)
    cand = np.full((n, cap), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    for i, l in enumerate(lists):
        cand[i, :len(l)] = l
        cnt[i] = len(l)
    return torch.from_numpy(cand).to(DEV), torch.from_numpy(cnt).to(DEV)


def test_all_candidates_equal_exhaustive(problem):
    P = problem
    ei, es, _ = search.ncc_search(P["g64"], P["g128"], P["y"], P["p"], P["rd"], P["offs"])
    cand, cnt = cand_tensors([np.arange(P["n_rot"])] * P["n_img"], 2048)
    resc = torch.zeros(P["n_img"], dtype=torch.int32, device=DEV)
    bi, bs, _ = search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], P["bank"], cand, cnt,
                                             rescored=resc)
    print("rescored per image", resc.cpu().numpy())
    assert torch.equal(bi, ei)
    assert torch.equal(bs, es)
    assert int(resc.min()) >= 1


def test_random_subsets_equal_restricted_exhaustive(problem, rng):
    P = problem
    n_off = int(P["offs"].shape[0])
    lists = [np.sort(rng.choice(P["n_rot"], int(rng.integers(1, 400)), replace=False)) for _ in range(P["n_img"])]
    cand, cnt = cand_tensors(lists, 512)
    bi, bs, _ = search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], P["bank"], cand, cnt)
    bi, bs = bi.cpu().numpy(), bs.cpu().numpy()
    for i, l in enumerate(lists):
        sub = torch.from_numpy(P["rmats"][l]).to(DEV)
        li, ls, _ = search.ncc_search(P["g64"], P["g128"], P["y"][i:i + 1].contiguous(), P["p"][i:i + 1].contiguous(),
                                      sub, P["offs"])
        li, ls = int(li[0]), float(ls[0])
        assert bi[i] == l[li // n_off] * n_off + li % n_off, i
        assert bs[i] == ls


def test_merge_of_disjoint_subsets_equals_union(problem, rng):
    P = problem
    perm = rng.permutation(P["n_rot"])
    a = [np.sort(perm[:600])] * P["n_img"]
    b = [np.sort(perm[600:1200])] * P["n_img"]
    u = [np.sort(perm[:1200])] * P["n_img"]
    ca, na = cand_tensors(a, 1024)
    cb, nb = cand_tensors(b, 1024)
    cu, nu = cand_tensors(u, 2048)
    best = search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], P["bank"], ca, na)
    search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], P["bank"], cb, nb, best=best,
                                 flags=search.REUSE_PREPARED | search.MERGE)
    ui, us, _ = search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], P["bank"], cu, nu)
    assert torch.equal(best[0], ui) and torch.equal(best[1], us)


def test_template_bank_rows_are_slices(problem):
    """Bank row r = the bare FP32 slice at rotation r (the forward's Fourier output without CTF / shift)."""
    P = problem
    n = int(P["g64"].shape[0])
    r = [0, 7, 1499]
    prm = _lib.pack_params(P["rmats"][r], np.zeros((3, 2)), 0.0, 0.0, 0.0, 0.0, 0.02, 2.0, 0.0, 1.0, 3)
    f = ops.simulate_images(P["g64"], ops.params_to_device(prm, DEV), _lib.OUTPUT_FOURIER)
    nh = n * (n // 2 + 1)
    got = P["bank"][r, :2 * nh].cpu().numpy().view(np.complex64).reshape(3, n, n // 2 + 1)
    want = f.cpu().numpy().copy()
    m = n // 2
    for b_ in (want, got):  # the self-conjugate bins differ only by the forward's fix-ups
        b_[:, m, 0] = b_[:, 0, m] = b_[:, m, m] = 0
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())


def test_candidates_equal_exhaustive_under_fill_rule(problem):
    """With every corner outside [0, N) filled (CRYOMI_GATHER_FILL), the bank, the candidate
    scoring and its FP64 re-scoring follow the same rule as the exhaustive search."""
    P = problem
    bank = search.template_bank(P["g64"], P["rd"], wrap_negative_indices=False)
    ei, es, _ = search.ncc_search(P["g64"], P["g128"], P["y"], P["p"], P["rd"], P["offs"], wrap_negative_indices=False)
    wi, ws, _ = search.ncc_search(P["g64"], P["g128"], P["y"], P["p"], P["rd"], P["offs"])
    cand, cnt = cand_tensors([np.arange(P["n_rot"])] * P["n_img"], 2048)
    bi, bs, _ = search.ncc_search_candidates(P["g128"], P["y"], P["p"], P["rd"], P["offs"], bank, cand, cnt,
                                             wrap_negative_indices=False)
    print("fill-rule scores differ from the wrap rule's by up to", float((es - ws).abs().max()))
    assert torch.equal(bi, ei)
    assert torch.equal(bs, es)
    assert not torch.equal(es, ws)  # the rules really differ on this grid
